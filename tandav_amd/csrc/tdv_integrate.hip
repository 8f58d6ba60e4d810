/*
 * The KDK integrator on the device, per particle on the integer timeline.
 *
 * Integer time (intime_t = uint32_t, globals.h:33; 32 time bins): the run
 * from Time.Begin to Time.End is 2^31 ticks of Time.Step_Min
 * (Setup_Time_Integration, timestep.c:130-167), in ln(a) on comoving runs.
 * A tick count maps to integration time as Time.Begin + it Step_Min, or
 * Time.Begin exp(it Step_Min) on comoving runs (timestep.c:340-372).
 *
 * Kick_First/Second_Halfstep (kick.c:9-70): each active particle kicks by
 * half of its own bin, from its own It_Kick_Pos:
 *   it_next = It_Kick_Pos + 2^Time_Bin / 2,  v += K(It_Kick_Pos, it_next) a.
 * Drift_To_Sync_Point (drift.c:12-64): each active particle drifts from its
 * own It_Drift_Pos to the common next sync time Int_Time.Next,
 *   x += D(It_Drift_Pos, it_next) v,
 * then positions are wrapped into the box (periodic.c:30-53).
 * K and D (Particle_Kick_Step / Particle_Drift_Step) are the difference of
 * the integration time at the two ends, or on comoving runs the difference
 * of natural cubic splines through tabulated integrals of 1/(H a^2) and
 * 1/(H a^3) (comoving.c:30-153): 512 knots a_i = 0.95 a_begin exp(i d),
 * d = ln(1.05 a_end / (0.95 a_begin)) / 511, i = 1..511, plus the knot
 * (0, 0) that the reference leaves unset; each integral from 0.95 a_begin.
 * The tables are formed on the host here by 10-point Gauss-Legendre
 * quadrature on every knot interval (the reference uses an adaptive
 * 41-point Gauss-Kronrod rule to 1e-8; both are converged far below that).
 *
 * set_new_particle_timebins (timestep.c:171-201): every particle's step
 * from its gravitational acceleration (cosmological_timestep, Dehnen &
 * Read 2011 eq. 21, timestep.c:513-526) in the reference's float / double
 * types, capped by Dt_Max_Global, to a power-of-two bin (timestep2timebin,
 * timestep.c:380-388), no longer than the largest active bin allows.
 */
#include "tdv_internal.h"

#include <cmath>
#include <cfloat>
#include <vector>

#define TL_KNOTS 512                 /* TABLESIZE, comoving.c:5 */
#define TL_NBITS 32                  /* N_INT_BINS for a 32-bit intime_t */

/* ---- host: comoving tables and the natural cubic spline ----------------- */

/* E(a) = H(a) / H0, cosmology.c:70-74 */
static double e_hubble(const tdv_timeline &t, double a)
{
	return sqrt(t.omega_lambda + (1.0 - t.omega_0) / (a * a)
			+ t.omega_matter / (a * a * a) + t.omega_rad / (a * a * a * a));
}

/* nodes and weights of the n-point Gauss-Legendre rule on [-1, 1] (Newton
 * iteration on P_n from the Chebyshev guesses) */
static void gauss_legendre(int n, double *x, double *w)
{
	for (int i = 0; i < n; i++) {
		double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1;
		for (int it = 0; it < 100; it++) {
			double p0 = 1, p1 = z;
			for (int k = 2; k <= n; k++) {
				const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
				p0 = p1;
				p1 = p2;
			}
			dp = n * (z * p1 - p0) / (z * z - 1);
			const double dz = p1 / dp;
			z -= dz;
			if (fabs(dz) < 1e-16)
				break;
		}
		x[i] = z;
		w[i] = 2.0 / ((1 - z * z) * dp * dp);
	}
}

/* setup_kick_drift_factors (comoving.c:109-153): knots and the integrals of
 * 1/(H a^3) (drift) and 1/(H a^2) (kick) from 0.95 a_begin */
static void comoving_tables(const tdv_timeline &t, double *a, double *kick,
		double *drift)
{
	const double beg = 0.95 * t.time_begin, end = 1.05 * t.time_end;
	const double d = log(end / beg) / (TL_KNOTS - 1);
	double gx[10], gw[10];
	gauss_legendre(10, gx, gw);
	a[0] = kick[0] = drift[0] = 0;   /* left unset by the reference */
	double lo = beg, sk = 0, sd = 0;
	for (int i = 1; i < TL_KNOTS; i++) {
		const double hi = beg * exp(d * i);
		const double mid = 0.5 * (lo + hi), half = 0.5 * (hi - lo);
		double ik = 0, id = 0;
		for (int q = 0; q < 10; q++) {
			const double x = mid + half * gx[q];
			const double h = t.hubble * e_hubble(t, x);
			ik += gw[q] / (h * x * x);
			id += gw[q] / (h * x * x * x);
		}
		sk += half * ik;
		sd += half * id;
		a[i] = hi;
		kick[i] = sk;
		drift[i] = sd;
		lo = hi;
	}
}

/* natural cubic spline through (x_i, y_i): c_i = y''(x_i) / 2, c_0 =
 * c_{n-1} = 0, from the symmetric tridiagonal system of the interior knots
 * (forward elimination, back substitution) */
static void natural_spline(const double *x, const double *y, double *c,
		int n)
{
	const int m = n - 2;
	std::vector<double> diag(m), off(m), rhs(m), z(m);
	for (int i = 0; i < m; i++) {
		const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
		diag[i] = 2.0 * (h0 + h1);
		off[i] = h1;
		rhs[i] = 3.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
	}
	/* A = L D L^T, L unit lower bidiagonal with l_i = off_i / d_i */
	std::vector<double> dd(m), l(m);
	dd[0] = diag[0];
	for (int i = 1; i < m; i++) {
		l[i - 1] = off[i - 1] / dd[i - 1];
		dd[i] = diag[i] - off[i - 1] * l[i - 1];
	}
	z[0] = rhs[0];
	for (int i = 1; i < m; i++)
		z[i] = rhs[i] - l[i - 1] * z[i - 1];
	c[0] = c[n - 1] = 0;
	c[m] = z[m - 1] / dd[m - 1];
	for (int i = m - 2; i >= 0; i--)
		c[i + 1] = z[i] / dd[i] - l[i] * c[i + 2];
}

/* ---- device: integration time, factors, kernels ------------------------- */

struct TimelineDev {
	int comoving;
	double begin, step_min;
	const double *x, *yk, *ck, *yd, *cd;  /* knots, kick y/c, drift y/c */
};

/* Integer_Time2Integration_Time (timestep.c:340-372) */
__device__ __forceinline__ double it2time(const TimelineDev &t, uint32_t it)
{
	return t.comoving ? t.begin * exp(it * t.step_min)
			: t.begin + it * t.step_min;
}

/* the spline at a: the interval [x_i, x_{i+1}] with x_i <= a < x_{i+1} by
 * bisection over the knots, then y_i + u (b + u (c_i + u e)) */
__device__ double spline_at(const double *x, const double *y,
		const double *c, double a)
{
	int lo = 0, hi = TL_KNOTS - 1;
	while (hi > lo + 1) {
		const int mid = (lo + hi) >> 1;
		if (x[mid] > a)
			hi = mid;
		else
			lo = mid;
	}
	const double h = x[lo + 1] - x[lo], u = a - x[lo];
	const double b = (y[lo + 1] - y[lo]) / h - h * (c[lo + 1] + 2.0 * c[lo])
			/ 3.0;
	const double e = (c[lo + 1] - c[lo]) / (3.0 * h);
	return y[lo] + u * (b + u * (c[lo] + u * e));
}

/* Particle_Kick_Step (kick.c:81-87, comoving.c:30-39) */
__device__ __forceinline__ double kick_step(const TimelineDev &t,
		uint32_t i0, uint32_t i1)
{
	const double a0 = it2time(t, i0), a1 = it2time(t, i1);
	if (!t.comoving)
		return a1 - a0;
	return spline_at(t.x, t.yk, t.ck, a1) - spline_at(t.x, t.yk, t.ck, a0);
}

/* Particle_Drift_Step (drift.c:115-121, comoving.c:41-50) */
__device__ __forceinline__ double drift_step(const TimelineDev &t,
		uint32_t i0, uint32_t i1)
{
	const double a0 = it2time(t, i0), a1 = it2time(t, i1);
	if (!t.comoving)
		return a1 - a0;
	return spline_at(t.x, t.yd, t.cd, a1) - spline_at(t.x, t.yd, t.cd, a0);
}

__global__ void k_kick_bins(int nact, const int *__restrict__ active,
		const int *__restrict__ bin, uint32_t *__restrict__ it_kick,
		const double *__restrict__ ax, const double *__restrict__ ay,
		const double *__restrict__ az, double *__restrict__ vx,
		double *__restrict__ vy, double *__restrict__ vz,
		double *__restrict__ kick_dt, TimelineDev t)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nact)
		return;
	const int i = active[g];
	const uint32_t step = ((uint32_t)1 << bin[i]) >> 1;
	const uint32_t i0 = it_kick[i], i1 = i0 + step;
	const double dt = kick_step(t, i0, i1);
	vx[i] = vx[i] + dt * ax[i];
	vy[i] = vy[i] + dt * ay[i];
	vz[i] = vz[i] + dt * az[i];
	it_kick[i] = i1;
	kick_dt[i] = dt;
}

/* lower bound of v in the sorted active list */
__device__ __forceinline__ int active_lower(const int *__restrict__ a, int n,
		int v)
{
	int lo = 0, hi = n;
	while (lo < hi) {
		const int mid = (lo + hi) >> 1;
		if (a[mid] < v)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

/* Gravity_Tree_Update_Kicks (tree_update.c:12-57) for every active
 * particle: each tree node on the particle's path below its subtree root
 * (the nodes whose particle range holds it) takes Dp += m dt a / M, marked
 * UPDATED; the node's particles in particle order, so the sums are the
 * reference's one-thread sums. */
__global__ void k_tree_kicks(size_t nn, int4 *__restrict__ nc,
		const double4 *__restrict__ na, double4 *__restrict__ dpn,
		int nact, const int *__restrict__ active,
		const double *__restrict__ m, const double *__restrict__ dtk,
		const double *__restrict__ ax, const double *__restrict__ ay,
		const double *__restrict__ az)
{
	const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nn)
		return;
	const int4 c = nc[j];
	if (c.w < 0 || ((c.z >> TDV_BIT_TOP) & 1))
		return;                          /* terminator, subtree root */
	const int lo = active_lower(active, nact, c.w);
	const int hi = active_lower(active, nact, c.w + c.y);
	if (lo >= hi)
		return;
	const double M = na[j].w;
	double4 d = dpn[j];
	for (int g = lo; g < hi; g++) {
		const int i = active[g];
		const double mdt = m[i] * dtk[i];
		d.x = d.x + mdt * ax[i] / M;
		d.y = d.y + mdt * ay[i] / M;
		d.z = d.z + mdt * az[i] / M;
	}
	dpn[j] = d;
	nc[j].z = c.z | (1 << TDV_BIT_UPDATED);   /* only this node's own */
}

/* the same for the top nodes (tree_update.c:43-55), in float: Dp is a
 * float in D[].TNode, and the node is marked kicked */
__global__ void k_top_kicks(const TopNodes T, int nact,
		const int *__restrict__ active, const double *__restrict__ m,
		const double *__restrict__ dtk, const double *__restrict__ ax,
		const double *__restrict__ ay, const double *__restrict__ az)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= *T.ntop)
		return;
	const int lo = active_lower(active, nact, T.first[t]);
	const int hi = active_lower(active, nact, T.first[t] + T.npart[t]);
	if (lo >= hi)
		return;
	const double M = (double)T.mass[t];
	float dx = T.dp[3 * t], dy = T.dp[3 * t + 1], dz = T.dp[3 * t + 2];
	for (int g = lo; g < hi; g++) {
		const int i = active[g];
		const double mdt = m[i] * dtk[i];
		dx = (float)((double)dx + mdt * ax[i] / M);
		dy = (float)((double)dy + mdt * ay[i] / M);
		dz = (float)((double)dz + mdt * az[i] / M);
	}
	T.dp[3 * t] = dx;
	T.dp[3 * t + 1] = dy;
	T.dp[3 * t + 2] = dz;
	T.kicked[t] = 1;
}

/* Gravity_Tree_Update_Drift (tree_update.c:64-104): the updated nodes move
 * by dt Dp and their Dp and flag are cleared; the kicked top nodes the
 * same, in float */
__global__ void k_tree_drift(size_t nn, int4 *__restrict__ nc,
		double4 *__restrict__ na, double4 *__restrict__ dpn, double dt)
{
	const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nn)
		return;
	const int4 c = nc[j];
	if (!((c.z >> TDV_BIT_UPDATED) & 1))
		return;
	const double4 d = dpn[j];
	double4 a = na[j];
	a.x = a.x + dt * d.x;
	a.y = a.y + dt * d.y;
	a.z = a.z + dt * d.z;
	na[j] = a;
	dpn[j] = make_double4(0, 0, 0, 0);
	nc[j].z = c.z & ~(1 << TDV_BIT_UPDATED);
}

__global__ void k_top_drift(const TopNodes T, double dt)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= *T.ntop || !T.kicked[t])
		return;
	for (int d = 0; d < 3; d++) {
		T.com[3 * t + d] = (float)((double)T.com[3 * t + d]
				+ dt * (double)T.dp[3 * t + d]);
		T.dp[3 * t + d] = 0.f;
	}
	T.kicked[t] = 0;
}

/* while (x < 0) x += box; while (x >= box) x -= box (periodic.c:33-49),
 * bounded: a position many boxes out is wrapped in one step */
__device__ __forceinline__ double wrap_into_box(double x, double box)
{
	if (!(fabs(x) < 64.0 * box))
		return x - floor(x / box) * box;
	while (x < 0)
		x += box;
	while (x >= box)
		x -= box;
	return x;
}

__global__ void k_drift_sync(int nact, const int *__restrict__ active,
		uint32_t *__restrict__ it_drift, uint32_t it_next,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, const double *__restrict__ vx,
		const double *__restrict__ vy, const double *__restrict__ vz,
		TimelineDev t, int periodic, double box, double4 *__restrict__ pm)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nact)
		return;
	const int i = active[g];
	const double dt = drift_step(t, it_drift[i], it_next);
	double px = x[i] + dt * vx[i], py = y[i] + dt * vy[i],
			pz = z[i] + dt * vz[i];
	if (periodic) {
		px = wrap_into_box(px, box);
		py = wrap_into_box(py, box);
		pz = wrap_into_box(pz, box);
	}
	x[i] = px;
	y[i] = py;
	z[i] = pz;
	it_drift[i] = it_next;
	const double4 o = pm[i];
	pm[i] = make_double4(px, py, pz, o.w);
}

struct TimebinDev {
	int comoving;
	double accuracy, epsilon, dt_max, step_min, step_max;
	double expansion, hubble, accel_factor;
	int max_active_bin;
};

/* timestep2timebin (timestep.c:380-388): N_INT_BINS - 1 - exponent of
 * Step_Max / dt */
__device__ __forceinline__ int dt2bin(double dt, double step_max)
{
	int e;
	(void)frexp(step_max / dt, &e);
	return TL_NBITS - 1 - e;
}

/* set_new_particle_timebins (timestep.c:171-201) with
 * get_physical_timestep and cosmological_timestep (:495-526) and
 * Acceleration_Physical (unit.c:55-71) */
__global__ void k_timebins(size_t n, const double *__restrict__ gx,
		const double *__restrict__ gy, const double *__restrict__ gz,
		int *__restrict__ bin, TimebinDev p, int *__restrict__ minmax,
		unsigned *__restrict__ err)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	int b = TL_NBITS - 1, bmax = 0;
	if (i < n) {
		double acc = sqrt(gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i]);
		if (p.comoving)
			acc *= p.accel_factor;
		const double f = p.comoving
				? p.accuracy * sqrt(2 * p.expansion * p.epsilon / acc)
				: p.accuracy * sqrt(2 * p.epsilon / acc);
		float dts = FLT_MAX;                   /* float, timestep.c:497 */
		dts = (float)fmin((double)dts, (double)(float)f);
		double dt = dts;
		if (p.comoving)
			dt = dt * p.hubble;                /* convert_dt_to_dlna */
		dt = fmin(dt, p.dt_max);
		if (!(dt >= p.step_min))               /* timestep.c:183-187 */
			atomicOr(err, TDV_EBIT_RANGE);
		const int want = dt2bin(dt, p.step_max);
		const int allowed = max(p.max_active_bin, bin[i]);
		b = min(want, allowed);
		bin[i] = b;
		bmax = b;
	}
	/* wave minimum and maximum, one atomic each per wave */
	for (int o = 32; o >= 1; o >>= 1) {
		b = min(b, __shfl_xor(b, o, 64));
		bmax = max(bmax, __shfl_xor(bmax, o, 64));
	}
	if ((threadIdx.x & 63) == 0) {
		atomicMin(&minmax[0], b);
		atomicMax(&minmax[1], bmax);
	}
}

__global__ void k_minmax_init(int *minmax)
{
	minmax[0] = TL_NBITS - 1;
	minmax[1] = 0;
}

/* ---- C ABI --------------------------------------------------------------- */

static int timeline_ok(const tdv_timeline *t)
{
	if (!t || !(t->time_end > t->time_begin))
		return 0;
	if (t->comoving)
		return t->time_begin > 0 && t->hubble > 0 && std::isfinite(t->omega_0)
				&& std::isfinite(t->omega_matter)
				&& std::isfinite(t->omega_lambda)
				&& std::isfinite(t->omega_rad);
	return 1;
}

static TimelineDev timeline_dev(const tdv_context *c)
{
	TimelineDev t;
	t.comoving = c->tl.comoving;
	t.begin = c->tl.time_begin;
	t.step_min = c->tl_step_min;
	t.x = c->tl_spline;
	t.yk = c->tl_spline + TL_KNOTS;
	t.ck = c->tl_spline + 2 * TL_KNOTS;
	t.yd = c->tl_spline + 3 * TL_KNOTS;
	t.cd = c->tl_spline + 4 * TL_KNOTS;
	return t;
}

extern "C" {

int tdv_comoving_tables(const tdv_timeline *tl, double *a, double *kick,
		double *drift)
{
	if (!timeline_ok(tl) || !tl->comoving || !a || !kick || !drift)
		return TDV_ERR_ARG;
	comoving_tables(*tl, a, kick, drift);
	return 0;
}

int tdv_set_timeline(tdv_context *c, const tdv_timeline *tl)
{
	if (!c || !timeline_ok(tl))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_timeline: time_end must "
				"exceed time_begin (> 0 and a finite background when "
				"comoving)");
	c->tl = *tl;
	/* Setup_Time_Integration, timestep.c:141-147 */
	c->tl_step_max = tl->comoving ? log(tl->time_end) - log(tl->time_begin)
			: tl->time_end - tl->time_begin;
	c->tl_step_min = c->tl_step_max / (double)((uint32_t)1 << (TL_NBITS - 1));
	std::vector<double> h(5 * TL_KNOTS, 0.0);
	if (tl->comoving) {
		double *x = h.data(), *yk = x + TL_KNOTS, *ck = x + 2 * TL_KNOTS,
				*yd = x + 3 * TL_KNOTS, *cd = x + 4 * TL_KNOTS;
		comoving_tables(*tl, x, yk, yd);
		natural_spline(x, yk, ck, TL_KNOTS);
		natural_spline(x, yd, cd, TL_KNOTS);
	}
	if (!c->tl_spline)                   /* + 2 ints for the bin min/max */
		TDV_HIP(c, hipMalloc((void **)&c->tl_spline, h.size() * 8 + 8));
	TDV_HIP(c, hipMemcpyAsync(c->tl_spline, h.data(), h.size() * 8,
			hipMemcpyHostToDevice, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	c->have_timeline = 1;
	return 0;
}

int tdv_set_integer_times(tdv_context *c, const uint32_t *it_drift,
		const uint32_t *it_kick)
{
	if (!c || !c->n)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_integer_times: no particles");
	const size_t b = c->n * sizeof(uint32_t);
	if (it_drift)
		TDV_HIP(c, hipMemcpyAsync(c->it_drift, it_drift, b,
				hipMemcpyHostToDevice, c->stream));
	if (it_kick)
		TDV_HIP(c, hipMemcpyAsync(c->it_kick, it_kick, b,
				hipMemcpyHostToDevice, c->stream));
	return tdv_check(c, hipStreamSynchronize(c->stream), "set_integer_times");
}

int tdv_get_integer_times(tdv_context *c, uint32_t *it_drift,
		uint32_t *it_kick)
{
	if (!c)
		return TDV_ERR_ARG;
	const size_t b = c->n * sizeof(uint32_t);
	if (it_drift && b)
		TDV_HIP(c, hipMemcpyAsync(it_drift, c->it_drift, b,
				hipMemcpyDeviceToHost, c->stream));
	if (it_kick && b)
		TDV_HIP(c, hipMemcpyAsync(it_kick, c->it_kick, b,
				hipMemcpyDeviceToHost, c->stream));
	return tdv_check(c, hipStreamSynchronize(c->stream), "get_integer_times");
}

int tdv_kick_halfstep(tdv_context *c, int update_tree)
{
	if (!c || !c->n || !c->have_timeline)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_kick_halfstep: no particles or "
				"no timeline (tdv_set_timeline)");
	if (update_tree && !c->have_tree)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_kick_halfstep: tree kicks "
				"without a tree");
	(void)hipEventRecord(c->ev[EV_KICK0], c->stream);
	if (c->nactive)
		hipLaunchKernelGGL(k_kick_bins, dim3(tdv_blocks((size_t)c->nactive,
				256)), dim3(256), 0, c->stream, c->nactive, c->active,
				c->time_bin, c->it_kick, c->acc[0], c->acc[1], c->acc[2],
				c->vel[0], c->vel[1], c->vel[2], c->kick_dt, timeline_dev(c));
	if (update_tree && c->nactive) {
		if (c->nnodes)
			hipLaunchKernelGGL(k_tree_kicks, dim3(tdv_blocks(c->nnodes, 256)),
					dim3(256), 0, c->stream, c->nnodes, c->node_c, c->node_a,
					c->node_dp, c->nactive, c->active, c->mass, c->kick_dt,
					c->acc[0], c->acc[1], c->acc[2]);
		hipLaunchKernelGGL(k_top_kicks, dim3(tdv_blocks((size_t)c->ntop, 64)),
				dim3(64), 0, c->stream, c->top, c->nactive, c->active, c->mass,
				c->kick_dt, c->acc[0], c->acc[1], c->acc[2]);
	}
	(void)hipEventRecord(c->ev[EV_KICK1], c->stream);
	TDV_HIP(c, hipGetLastError());
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	(void)hipEventElapsedTime(&c->stage_ms[7], c->ev[EV_KICK0],
			c->ev[EV_KICK1]);
	return 0;
}

int tdv_drift_to_sync_point(tdv_context *c, uint32_t it_next)
{
	if (!c || !c->n || !c->have_timeline)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_drift_to_sync_point: no "
				"particles or no timeline (tdv_set_timeline)");
	if (it_next > ((uint32_t)1 << (TL_NBITS - 1)))    /* drift.c:29-34 */
		return tdv_fail(c, TDV_ERR_ARG, "tdv_drift_to_sync_point: next "
				"time beyond the end of the timeline");
	(void)hipEventRecord(c->ev[EV_KICK0], c->stream);
	if (c->nactive)
		hipLaunchKernelGGL(k_drift_sync, dim3(tdv_blocks((size_t)c->nactive,
				256)), dim3(256), 0, c->stream, c->nactive, c->active,
				c->it_drift, it_next, c->pos[0], c->pos[1], c->pos[2],
				c->vel[0], c->vel[1], c->vel[2], timeline_dev(c),
				c->par.periodic, c->par.boxsize, c->pm);
	(void)hipEventRecord(c->ev[EV_KICK1], c->stream);
	TDV_HIP(c, hipGetLastError());
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	(void)hipEventElapsedTime(&c->stage_ms[7], c->ev[EV_KICK0],
			c->ev[EV_KICK1]);
	return 0;
}

int tdv_set_new_timebins(tdv_context *c, const tdv_timestep_params *p,
		int *bin_min, int *bin_max)
{
	if (!c || !p || !c->n || !c->have_timeline)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_new_timebins: no particles, "
				"parameters or timeline");
	if (p->max_active_bin < 0 || p->max_active_bin >= TL_NBITS)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_new_timebins: "
				"max_active_bin outside 0..31");
	TimebinDev d;
	d.comoving = c->tl.comoving;
	d.accuracy = p->time_int_accuracy;
	d.epsilon = 105.0 / 32.0 * p->softening;   /* timestep.c:518 */
	d.dt_max = (double)(float)p->dt_max_global;  /* a float, timestep.c:26 */
	d.step_min = c->tl_step_min;
	d.step_max = c->tl_step_max;
	d.expansion = p->expansion_factor;
	d.hubble = p->hubble_parameter;
	d.accel_factor = p->grav_accel_factor;
	d.max_active_bin = p->max_active_bin;
	int *mm = reinterpret_cast<int *>(c->tl_spline + 5 * TL_KNOTS);
	hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(1), 0, c->stream, mm);
	hipLaunchKernelGGL(k_timebins, dim3(tdv_blocks(c->n, 256)), dim3(256), 0,
			c->stream, c->n, c->gacc[0], c->gacc[1], c->gacc[2], c->time_bin,
			d, mm, c->d_err);
	int h[2];
	TDV_HIP(c, hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost,
			c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	if (bin_min)
		*bin_min = h[0];
	if (bin_max)
		*bin_max = h[1];
	unsigned e = 0;
	TDV_HIP(c, hipMemcpy(&e, c->d_err, sizeof(e), hipMemcpyDeviceToHost));
	if (e & TDV_EBIT_RANGE) {
		TDV_HIP(c, hipMemset(c->d_err, 0, sizeof(unsigned)));
		return tdv_fail(c, TDV_ERR_RANGE, "tdv_set_new_timebins: a time "
				"step below the timeline's resolution or not finite");
	}
	return tdv_check_err_flags(c);
}

} /* extern "C" */

extern "C" int tdv_tree_update_drift(tdv_context *c, double dt)
{
	if (!c || !c->have_tree || !std::isfinite(dt))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_tree_update_drift: no tree or "
				"dt not finite");
	if (c->nnodes)
		hipLaunchKernelGGL(k_tree_drift, dim3(tdv_blocks(c->nnodes, 256)),
				dim3(256), 0, c->stream, c->nnodes, c->node_c, c->node_a,
				c->node_dp, dt);
	hipLaunchKernelGGL(k_top_drift, dim3(tdv_blocks((size_t)c->ntop, 64)),
			dim3(64), 0, c->stream, c->top, dt);
	TDV_HIP(c, hipGetLastError());
	return tdv_check(c, hipStreamSynchronize(c->stream), "tree drift");
}
