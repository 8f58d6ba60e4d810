/*
 * FP64 tree walks, one lane per active particle, reproducing the
 * reference's per-particle decisions and summation order:
 *   k_walk64        Gravity_Tree_Acceleration (Gravity/tree_accel.c:32-394):
 *                   top-node loop, gravity_tree_walk (relative criterion +
 *                   "inside" test) or gravity_tree_walk_BH, interact()
 *   k_walk64_ewald  Gravity_Tree_Periodic (Gravity/tree_periodic.c:23-450)
 *                   with Ewald_Correction (Gravity/periodic.c:35-113)
 * Opening decisions use the same FP64 expressions in the same order as the
 * reference, so the interaction set and P.Cost are identical; the build is
 * compiled without FP contraction so every product and sum rounds once.
 * Lanes of a wavefront are PH neighbours whose walks visit nearly the same
 * nodes, so the wavefront walks in lockstep (see k_walk64): node and
 * particle records are loaded once per wavefront through the scalar cache.
 */
#include "tdv_internal.h"
#include "tdv_device.h"
#include <climits>

struct WalkParams {
	int ntop, vs, compat, periodic, use_bh;
	double G, theta_bh, theta_bh2, alpha_rel, eps2, eps3, box, boxhalf;
	double dom_size, box2ewald;
};

/* Periodic_Nearest (periodic.c:62-80) */
__device__ __forceinline__ double wrap(double d, const WalkParams &w)
{
	if (!w.periodic)
		return d;
	if (d > w.boxhalf)
		d -= w.box;
	else if (d < -w.boxhalf)
		d += w.box;
	return d;
}

struct Acc {
	double x, y, z, cost;
};

/* sum of P.Cost over the launch: one atomic per wavefront */
__device__ __forceinline__ void count_interactions(double cost,
		unsigned long long *counter)
{
	unsigned long long v = (unsigned long long)cost;
	for (int d = 32; d >= 1; d >>= 1)
		v += __shfl_xor(v, d, 64);
	if ((threadIdx.x & 63) == 0 && v)
		atomicAdd(counter, v);
}

/* interact (Gravity/tree_accel.c:362-394): Dehnen K1 kernel, h < 0 */
__device__ __forceinline__ void interact(Acc &a, double m, double dx,
		double dy, double dz, double r2, const WalkParams &w)
{
	double fac = w.G * m;
	if (r2 < w.eps2) {
		const double u2 = r2 / w.eps2;
		fac *= (175 - u2 * (294 - u2 * 135)) / (16 * w.eps3);
	} else {
		const double r_inv = 1 / sqrt(r2);
		fac *= r_inv * r_inv * r_inv;
	}
	a.x += fac * dx;
	a.y += fac * dy;
	a.z += fac * dz;
	a.cost += 1;
}

/* P-P with the particles [first, last) for the lanes with `on`, in the
 * reference's order; the range is the wavefront's, so the loads are
 * uniform */
__device__ __forceinline__ void pp_range_on(Acc &a, const double4 *pm,
		int first, int last, double px, double py, double pz, bool on,
		const WalkParams &w)
{
	for (int j = first; j < last; j++) {
		const double4 q = pm[j];
		if (!on)
			continue;
		const double dx = wrap(q.x - px, w), dy = wrap(q.y - py, w),
				dz = wrap(q.z - pz, w);
		const double r2 = dx * dx + dy * dy + dz * dz;
		if (r2 != 0)
			interact(a, q.w, dx, dy, dz, r2, w);
	}
}

/* the smallest value over the wavefront, as a wave-uniform value */
/* min of v and the value DPP control `ctrl` moves into this lane (INT_MAX
 * where nothing arrives) */
template <int ctrl, int rows>
__device__ __forceinline__ int dpp_min(int v)
{
	const int o = __builtin_amdgcn_update_dpp(INT_MAX, v, ctrl, rows, 0xf,
			false);
	return o < v ? o : v;
}

/* The wavefront's minimum, in registers: a prefix minimum by DPP (row_shr
 * 1, 2, 4, 8 within rows of 16, row_bcast 15 and 31 across rows) leaves
 * it in lane 63.  Every lane of the wavefront takes part. */
__device__ __forceinline__ int wave_min(int v)
{
	v = dpp_min<0x111, 0xf>(v);
	v = dpp_min<0x112, 0xf>(v);
	v = dpp_min<0x114, 0xf>(v);
	v = dpp_min<0x118, 0xf>(v);
	v = dpp_min<0x142, 0xa>(v);
	v = dpp_min<0x143, 0xc>(v);
	return __builtin_amdgcn_readlane(v, 63);
}

/* One lane per active particle, every lane taking exactly the reference's
 * walk: its own decisions, its own interactions in its own order.  The
 * wavefront walks in lockstep: each lane keeps the next node of its walk,
 * the wavefront takes the smallest of them (walks only move forward in
 * the depth-first array), loads that node once through the scalar cache,
 * and the lanes whose next node it is act on it; the others wait.  PH
 * neighbours' walks share most nodes, so the wavefront visits little more
 * than one walk's nodes, without the per-lane gathers and divergent
 * branches of independent loops. */
__global__ void __launch_bounds__(256) k_walk64(int nact,
		const int *__restrict__ active, const double4 *__restrict__ pm,
		const double *__restrict__ last_acc, TopNodes T,
		const double4 *__restrict__ na, const double4 *__restrict__ nb,
		const int4 *__restrict__ nc, WalkParams w, double *__restrict__ ax,
		double *__restrict__ ay, double *__restrict__ az,
		double *__restrict__ gx, double *__restrict__ gy,
		double *__restrict__ gz, double *__restrict__ cost, int npart_total,
		unsigned *err, unsigned long long *counter)
{
	const int gid = blockIdx.x * blockDim.x + threadIdx.x;
	const bool live = gid < nact;
	const int ip = live ? active[gid] : 0;
	const double4 p = pm[ip];
	const double fac = last_acc[ip] / w.G * w.alpha_rel;
	Acc a = { 0, 0, 0, 0 };

	for (int j = 0; j < w.ntop; j++) {
		/* interact_with_topnode, Gravity/tree_accel.c:137-176 */
		const double nSize = w.dom_size / ((double)(1ull << T.level[j]));
		bool walk = false;
		if (live) {
			const double dx = (double)T.pos[3 * j] - p.x;
			const double dy = (double)T.pos[3 * j + 1] - p.y;
			const double dz = (double)T.pos[3 * j + 2] - p.z;
			const bool inside = fabs(dx) < 0.6 * nSize && fabs(dy) < 0.6 * nSize
					&& fabs(dz) < 0.6 * nSize;
			walk = inside;
			if (!inside) {
				const double rx = wrap((double)T.com[3 * j] - p.x, w);
				const double ry = wrap((double)T.com[3 * j + 1] - p.y, w);
				const double rz = wrap((double)T.com[3 * j + 2] - p.z, w);
				const double r2 = rx * rx + ry * ry + rz * rz;
				const double m = (double)T.mass[j];
				const bool open = w.use_bh ? (nSize * nSize > r2 * w.theta_bh)
						: (m * nSize * nSize > r2 * r2 * fac);
				if (open)
					walk = true;
				else
					interact(a, m, rx, ry, rz, r2, w);
			}
		}
		if (!__any(walk))
			continue;
		if (T.npart[j] <= w.vs) {
			/* interact_with_topnode_particles, tree_accel.c:183-203 */
			const int f = w.compat ? T.target[j] : T.first[j];
			if (f < 0 || f + T.npart[j] > npart_total) {
				if (walk)
					atomicOr(err, TDV_EBIT_RANGE);
				continue;
			}
			pp_range_on(a, pm, f, f + T.npart[j], p.x, p.y, p.z, walk, w);
			continue;
		}
		/* gravity_tree_walk(_BH), Gravity/tree_accel.c:211-353 */
		const int start = T.target[j], term = T.term[j];
		int next = walk ? start : INT_MAX;
		for (;;) {
			const int node = wave_min(next);
			if (node == INT_MAX)
				break;
			if (node < start || node > term) {  /* never for a valid tree */
				atomicOr(err, TDV_EBIT_RANGE);
				break;
			}
			const bool mine = next == node;
			const int4 c = nc[node];
			if (c.x == 0) {                     /* the terminator */
				if (node == start)
					atomicOr(err, TDV_EBIT_RANGE);
				if (mine)
					next = INT_MAX;
				continue;
			}
			if (c.x < 0) {
				const int f = -c.x - 1;
				if (f + c.y > npart_total) {
					atomicOr(err, TDV_EBIT_RANGE);
					break;
				}
				pp_range_on(a, pm, f, f + c.y, p.x, p.y, p.z, mine, w);
				if (mine)
					next = node + 1;
				continue;
			}
			const double4 A = na[node];
			const double4 B = nb[node];
			if (!mine)
				continue;
			const double rx = wrap(A.x - p.x, w), ry = wrap(A.y - p.y, w),
					rz = wrap(A.z - p.z, w);
			const double r2 = rx * rx + ry * ry + rz * rz;
			const double m = A.w;
			const double s = B.w;
			bool open;
			if (w.use_bh) {
				open = s * s > r2 * w.theta_bh;
			} else {
				open = m * s * s > r2 * r2 * fac;
				if (!open) {
					const double sx = wrap(B.x - p.x, w),
							sy = wrap(B.y - p.y, w), sz = wrap(B.z - p.z, w);
					open = fabs(sx) < 0.6 * s && fabs(sy) < 0.6 * s
							&& fabs(sz) < 0.6 * s;
				}
			}
			if (open) {
				next = node + 1;
			} else {
				interact(a, m, rx, ry, rz, r2, w);
				next = node + c.x;
			}
		}
	}
	/* add_recv_to, Gravity/tree_accel.c:110-129 */
	if (live) {
		ax[ip] = ax[ip] + a.x;
		ay[ip] = ay[ip] + a.y;
		az[ip] = az[ip] + a.z;
		gx[ip] = a.x;
		gy[ip] = a.y;
		gz[ip] = a.z;
		cost[ip] = a.cost;
	}
	count_interactions(a.cost, counter);
}

__device__ __forceinline__ void ewald_cube(Acc &a, const double *tab,
		double dx, double dy, double dz, double m, const WalkParams &w)
{
	double f[3];
	ewald_corr(tab, dx, dy, dz, w.box2ewald, f);
	a.x += w.G * m * f[0];
	a.y += w.G * m * f[1];
	a.z += w.G * m * f[2];
	a.cost += 1;
}

__device__ __forceinline__ void ewald_pp_range_on(Acc &a, const double *tab,
		const double4 *pm, int first, int last, double px, double py,
		double pz, bool on, const WalkParams &w)
{
	for (int j = first; j < last; j++) {
		const double4 q = pm[j];
		if (!on)
			continue;
		const double dx = wrap(q.x - px, w), dy = wrap(q.y - py, w),
				dz = wrap(q.z - pz, w);
		if (dx != 0 && dy != 0 && dz != 0)
			ewald_cube(a, tab, dx, dy, dz, q.w, w);
	}
}

/* Gravity_Tree_Periodic per lane, in lockstep as k_walk64 */
__global__ void __launch_bounds__(256) k_walk64_ewald(int nact,
		const int *__restrict__ active, const double4 *__restrict__ pm,
		const double *__restrict__ last_acc, TopNodes T,
		const double4 *__restrict__ na, const double4 *__restrict__ nb,
		const int4 *__restrict__ nc, const double *__restrict__ tab,
		WalkParams w, double *__restrict__ ax, double *__restrict__ ay,
		double *__restrict__ az, double *__restrict__ gx,
		double *__restrict__ gy, double *__restrict__ gz,
		double *__restrict__ cost, int npart_total, unsigned *err,
		unsigned long long *counter)
{
	const int gid = blockIdx.x * blockDim.x + threadIdx.x;
	const bool live = gid < nact;
	const int ip = live ? active[gid] : 0;
	const double4 p = pm[ip];
	const double fac = last_acc[ip] / w.G * w.alpha_rel;
	Acc a = { 0, 0, 0, 0 };

	for (int j = 0; j < w.ntop; j++) {
		/* interact_with_topnode, Gravity/tree_periodic.c:112-177 */
		const double s = w.dom_size / (double)(1ull << T.level[j]);
		bool walk = false;
		if (live) {
			bool want_open = false;
			const double rx = wrap((double)T.com[3 * j] - p.x, w);
			const double ry = wrap((double)T.com[3 * j + 1] - p.y, w);
			const double rz = wrap((double)T.com[3 * j + 2] - p.z, w);
			const double r2 = rx * rx + ry * ry + rz * rz;
			if (w.use_bh) {
				if (s * s > r2 * w.theta_bh)
					want_open = true;
			} else {
				const double m = (double)T.mass[j];
				if (m * s * s > r2 * r2 * fac)
					want_open = true;
			}
			if (fabs(p.x - (double)T.pos[3 * j]) < 0.6 * s
					&& fabs(p.y - (double)T.pos[3 * j + 1]) < 0.6 * s
					&& fabs(p.z - (double)T.pos[3 * j + 2]) < 0.6 * s)
				want_open = true;
			/* a node that wants opening is walked only when it straddles
			 * the half box or is larger than 0.2 box; otherwise the
			 * correction of its centre of mass is used */
			if (want_open) {
				const double lim = 0.5 * (w.box - s);
				walk = fabs(rx) > lim || fabs(ry) > lim || fabs(rz) > lim
						|| s > 0.2 * w.box;
			}
			if (!walk)
				ewald_cube(a, tab, rx, ry, rz, (double)T.mass[j], w);
		}
		if (!__any(walk))
			continue;
		if (T.npart[j] <= w.vs) {
			const int f = w.compat ? T.target[j] : T.first[j];
			if (f < 0 || f + T.npart[j] > npart_total) {
				if (walk)
					atomicOr(err, TDV_EBIT_RANGE);
				continue;
			}
			ewald_pp_range_on(a, tab, pm, f, f + T.npart[j], p.x, p.y, p.z,
					walk, w);
			continue;
		}
		/* gravity_tree_walk_ewald(_BH), Gravity/tree_periodic.c:216-422 */
		const int start = T.target[j];
		const int end = start + nc[start].x;
		if (end != T.term[j]) {
			if (walk)
				atomicOr(err, TDV_EBIT_RANGE);
			continue;
		}
		int next = walk ? start : INT_MAX;
		for (;;) {
			const int node = wave_min(next);
			if (node == INT_MAX)
				break;
			if (node < start || node >= end) {   /* never for a valid tree */
				atomicOr(err, TDV_EBIT_RANGE);
				break;
			}
			const bool mine = next == node;
			const int4 c = nc[node];
			if (c.x < 0) {
				const int f = -c.x - 1;
				if (f + c.y > npart_total) {
					atomicOr(err, TDV_EBIT_RANGE);
					break;
				}
				ewald_pp_range_on(a, tab, pm, f, f + c.y, p.x, p.y, p.z, mine,
						w);
				if (mine)
					next = node + 1 == end ? INT_MAX : node + 1;
				continue;
			}
			const double4 A = na[node];
			const double4 B = nb[node];
			if (!mine)
				continue;
			const double dx = wrap(A.x - p.x, w), dy = wrap(A.y - p.y, w),
					dz = wrap(A.z - p.z, w);
			const double d2 = dx * dx + dy * dy + dz * dz;
			const double m = A.w;
			const double ns = B.w;
			const double sx = B.x - p.x, sy = B.y - p.y, sz = B.z - p.z;
			bool open = w.use_bh ? (ns * ns > d2 * w.theta_bh2)
					: (m * ns * ns > d2 * d2 * fac);
			if (!open && fabs(sx) < 0.6 * ns && fabs(sy) < 0.6 * ns
					&& fabs(sz) < 0.6 * ns)
				open = true;
			int to;
			if (open) {
				const double lim = 0.5 * (w.box - ns);
				if (fabs(wrap(sx, w)) > lim || fabs(wrap(sy, w)) > lim
						|| fabs(wrap(sz, w)) > lim || ns > 0.2 * w.box)
					to = node + 1;
				else
					to = -1;
			} else {
				to = -1;
			}
			if (to < 0) {
				ewald_cube(a, tab, dx, dy, dz, m, w);
				to = node + c.x;
			}
			next = to == end ? INT_MAX : to;
		}
	}
	/* add_recv_to, Gravity/tree_periodic.c:91-110 */
	if (live) {
		ax[ip] = ax[ip] + a.x;
		ay[ip] = ay[ip] + a.y;
		az[ip] = az[ip] + a.z;
		gx[ip] = gx[ip] + a.x;
		gy[ip] = gy[ip] + a.y;
		gz[ip] = gz[ip] + a.z;
		cost[ip] = cost[ip] + a.cost;
	}
	count_interactions(a.cost, counter);
}

static WalkParams walk_params(const tdv_context *c, int use_bh)
{
	WalkParams w;
	w.ntop = c->ntop;
	w.vs = c->par.vector_size;
	w.compat = c->par.compat;
	w.periodic = c->par.periodic;
	w.use_bh = use_bh;
	w.G = c->par.gravity;
	w.theta_bh = c->par.theta_bh;
	w.theta_bh2 = c->par.theta_bh * c->par.theta_bh;
	w.alpha_rel = c->par.alpha_rel;
	w.eps2 = c->Epsilon2;
	w.eps3 = c->Epsilon3;
	w.box = c->par.boxsize;
	w.boxhalf = c->par.boxsize / 2.0;
	w.dom_size = c->dom_size;
	w.box2ewald = 2 * TDV_NE / c->par.boxsize;
	return w;
}

int tdv_walk64(tdv_context *c, int use_bh)
{
	if (!c->nactive)
		return 0;
	const WalkParams w = walk_params(c, use_bh);
	hipLaunchKernelGGL(k_walk64, dim3(tdv_blocks((size_t)c->nactive, 256)),
			dim3(256), 0, c->stream, c->nactive, c->active, c->pm, c->last_acc,
			c->top, c->node_a, c->node_b, c->node_c, w, c->acc[0], c->acc[1],
			c->acc[2], c->gacc[0], c->gacc[1], c->gacc[2], c->cost, (int)c->n,
			c->d_err, c->d_count);
	return tdv_check(c, hipGetLastError(), "k_walk64");
}

int tdv_walk64_ewald(tdv_context *c, int use_bh)
{
	if (!c->nactive)
		return 0;
	if (!c->have_ewald)
		return tdv_fail(c, TDV_ERR_ARG, "periodic walk without an Ewald table");
	const WalkParams w = walk_params(c, use_bh);
	hipLaunchKernelGGL(k_walk64_ewald,
			dim3(tdv_blocks((size_t)c->nactive, 256)), dim3(256), 0, c->stream,
			c->nactive, c->active, c->pm, c->last_acc, c->top, c->node_a,
			c->node_b, c->node_c, c->ewald, w, c->acc[0], c->acc[1], c->acc[2],
			c->gacc[0], c->gacc[1], c->gacc[2], c->cost, (int)c->n, c->d_err,
			c->d_count + 1);
	return tdv_check(c, hipGetLastError(), "k_walk64_ewald");
}

/* safe_last_accel, accel.c:28-43 */
__global__ void k_safe_last_accel(int nact, const int *__restrict__ active,
		double *__restrict__ ax, double *__restrict__ ay,
		double *__restrict__ az, double *__restrict__ last)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nact)
		return;
	const int i = active[g];
	last[i] = sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
	ax[i] = 0;
	ay[i] = 0;
	az[i] = 0;
}

void tdv_launch_safe_last_accel(tdv_context *c)
{
	if (!c->nactive)
		return;
	hipLaunchKernelGGL(k_safe_last_accel,
			dim3(tdv_blocks((size_t)c->nactive, 256)), dim3(256), 0, c->stream,
			c->nactive, c->active, c->acc[0], c->acc[1], c->acc[2],
			c->last_acc);
}

__global__ void k_ewald_correction_batch(size_t n, const double *__restrict__ dr,
		double *__restrict__ f, const double *__restrict__ tab, double b2e)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double r[3];
	ewald_corr(tab, dr[3 * i], dr[3 * i + 1], dr[3 * i + 2], b2e, r);
	f[3 * i] = r[0];
	f[3 * i + 1] = r[1];
	f[3 * i + 2] = r[2];
}

void tdv_launch_ewald_correction(tdv_context *c, size_t n, const double *dr,
		double *f)
{
	hipLaunchKernelGGL(k_ewald_correction_batch, dim3(tdv_blocks(n, 256)),
			dim3(256), 0, c->stream, n, dr, f, c->ewald,
			2 * TDV_NE / c->par.boxsize);
}
