/*
 * Ewald correction table in FP64 on the GPU -- compute_ewald_correction_table
 * (Gravity/periodic.c:235-300) with compute_ewald_force / _potential
 * (Gravity/periodic.c:352-461): Hernquist, Bouchet & Suto (1991) sums,
 * alpha = 2, |n|, |h| <= 4 per component, at r = (i, j, k) / 128 box units
 * on the 65^3 octant grid.  One lane per grid node; the output is the
 * unnormalised table the reference writes to ewald_tables.dat.
 */
#include "tdv_internal.h"

#define EW_PI 3.14159265358979323846

__global__ void __launch_bounds__(256) k_ewald_table(double *__restrict__ fx,
		double *__restrict__ fy, double *__restrict__ fz,
		double *__restrict__ fp)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= TDV_NE3)
		return;
	const int i = n / (TDV_NE1 * TDV_NE1), j = (n / TDV_NE1) % TDV_NE1,
			k = n % TDV_NE1;
	const double alpha = 2.0;
	const double x[3] = { (0.5 * i) / TDV_NE, (0.5 * j) / TDV_NE,
		(0.5 * k) / TDV_NE };
	double f[3] = { 0, 0, 0 };
	double pot;
	if (i == 0 && j == 0 && k == 0) {
		pot = 2.8372975;
	} else {
		const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
		for (int d = 0; d < 3; d++)
			f[d] = x[d] / (r * r * r);
		double s1 = 0;
		for (int a = -4; a < 5; a++)
			for (int b = -4; b < 5; b++)
				for (int c = -4; c < 5; c++) {
					const double dx[3] = { x[0] - a, x[1] - b, x[2] - c };
					const double rr = sqrt(dx[0] * dx[0] + dx[1] * dx[1]
							+ dx[2] * dx[2]);
					const double ec = erfc(alpha * rr);
					const double val = ec + 2 * alpha * rr / sqrt(EW_PI)
							* exp(-((alpha * rr) * (alpha * rr)));
					for (int d = 0; d < 3; d++)
						f[d] -= dx[d] / (rr * rr * rr) * val;
					s1 += ec / rr;
				}
		double s2 = 0;
		for (int a = -4; a < 5; a++)
			for (int b = -4; b < 5; b++)
				for (int c = -4; c < 5; c++) {
					const int h2 = a * a + b * b + c * c;
					if (h2 == 0)
						continue;
					const double hx = x[0] * a + x[1] * b + x[2] * c;
					const double e = exp(-EW_PI * EW_PI * h2 / (alpha * alpha));
					double sn, cs;
					sincos(2 * EW_PI * hx, &sn, &cs);
					const double val = 2.0 / h2 * e * sn;
					f[0] -= a * val;
					f[1] -= b * val;
					f[2] -= c * val;
					s2 += 1.0 / (EW_PI * h2) * e * cs;
				}
		pot = EW_PI / (alpha * alpha) - s1 - s2 + 1.0 / r;
	}
	fx[n] = f[0];
	fy[n] = f[1];
	fz[n] = f[2];
	fp[n] = pot;
}

int tdv_ewald_table_device(tdv_context *c, double *out4)
{
	hipLaunchKernelGGL(k_ewald_table, dim3(tdv_blocks(TDV_NE3, 256)), dim3(256),
			0, c->stream, out4, out4 + TDV_NE3, out4 + 2 * TDV_NE3,
			out4 + 3 * TDV_NE3);
	return tdv_check(c, hipGetLastError(), "k_ewald_table");
}

/*
 * Second derivatives of the Ewald correction, for taking a node's
 * correction at its centre of mass to second order (the FP32 group walk).
 *
 * F(x) is the table's function for any x, not only on the octant grid: the
 * same real-space (|n| <= 4) and Fourier (|h| <= 4) sums as k_ewald_table,
 * alpha = 2, with the n = 0 real-space term and the Newton term x / r^3
 * combined as x (erf(a r) - 2 a r / sqrt(pi) exp(-a^2 r^2)) / r^3, a series
 * below a r = 0.5, so F is smooth through x = 0.  The walk's correction is
 * E(d) = -F(d) (the table's sign rule), continued smoothly past the half-box
 * faces (the next singularity is the image at distance box).  Its second
 * derivatives S_abc = d_b d_c E_a are symmetric in (a, b, c) and traceless
 * in (b, c) away from the images; the ten independent ones come from central
 * differences of F with step H (box units): nine pure second differences
 * (three evaluations per axis) and d_y d_z E_x from four diagonal ones.
 * Layout: 10 x 65^3 doubles, component-major, in the order
 * xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz, per unit box.
 */
#define EW_H (1.0 / 512.0)

__device__ void ewald_F_any(const double x[3], double f[3])
{
	const double alpha = 2.0;
	const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
	const double r = sqrt(r2);
	const double z = alpha * r;
	/* x g(z) / r^3 with g(z) = erf(z) - 2 z / sqrt(pi) exp(-z^2) */
	double gr3;
	if (z < 0.5) {
		/* g(z) = 2/sqrt(pi) sum_{n>=1} (-1)^(n+1) 2n z^(2n+1) / (n! (2n+1)) */
		const double z2 = z * z;
		double term = 1.0, sum = 0.0, zp = 1.0;  /* zp = z^(2n-2) */
		for (int n = 1; n <= 14; n++) {
			term = (n == 1) ? 1.0 : term / n;  /* 1 / n! */
			const double c = ((n & 1) ? 1.0 : -1.0) * 2.0 * n / (2.0 * n + 1.0);
			sum += c * term * zp;
			zp *= z2;
		}
		/* g / r^3 = 2/sqrt(pi) alpha^3 sum_n c_n z^(2n-2) / n! */
		gr3 = 2.0 / sqrt(EW_PI) * alpha * alpha * alpha * sum;
	} else {
		gr3 = (erf(z) - 2.0 * z / sqrt(EW_PI) * exp(-z * z)) / (r2 * r);
	}
	for (int d = 0; d < 3; d++)
		f[d] = x[d] * gr3;
	for (int a = -4; a < 5; a++)
		for (int b = -4; b < 5; b++)
			for (int c = -4; c < 5; c++) {
				if (a == 0 && b == 0 && c == 0)
					continue;
				const double dx[3] = { x[0] - a, x[1] - b, x[2] - c };
				const double rr = sqrt(dx[0] * dx[0] + dx[1] * dx[1]
						+ dx[2] * dx[2]);
				const double val = erfc(alpha * rr) + 2 * alpha * rr
						/ sqrt(EW_PI) * exp(-((alpha * rr) * (alpha * rr)));
				for (int d = 0; d < 3; d++)
					f[d] -= dx[d] / (rr * rr * rr) * val;
			}
	for (int a = -4; a < 5; a++)
		for (int b = -4; b < 5; b++)
			for (int c = -4; c < 5; c++) {
				const int h2 = a * a + b * b + c * c;
				if (h2 == 0)
					continue;
				const double hx = x[0] * a + x[1] * b + x[2] * c;
				const double e = exp(-EW_PI * EW_PI * h2 / (alpha * alpha));
				const double val = 2.0 / h2 * e * sin(2 * EW_PI * hx);
				f[0] -= a * val;
				f[1] -= b * val;
				f[2] -= c * val;
			}
}

__global__ void __launch_bounds__(64) k_ewald_derivs(double *__restrict__ s,
		double *__restrict__ jac)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= TDV_NE3)
		return;
	const int i = n / (TDV_NE1 * TDV_NE1), j = (n / TDV_NE1) % TDV_NE1,
			k = n % TDV_NE1;
	const double x0[3] = { (0.5 * i) / TDV_NE, (0.5 * j) / TDV_NE,
		(0.5 * k) / TDV_NE };
	const double H = EW_H;
	double c0[3], p[3], m[3];
	ewald_F_any(x0, c0);
	double d2[3][3];                         /* d2[axis][component] of F */
	for (int ax = 0; ax < 3; ax++) {
		double xp[3] = { x0[0], x0[1], x0[2] }, xm[3] = { x0[0], x0[1], x0[2] };
		xp[ax] += H;
		xm[ax] -= H;
		ewald_F_any(xp, p);
		ewald_F_any(xm, m);
		for (int q = 0; q < 3; q++)
			d2[ax][q] = (p[q] - 2.0 * c0[q] + m[q]) / (H * H);
	}
	double pp[3], pm[3], mp[3], mm[3];
	{
		double x1[3] = { x0[0], x0[1] + H, x0[2] + H };
		ewald_F_any(x1, pp);
		double x2[3] = { x0[0], x0[1] + H, x0[2] - H };
		ewald_F_any(x2, pm);
		double x3[3] = { x0[0], x0[1] - H, x0[2] + H };
		ewald_F_any(x3, mp);
		double x4[3] = { x0[0], x0[1] - H, x0[2] - H };
		ewald_F_any(x4, mm);
	}
	const double dyz_x = (pp[0] - pm[0] - mp[0] + mm[0]) / (4.0 * H * H);
	/* E = -F */
	const double v[10] = { -d2[0][0], -d2[1][1], -d2[2][2], -d2[0][1],
		-d2[0][2], -d2[1][0], -d2[1][2], -d2[2][0], -d2[2][1], -dyz_x };
	for (int q = 0; q < 10; q++)
		s[(size_t)q * TDV_NE3 + n] = v[q];
	if (jac) {
		/* J_ab = d_b E_a from the same central differences, symmetrised
		 * (E is a gradient): xx yy zz xy xz yz */
		double pj[3][3];                  /* pj[axis][component] = d F */
		for (int ax = 0; ax < 3; ax++) {
			double xp[3] = { x0[0], x0[1], x0[2] },
					xm[3] = { x0[0], x0[1], x0[2] };
			xp[ax] += H;
			xm[ax] -= H;
			ewald_F_any(xp, p);
			ewald_F_any(xm, m);
			for (int q = 0; q < 3; q++)
				pj[ax][q] = (p[q] - m[q]) / (2.0 * H);
		}
		const double jv[6] = { -pj[0][0], -pj[1][1], -pj[2][2],
			-0.5 * (pj[1][0] + pj[0][1]), -0.5 * (pj[2][0] + pj[0][2]),
			-0.5 * (pj[2][1] + pj[1][2]) };
		for (int q = 0; q < 6; q++)
			jac[(size_t)q * TDV_NE3 + n] = jv[q];
	}
}

int tdv_ewald_derivs_device(tdv_context *c, double *out10, double *out6)
{
	hipLaunchKernelGGL(k_ewald_derivs, dim3(tdv_blocks(TDV_NE3, 64)), dim3(64),
			0, c->stream, out10, out6);
	return tdv_check(c, hipGetLastError(), "k_ewald_derivs");
}

/*
 * FP64 direct summation with the exact Ewald correction of every pair: the
 * same pairwise Newton kernel as k_direct (nearest image, K1 softening),
 * plus -F(d / Box) / Box^2 from the Ewald sums themselves (ewald_F_any)
 * instead of the 65^3 table's trilinear interpolation.  An error oracle for
 * the walks that is free of the table's interpolation error.  Every pair
 * costs a full Ewald sum, so the sources of each target are cut into
 * chunks (blockIdx.y) to fill the chip whatever the number of targets;
 * k_exact_chunks_sum adds the chunks' sums in order.
 */
__global__ void __launch_bounds__(256) k_direct_exact(int nt,
		const int *__restrict__ idx, const double4 *__restrict__ pm, int n,
		int chunk, double box, double G, double eps2, double eps3,
		double *__restrict__ part)
{
	__shared__ double red[3][256];
	const int t = blockIdx.x;
	if (t >= nt)
		return;
	const int i = idx[t];
	const double4 p = pm[i];
	const double half = box / 2.0, ib = 1.0 / box;
	const long long j0 = (long long)blockIdx.y * chunk;
	const long long j1 = j0 + chunk < n ? j0 + chunk : n;
	double ax = 0, ay = 0, az = 0;
	for (long long jj = j0 + threadIdx.x; jj < j1; jj += blockDim.x) {
		const int j = (int)jj;
		if (j == i)
			continue;
		const double4 q = pm[j];
		double d[3] = { q.x - p.x, q.y - p.y, q.z - p.z };
		for (int k = 0; k < 3; k++) {
			if (d[k] > half)
				d[k] -= box;
			else if (d[k] < -half)
				d[k] += box;
		}
		const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		if (r2 == 0)
			continue;
		double fac = G * q.w;
		if (r2 < eps2) {
			const double u2 = r2 / eps2;
			fac *= (175 - u2 * (294 - u2 * 135)) / (16 * eps3);
		} else {
			const double ri = 1 / sqrt(r2);
			fac *= ri * ri * ri;
		}
		const double du[3] = { d[0] * ib, d[1] * ib, d[2] * ib };
		double f[3];
		ewald_F_any(du, f);
		const double ge = G * q.w * ib * ib;
		ax += fac * d[0] - ge * f[0];
		ay += fac * d[1] - ge * f[1];
		az += fac * d[2] - ge * f[2];
	}
	red[0][threadIdx.x] = ax;
	red[1][threadIdx.x] = ay;
	red[2][threadIdx.x] = az;
	__syncthreads();
	for (int s = blockDim.x / 2; s > 0; s >>= 1) {
		if ((int)threadIdx.x < s)
			for (int d = 0; d < 3; d++)
				red[d][threadIdx.x] += red[d][threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		for (int d = 0; d < 3; d++)
			part[((size_t)t * gridDim.y + blockIdx.y) * 3 + d] = red[d][0];
}

__global__ void k_exact_chunks_sum(int nt, int nchunk,
		const double *__restrict__ part, double *__restrict__ out)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nt)
		return;
	double a[3] = { 0, 0, 0 };
	for (int ch = 0; ch < nchunk; ch++)
		for (int d = 0; d < 3; d++)
			a[d] += part[((size_t)t * nchunk + ch) * 3 + d];
	for (int d = 0; d < 3; d++)
		out[3 * t + d] = a[d];
}

int tdv_direct_sum_exact(tdv_context *c, int nt, const int *idx_dev,
		double *out_dev)
{
	if (nt <= 0)
		return 0;
	if (!c->par.periodic)
		return tdv_direct_sum(c, nt, idx_dev, out_dev);
	/* about 4096 workgroups in all, chunks of at least 4096 sources */
	const size_t n = c->n;
	size_t nchunk = (4096 + (size_t)nt - 1) / (size_t)nt;
	if (nchunk > (n + 4095) / 4096)
		nchunk = (n + 4095) / 4096;
	if (nchunk < 1)
		nchunk = 1;
	if (nchunk > 65535)
		nchunk = 65535;
	const size_t chunk = (n + nchunk - 1) / nchunk;
	double *part = nullptr;
	TDV_HIP(c, hipMalloc((void **)&part, (size_t)nt * nchunk * 3 * sizeof(double)));
	hipLaunchKernelGGL(k_direct_exact, dim3(nt, (unsigned)nchunk), dim3(256), 0,
			c->stream, nt, idx_dev, c->pm, (int)n, (int)chunk, c->par.boxsize,
			c->par.gravity, c->Epsilon2, c->Epsilon3, part);
	hipLaunchKernelGGL(k_exact_chunks_sum, dim3(tdv_blocks((size_t)nt, 256)),
			dim3(256), 0, c->stream, nt, (int)nchunk, part, out_dev);
	const int r = tdv_check(c, hipGetLastError(), "k_direct_exact");
	const int rs = tdv_check(c, hipStreamSynchronize(c->stream),
			"k_direct_exact");
	(void)hipFree(part);
	return r ? r : rs;
}

/*
 * The same exact periodic force, cheaper: the Ewald split at a larger
 * screening parameter.  For a unit mass at separation x (box units) the
 * nearest-image Newton term plus the correction is the lattice sum
 *   P(x) = sum_n (x - n) v(|x - n|) / |x - n|^3 + sum_{h != 0} c_h h sin(2 pi h.x),
 *   v(r) = erfc(a r) + 2 a r / sqrt(pi) exp(-a^2 r^2),
 *   c_h = 2 / h^2 exp(-pi^2 h^2 / a^2),
 * whatever the screening a (the table's sums use a = 2 with |n|, |h| <= 4
 * per component; both are converged far below FP64 rounding).  With a = 4
 * the real-space images beyond |x - n| = 6.5 / a and the wave vectors with
 * h^2 >= 64 each contribute below 2e-18, so they are left out.  The
 * real-space part stays per pair (about 17 images); the Fourier part of all
 * sources together is
 *   sum_j m_j sin(2 pi h.(x_j - x_i)) = S_h cos(2 pi h.x_i) - C_h sin(2 pi h.x_i)
 * with the structure factors S_h = sum_j m_j sin(2 pi h.x_j),
 * C_h = sum_j m_j cos(2 pi h.x_j) formed once over all particles (half of
 * the h, since the terms of h and -h are equal).  The target itself adds
 * nothing in exact arithmetic (its real-space images cancel in pairs, its
 * Fourier term is sin 0), so it is skipped in the pair sum as in
 * k_direct_exact.
 */
#define SPLIT_ALPHA 4.0
#define SPLIT_RC (6.5 / SPLIT_ALPHA)
#define SPLIT_H2MAX 63
#define SPLIT_HT 5                  /* wave vectors per thread (<= 1280) */

/* x g(a r) / r^3, g(z) = erf(z) - 2 z / sqrt(pi) exp(-z^2): the n = 0
 * real-space term minus Newton's x / r^3, as a series below a r = 0.5 */
__device__ __forceinline__ double split_g_r3(double r2, double a)
{
	const double r = sqrt(r2);
	const double z = a * r;
	if (z < 0.5) {
		const double z2 = z * z;
		double term = 1.0, sum = 0.0, zp = 1.0;
		for (int n = 1; n <= 14; n++) {
			term = (n == 1) ? 1.0 : term / n;
			const double c = ((n & 1) ? 1.0 : -1.0) * 2.0 * n / (2.0 * n + 1.0);
			sum += c * term * zp;
			zp *= z2;
		}
		return 2.0 / sqrt(EW_PI) * a * a * a * sum;
	}
	return (erf(z) - 2.0 * z / sqrt(EW_PI) * exp(-z * z)) / (r2 * r);
}

__global__ void __launch_bounds__(256) k_direct_split(int nt,
		const int *__restrict__ idx, const double4 *__restrict__ pm, int n,
		int chunk, double box, double G, double eps2, double eps3,
		double *__restrict__ part)
{
	__shared__ double red[3][256];
	const int t = blockIdx.x;
	if (t >= nt)
		return;
	const int i = idx[t];
	const double4 p = pm[i];
	const double half = box / 2.0, ib = 1.0 / box;
	const double a = SPLIT_ALPHA, rc = SPLIT_RC, rc2 = rc * rc;
	const long long j0 = (long long)blockIdx.y * chunk;
	const long long j1 = j0 + chunk < n ? j0 + chunk : n;
	double ax = 0, ay = 0, az = 0;
	for (long long jj = j0 + threadIdx.x; jj < j1; jj += blockDim.x) {
		const int j = (int)jj;
		if (j == i)
			continue;
		const double4 q = pm[j];
		double d[3] = { q.x - p.x, q.y - p.y, q.z - p.z };
		for (int k = 0; k < 3; k++) {
			if (d[k] > half)
				d[k] -= box;
			else if (d[k] < -half)
				d[k] += box;
		}
		const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		if (r2 == 0)
			continue;
		double fac = G * q.w;
		if (r2 < eps2) {
			const double u2 = r2 / eps2;
			fac *= (175 - u2 * (294 - u2 * 135)) / (16 * eps3);
		} else {
			const double ri = 1 / sqrt(r2);
			fac *= ri * ri * ri;
		}
		const double x[3] = { d[0] * ib, d[1] * ib, d[2] * ib };
		const double xr2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
		const double g0 = split_g_r3(xr2, a);
		double e[3] = { -x[0] * g0, -x[1] * g0, -x[2] * g0 };
		const int n0lo = (int)ceil(x[0] - rc), n0hi = (int)floor(x[0] + rc);
		const int n1lo = (int)ceil(x[1] - rc), n1hi = (int)floor(x[1] + rc);
		const int n2lo = (int)ceil(x[2] - rc), n2hi = (int)floor(x[2] + rc);
		for (int u = n0lo; u <= n0hi; u++)
			for (int v = n1lo; v <= n1hi; v++)
				for (int w = n2lo; w <= n2hi; w++) {
					if (u == 0 && v == 0 && w == 0)
						continue;
					const double y[3] = { x[0] - u, x[1] - v, x[2] - w };
					const double yr2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
					if (yr2 >= rc2)
						continue;
					const double yr = sqrt(yr2);
					const double val = (erfc(a * yr) + 2 * a * yr / sqrt(EW_PI)
							* exp(-(a * yr) * (a * yr))) / (yr2 * yr);
					e[0] += y[0] * val;
					e[1] += y[1] * val;
					e[2] += y[2] * val;
				}
		const double ge = G * q.w * ib * ib;
		ax += fac * d[0] + ge * e[0];
		ay += fac * d[1] + ge * e[1];
		az += fac * d[2] + ge * e[2];
	}
	red[0][threadIdx.x] = ax;
	red[1][threadIdx.x] = ay;
	red[2][threadIdx.x] = az;
	__syncthreads();
	for (int s = blockDim.x / 2; s > 0; s >>= 1) {
		if ((int)threadIdx.x < s)
			for (int d = 0; d < 3; d++)
				red[d][threadIdx.x] += red[d][threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		for (int d = 0; d < 3; d++)
			part[((size_t)t * gridDim.y + blockIdx.y) * 3 + d] = red[d][0];
}

/* structure factors per chunk of particles: thread k takes wave vectors
 * k, k + 256, ... (hv[h] = {hx, hy, hz, 0}); part[chunk * nh + h] = (S, C) */
__global__ void __launch_bounds__(256) k_sfac_chunk(
		const double4 *__restrict__ pm, long long n, long long chunk,
		const int4 *__restrict__ hv, int nh, double tpib,
		double2 *__restrict__ part)
{
	__shared__ double4 tile[256];
	const long long j0 = (long long)blockIdx.x * chunk;
	const long long j1 = j0 + chunk < n ? j0 + chunk : n;
	double s[SPLIT_HT], cs[SPLIT_HT], hx[SPLIT_HT], hy[SPLIT_HT], hz[SPLIT_HT];
#pragma unroll
	for (int k = 0; k < SPLIT_HT; k++) {
		const int h = threadIdx.x + 256 * k;
		const int4 v = h < nh ? hv[h] : make_int4(0, 0, 0, 0);
		hx[k] = v.x * tpib;
		hy[k] = v.y * tpib;
		hz[k] = v.z * tpib;
		s[k] = 0;
		cs[k] = 0;
	}
	for (long long b = j0; b < j1; b += 256) {
		__syncthreads();
		const long long j = b + threadIdx.x;
		tile[threadIdx.x] = j < j1 ? pm[j] : make_double4(0, 0, 0, 0);
		__syncthreads();
		const int cnt = (int)(j1 - b < 256 ? j1 - b : 256);
		for (int q = 0; q < cnt; q++) {
			const double4 P = tile[q];
#pragma unroll
			for (int k = 0; k < SPLIT_HT; k++) {
				double sn, c;
				sincos(hx[k] * P.x + hy[k] * P.y + hz[k] * P.z, &sn, &c);
				s[k] += P.w * sn;
				cs[k] += P.w * c;
			}
		}
	}
#pragma unroll
	for (int k = 0; k < SPLIT_HT; k++) {
		const int h = threadIdx.x + 256 * k;
		if (h < nh)
			part[(size_t)blockIdx.x * nh + h] = make_double2(s[k], cs[k]);
	}
}

/* chunk partials summed in chunk order, per wave vector */
__global__ void k_sfac_sum(int nchunk, int nh, const double2 *__restrict__ part,
		double2 *__restrict__ sf)
{
	const int h = blockIdx.x * blockDim.x + threadIdx.x;
	if (h >= nh)
		return;
	double s = 0, c = 0;
	for (int k = 0; k < nchunk; k++) {
		const double2 v = part[(size_t)k * nh + h];
		s += v.x;
		c += v.y;
	}
	sf[h] = make_double2(s, c);
}

/* per target: the real-space chunks in order, then the Fourier part
 * 2 sum_h c_h h (S_h cos b - C_h sin b), b = 2 pi h.x_i, times G / box^2 */
__global__ void k_split_finish(int nt, int nchunk, const int *__restrict__ idx,
		const double4 *__restrict__ pm, const double *__restrict__ part,
		const int4 *__restrict__ hv, const double2 *__restrict__ sf, int nh,
		double box, double G, double *__restrict__ out)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nt)
		return;
	double a[3] = { 0, 0, 0 };
	for (int ch = 0; ch < nchunk; ch++)
		for (int d = 0; d < 3; d++)
			a[d] += part[((size_t)t * nchunk + ch) * 3 + d];
	const double4 p = pm[idx[t]];
	const double tpib = 2.0 * EW_PI / box;
	const double al = SPLIT_ALPHA;
	double f[3] = { 0, 0, 0 };
	for (int h = 0; h < nh; h++) {
		const int4 v = hv[h];
		const int h2 = v.x * v.x + v.y * v.y + v.z * v.z;
		const double ch = 2.0 / h2 * exp(-EW_PI * EW_PI * h2 / (al * al));
		double sn, cs;
		sincos(tpib * (v.x * p.x + v.y * p.y + v.z * p.z), &sn, &cs);
		const double w = 2.0 * ch * (sf[h].x * cs - sf[h].y * sn);
		f[0] += w * v.x;
		f[1] += w * v.y;
		f[2] += w * v.z;
	}
	const double g = G / (box * box);
	for (int d = 0; d < 3; d++)
		out[3 * t + d] = a[d] + g * f[d];
}

int tdv_direct_sum_split(tdv_context *c, int nt, const int *idx_dev,
		double *out_dev)
{
	if (nt <= 0)
		return 0;
	if (!c->par.periodic)
		return tdv_direct_sum(c, nt, idx_dev, out_dev);
	/* wave vectors of the half space h > 0 (lexicographic), h^2 <= 63 */
	std::vector<int4> hv;
	for (int x = 0; x <= 7; x++)
		for (int y = -7; y <= 7; y++)
			for (int z = -7; z <= 7; z++) {
				const int h2 = x * x + y * y + z * z;
				if (h2 == 0 || h2 > SPLIT_H2MAX)
					continue;
				if (x == 0 && (y < 0 || (y == 0 && z < 0)))
					continue;
				hv.push_back(make_int4(x, y, z, 0));
			}
	const int nh = (int)hv.size();
	if (nh > 256 * SPLIT_HT)
		return tdv_fail(c, TDV_ERR_ARG, "too many wave vectors");
	const size_t n = c->n;
	size_t nchunk = (4096 + (size_t)nt - 1) / (size_t)nt;
	if (nchunk > (n + 4095) / 4096)
		nchunk = (n + 4095) / 4096;
	if (nchunk < 1)
		nchunk = 1;
	if (nchunk > 65535)
		nchunk = 65535;
	const size_t chunk = (n + nchunk - 1) / nchunk;
	/* structure-factor chunks: about 2048 workgroups */
	const size_t schunk = ((n + 2047) / 2048 + 255) / 256 * 256;
	const size_t nsch = (n + schunk - 1) / schunk;
	double *part = nullptr;
	int4 *dh = nullptr;
	double2 *spart = nullptr, *sf = nullptr;
	int r = tdv_check(c, hipMalloc((void **)&part,
			(size_t)nt * nchunk * 3 * sizeof(double)), "hipMalloc");
	if (!r)
		r = tdv_check(c, hipMalloc((void **)&dh, nh * sizeof(int4)), "hipMalloc");
	if (!r)
		r = tdv_check(c, hipMalloc((void **)&spart,
				nsch * nh * sizeof(double2)), "hipMalloc");
	if (!r)
		r = tdv_check(c, hipMalloc((void **)&sf, nh * sizeof(double2)),
				"hipMalloc");
	if (!r)
		r = tdv_check(c, hipMemcpyAsync(dh, hv.data(), nh * sizeof(int4),
				hipMemcpyHostToDevice, c->stream), "copy");
	if (!r) {
		hipLaunchKernelGGL(k_sfac_chunk, dim3((unsigned)nsch), dim3(256), 0,
				c->stream, c->pm, (long long)n, (long long)schunk, dh, nh,
				2.0 * EW_PI / c->par.boxsize, spart);
		hipLaunchKernelGGL(k_sfac_sum, dim3(tdv_blocks(nh, 256)), dim3(256), 0,
				c->stream, (int)nsch, nh, spart, sf);
		hipLaunchKernelGGL(k_direct_split, dim3(nt, (unsigned)nchunk),
				dim3(256), 0, c->stream, nt, idx_dev, c->pm, (int)n, (int)chunk,
				c->par.boxsize, c->par.gravity, c->Epsilon2, c->Epsilon3, part);
		hipLaunchKernelGGL(k_split_finish, dim3(tdv_blocks((size_t)nt, 64)),
				dim3(64), 0, c->stream, nt, (int)nchunk, idx_dev, c->pm, part,
				dh, sf, nh, c->par.boxsize, c->par.gravity, out_dev);
		r = tdv_check(c, hipGetLastError(), "k_direct_split");
	}
	const int rs = tdv_check(c, hipStreamSynchronize(c->stream),
			"k_direct_split");
	(void)hipFree(part);
	(void)hipFree(dh);
	(void)hipFree(spart);
	(void)hipFree(sf);
	return r ? r : rs;
}
