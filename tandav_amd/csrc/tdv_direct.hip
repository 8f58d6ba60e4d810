/*
 * FP64 direct summation for sampled target particles -- the engine's form of
 * Gravity_Forcetest (Gravity/forcetest.c:24-174): for each target, the sum
 * over every other particle of the path's own pairwise kernel (K1 with the
 * negative softening length, Gravity/tree_accel.c:362-394), plus, on
 * periodic runs, the tabulated Ewald correction of every pair (nearest image
 * + Ewald_Correction, Gravity/periodic.c:35-113).  One workgroup per target;
 * lanes stride over the sources and the partial sums meet in a block
 * reduction.  Used to state the acceleration error of the walks.
 */
#include "tdv_internal.h"

__device__ __forceinline__ double dwrap(double d, int periodic, double box,
		double half)
{
	if (!periodic)
		return d;
	if (d > half)
		d -= box;
	else if (d < -half)
		d += box;
	return d;
}

__device__ __forceinline__ void dewald(const double *__restrict__ tab,
		double dx, double dy, double dz, double b2e, double f[3])
{
	double sg[3] = { -1, -1, -1 };
	if (dx < 0) {
		dx = -dx;
		sg[0] = 1;
	}
	if (dy < 0) {
		dy = -dy;
		sg[1] = 1;
	}
	if (dz < 0) {
		dz = -dz;
		sg[2] = 1;
	}
	double u = dx * b2e, v = dy * b2e, q = dz * b2e;
	int i = (int)u, j = (int)v, k = (int)q;
	i = i >= TDV_NE ? TDV_NE - 1 : i;
	j = j >= TDV_NE ? TDV_NE - 1 : j;
	k = k >= TDV_NE ? TDV_NE - 1 : k;
	u -= i;
	v -= j;
	q -= k;
	const double w0 = (1 - u) * (1 - v) * (1 - q), w1 = (1 - u) * (1 - v) * q,
			w2 = (1 - u) * v * (1 - q), w3 = (1 - u) * v * q,
			w4 = u * (1 - v) * (1 - q), w5 = u * (1 - v) * q,
			w6 = u * v * (1 - q), w7 = u * v * q;
	const int b = (i * TDV_NE1 + j) * TDV_NE1 + k;
	const int di = TDV_NE1 * TDV_NE1, dj = TDV_NE1;
	for (int d = 0; d < 3; d++) {
		const double *F = tab + d * TDV_NE3 + b;
		f[d] = (w0 * F[0] + w1 * F[1] + w2 * F[dj] + w3 * F[dj + 1]
				+ w4 * F[di] + w5 * F[di + 1] + w6 * F[di + dj]
				+ w7 * F[di + dj + 1]) * sg[d];
	}
}

__global__ void __launch_bounds__(256) k_direct(int nt,
		const int *__restrict__ idx, const double4 *__restrict__ pm, int n,
		int periodic, double box, double G, double eps2, double eps3,
		const double *__restrict__ tab, double b2e, double *__restrict__ out)
{
	__shared__ double red[3][256];
	const int t = blockIdx.x;
	if (t >= nt)
		return;
	const int i = idx[t];
	const double4 p = pm[i];
	const double half = box / 2.0;
	double ax = 0, ay = 0, az = 0;
	for (int j = threadIdx.x; j < n; j += blockDim.x) {
		if (j == i)
			continue;
		const double4 q = pm[j];
		const double dx = dwrap(q.x - p.x, periodic, box, half);
		const double dy = dwrap(q.y - p.y, periodic, box, half);
		const double dz = dwrap(q.z - p.z, periodic, box, half);
		const double r2 = dx * dx + dy * dy + dz * dz;
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
		ax += fac * dx;
		ay += fac * dy;
		az += fac * dz;
		if (periodic) {
			double f[3];
			dewald(tab, dx, dy, dz, b2e, f);
			ax += G * q.w * f[0];
			ay += G * q.w * f[1];
			az += G * q.w * f[2];
		}
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
			out[3 * t + d] = red[d][0];
}

int tdv_direct_sum(tdv_context *c, int nt, const int *idx_dev, double *out_dev)
{
	if (nt <= 0)
		return 0;
	if (c->par.periodic && !c->have_ewald)
		return tdv_fail(c, TDV_ERR_ARG, "direct sum without an Ewald table");
	hipLaunchKernelGGL(k_direct, dim3(nt), dim3(256), 0, c->stream, nt,
			idx_dev, c->pm, (int)c->n, c->par.periodic, c->par.boxsize,
			c->par.gravity, c->Epsilon2, c->Epsilon3, c->ewald,
			2 * TDV_NE / (c->par.periodic ? c->par.boxsize : 1.0), out_dev);
	return tdv_check(c, hipGetLastError(), "k_direct");
}
