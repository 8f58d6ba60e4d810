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
