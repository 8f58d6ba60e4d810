/*
 * Ewald correction tables -- restates compute_ewald_correction_table,
 * compute_ewald_force and compute_ewald_potential (Gravity/periodic.c:
 * 235-300, 352-461): Hernquist, Bouchet & Suto (1991) sums with alpha = 2
 * and |n|,|h| <= 4 per component, evaluated at r = (i,j,k)/(2*64) box units
 * on the 65^3 octant grid.  Output is unnormalised (box units), layout
 * [i][j][k].  TEST INFRASTRUCTURE (see oracle.h).
 */
#include <math.h>
#include <omp.h>
#include "oracle.h"

#define N_EWALD 64
#define NE1 (N_EWALD + 1)
#define PI 3.14159265358979323846     /* constants.h:15 */

static const double Alpha = 2.0;

static void ewald_force(int i, int j, int k, const double x[3], double f[3])
{
	f[0] = f[1] = f[2] = 0;
	if (i == 0 && j == 0 && k == 0)
		return;
	double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
	for (int d = 0; d < 3; d++)
		f[d] = x[d] / (r * r * r);
	for (int a = -4; a < 5; a++)
		for (int b = -4; b < 5; b++)
			for (int c = -4; c < 5; c++) {
				double dx[3] = { x[0] - a, x[1] - b, x[2] - c };
				double rr = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
				double val = erfc(Alpha * rr) + 2 * Alpha * rr / sqrt(PI)
						* exp(-((Alpha * rr) * (Alpha * rr)));
				for (int d = 0; d < 3; d++)
					f[d] -= dx[d] / (rr * rr * rr) * val;
			}
	for (int a = -4; a < 5; a++)
		for (int b = -4; b < 5; b++)
			for (int c = -4; c < 5; c++) {
				double hdotx = x[0] * a + x[1] * b + x[2] * c;
				int h2 = a * a + b * b + c * c;
				if (h2 > 0) {
					double val = 2.0 / h2 * exp(-PI * PI * h2 / (Alpha * Alpha))
							* sin(2 * PI * hdotx);
					f[0] -= a * val;
					f[1] -= b * val;
					f[2] -= c * val;
				}
			}
}

static double ewald_potential(const double r[3])
{
	if (r[0] == 0 && r[1] == 0 && r[2] == 0)
		return 2.8372975;             /* Gravity/periodic.c:419 */
	double s1 = 0;
	for (int a = -4; a < 5; a++)
		for (int b = -4; b < 5; b++)
			for (int c = -4; c < 5; c++) {
				double dx[3] = { r[0] - a, r[1] - b, r[2] - c };
				double dr = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
				s1 += erfc(Alpha * dr) / dr;
			}
	double s2 = 0;
	for (int a = -4; a < 5; a++)
		for (int b = -4; b < 5; b++)
			for (int c = -4; c < 5; c++) {
				double hdotr = r[0] * a + r[1] * b + r[2] * c;
				int h2 = a * a + b * b + c * c;
				if (h2 > 0)
					s2 += 1.0 / (PI * h2) * exp(-PI * PI * h2 / (Alpha * Alpha))
						* cos(2 * PI * hdotr);
			}
	return PI / (Alpha * Alpha) - s1 - s2
		+ 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

void orc_ewald_table(double *fx, double *fy, double *fz, double *fp,
		int nthreads)
{
	if (nthreads < 1)
		nthreads = 1;
	#pragma omp parallel for collapse(2) schedule(dynamic) num_threads(nthreads)
	for (int i = 0; i < NE1; i++)
		for (int j = 0; j < NE1; j++)
			for (int k = 0; k < NE1; k++) {
				int n = (i * NE1 + j) * NE1 + k;
				double r[3] = { (0.5 * i) / N_EWALD, (0.5 * j) / N_EWALD,
					(0.5 * k) / N_EWALD };
				double f[3];
				ewald_force(i, j, k, r, f);
				fx[n] = f[0];
				fy[n] = f[1];
				fz[n] = f[2];
				fp[n] = ewald_potential(r);
			}
}
