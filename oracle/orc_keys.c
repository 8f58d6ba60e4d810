/*
 * Peano-Hilbert keys -- restates peano.c:122-276 (Peano_Key,
 * Reversed_Peano_Key).  TEST INFRASTRUCTURE (see oracle.h).
 *
 * Both layouts start from the same 3 x 64-bit "transpose" words produced by
 * Skilling's (2004) inverse-undo + Gray-encode transform of the quantised
 * position X_d = (uint64)((x_d - origin_d) * 2^63/size).
 *   standard key: level k (1..42) triplet (X0,X1,X2 bit 63-k) at key bits
 *                 [128-3k, 130-3k]; bits 1..0 are zero (peano.c:172-194).
 *   reversed key: level k (0..42) triplet at key bits [3k, 3k+2]; level 0 is
 *                 000 and the top bit of level 42 falls off (peano.c:254-275).
 */
#include <omp.h>
#include "oracle.h"

/* Skilling's transform, peano.c:131-170. */
static void hilbert_transpose(uint64_t X[3])
{
	for (uint64_t q = (uint64_t)1 << 63; q > 1; q >>= 1) {
		const uint64_t low = q - 1;
		if (X[0] & q)
			X[0] ^= low;
		for (int d = 1; d < 3; d++) {
			if (X[d] & q) {
				X[0] ^= low;
			} else {
				uint64_t t = (X[0] ^ X[d]) & low;
				X[0] ^= t;
				X[d] ^= t;
			}
		}
	}
	X[1] ^= X[0];
	X[2] ^= X[1];
	uint64_t par = X[2];
	for (int s = 1; s < 64; s <<= 1)
		par ^= par >> s;
	uint64_t t = par ^ X[2];
	X[2] = par;
	X[1] ^= t;
	X[0] ^= t;
}

static void quantise(double px, double py, double pz, const double origin[3],
		double size, uint64_t X[3])
{
	const double fac = (double)((uint64_t)1 << 63) / size;
	X[0] = (uint64_t)((px - origin[0]) * fac);
	X[1] = (uint64_t)((py - origin[1]) * fac);
	X[2] = (uint64_t)((pz - origin[2]) * fac);
}

orc_key orc_peano_key(double px, double py, double pz,
		const double origin[3], double size)
{
	uint64_t X[3];
	quantise(px, py, pz, origin, size, X);
	hilbert_transpose(X);

	/* 43 columns MSB first; column 0 and the top bit of column 1 fall off
	 * the 128-bit word after the final shift by 2. */
	orc_key key = 0;
	for (int b = 63; b >= 21; b--) {
		unsigned col = (unsigned)(((X[0] >> b) & 1) << 2
				| ((X[1] >> b) & 1) << 1 | ((X[2] >> b) & 1));
		key = (key << 3) | col;
	}
	return key << 2;
}

orc_key orc_reversed_peano_key(double px, double py, double pz,
		const double origin[3], double size)
{
	uint64_t X[3];
	quantise(px, py, pz, origin, size, X);
	hilbert_transpose(X);

	/* columns from bit 20 (first, ends up highest) to bit 62; then one
	 * empty level-0 triplet at the bottom. */
	orc_key key = 0;
	for (int b = 20; b <= 62; b++) {
		unsigned col = (unsigned)(((X[0] >> b) & 1) << 2
				| ((X[1] >> b) & 1) << 1 | ((X[2] >> b) & 1));
		key = (key << 3) | col;
	}
	return key << 3;
}

void orc_peano_keys(size_t n, const double *x, const double *y,
		const double *z, const double origin[3], double size, orc_key *out)
{
	#pragma omp parallel for schedule(static)
	for (long i = 0; i < (long)n; i++)
		out[i] = orc_peano_key(x[i], y[i], z[i], origin, size);
}

void orc_reversed_peano_keys(size_t n, const double *x, const double *y,
		const double *z, const double origin[3], double size, orc_key *out)
{
	#pragma omp parallel for schedule(static)
	for (long i = 0; i < (long)n; i++)
		out[i] = orc_reversed_peano_key(x[i], y[i], z[i], origin, size);
}
