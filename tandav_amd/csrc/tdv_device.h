/*
 * Device helpers shared by the kernels: 128-bit key words, the Hilbert
 * transform and the PH key of a position, key-divergence levels, the
 * Ewald-table lookup (the key and the lookup compile for the host too, for
 * the per-element calls).
 */
#ifndef TDV_DEVICE_H
#define TDV_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdv_internal.h"       /* TDV_NE, TDV_NE1, TDV_NE3 */

/* Keys live as two 64-bit words per particle: k[2i] = low, k[2i+1] = high,
 * the memory image of the reference's __uint128_t peanoKey on x86-64. */
struct Key2 {
	uint64_t lo, hi;
};

__device__ __forceinline__ Key2 load_key(const uint64_t *k, size_t i)
{
	const ulonglong2 v = reinterpret_cast<const ulonglong2 *>(k)[i];
	return Key2{ v.x, v.y };
}

__device__ __forceinline__ void store_key(uint64_t *k, size_t i, Key2 v)
{
	reinterpret_cast<ulonglong2 *>(k)[i] = make_ulonglong2(v.lo, v.hi);
}

/* First level (1..42) at which two standard keys differ; 43 if equal.
 * Level L occupies key bits [128-3L, 130-3L]. */
__device__ __forceinline__ int div_level(Key2 a, Key2 b)
{
	const uint64_t xh = a.hi ^ b.hi, xl = a.lo ^ b.lo;
	const int clz = xh ? __clzll(xh) : (xl ? 64 + __clzll(xl) : 128);
	return 1 + clz / 3;
}

/* Triplet of level L (1..42) of a standard key. */
__device__ __forceinline__ unsigned key_triplet(Key2 k, int L)
{
	const int b = 128 - 3 * L;          /* lowest bit of the triplet */
	if (b >= 64)
		return (unsigned)(k.hi >> (b - 64)) & 7u;
	if (b <= 61)
		return (unsigned)(k.lo >> b) & 7u;
	/* straddles the words (b = 62, 63) */
	const uint64_t v = (k.lo >> b) | (k.hi << (64 - b));
	return (unsigned)v & 7u;
}

/* Skilling (2004) inverse-undo + Gray encode on the three transpose words,
 * the transform of Peano_Key (peano.c:131-170), for the bits the key keeps.
 * The reference runs the undo down to bit 1; step s changes only bits below
 * s, so the steps below 22 touch only bits 0-20, which the 42-level key
 * (bits 21-62 of each word) drops, and the Gray encode combines bits only
 * downwards: stopping at s = 22 gives the same key in two thirds of the
 * steps. */
__host__ __device__ inline void hilbert_transpose(uint64_t &X0, uint64_t &X1,
		uint64_t &X2)
{
#pragma unroll
	for (int s = 63; s >= 22; s--) {      /* unrolled: constant masks */
		const uint64_t q = 1ull << s, low = q - 1;
		if (X0 & q)
			X0 ^= low;
		if (X1 & q) {
			X0 ^= low;
		} else {
			const uint64_t t = (X0 ^ X1) & low;
			X0 ^= t;
			X1 ^= t;
		}
		if (X2 & q) {
			X0 ^= low;
		} else {
			const uint64_t t = (X0 ^ X2) & low;
			X0 ^= t;
			X2 ^= t;
		}
	}
	X1 ^= X0;
	X2 ^= X1;
	uint64_t par = X2;
	par ^= par >> 1;
	par ^= par >> 2;
	par ^= par >> 4;
	par ^= par >> 8;
	par ^= par >> 16;
	par ^= par >> 32;
	const uint64_t t = par ^ X2;
	X2 = par;
	X1 ^= t;
	X0 ^= t;
}

/* The standard or reversed 128-bit key of one position (peano.c:122-276):
 * the same code in k_peano_keys and on the host (tdv_peano_keys_host). */
__host__ __device__ inline void peano_key_words(double x, double y, double z,
		double o0, double o1, double o2, double fac, int reversed,
		uint64_t *lo_out, uint64_t *hi_out)
{
	uint64_t X0 = (uint64_t)((x - o0) * fac);
	uint64_t X1 = (uint64_t)((y - o1) * fac);
	uint64_t X2 = (uint64_t)((z - o2) * fac);
	hilbert_transpose(X0, X1, X2);
	uint64_t lo = 0, hi = 0;
	if (!reversed) {
		/* column of bit b (21..62) -> key bits [3b-61, 3b-59], X0 highest */
#pragma unroll 6
		for (int b = 62; b >= 21; b--) {
			const uint64_t col = ((X0 >> b) & 1) << 2 | ((X1 >> b) & 1) << 1
					| ((X2 >> b) & 1);
			const int p = 3 * b - 61;
			if (p >= 64)
				hi |= col << (p - 64);
			else if (p <= 61)
				lo |= col << p;
			else {
				lo |= col << p;
				hi |= col >> (64 - p);
			}
		}
	} else {
		/* level k = 63-b at bits [3k, 3k+2]; level 42 loses its top bit */
#pragma unroll 6
		for (int b = 21; b <= 62; b++) {
			const uint64_t col = ((X0 >> b) & 1) << 2 | ((X1 >> b) & 1) << 1
					| ((X2 >> b) & 1);
			const int p = 3 * (63 - b);
			if (p >= 64) {
				if (p - 64 <= 63)
					hi |= col << (p - 64);
			} else if (p <= 61) {
				lo |= col << p;
			} else {
				lo |= col << p;
				hi |= col >> (64 - p);
			}
		}
	}
	*lo_out = lo;
	*hi_out = hi;
}

/* Ewald_Correction, Gravity/periodic.c:35-113 (table already / Box^2):
 * the same code on the device (the walks, tdv_ewald_correction) and on
 * the host (tdv_ewald_correction_host) */
__host__ __device__ inline void ewald_corr(const double *__restrict__ tab,
		double dx, double dy, double dz, double box2ewald, double f[3])
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
	double u = dx * box2ewald, v = dy * box2ewald, q = dz * box2ewald;
	int i = (int)u, j = (int)v, k = (int)q;
	if (i >= TDV_NE)
		i = TDV_NE - 1;
	if (j >= TDV_NE)
		j = TDV_NE - 1;
	if (k >= TDV_NE)
		k = TDV_NE - 1;
	u -= i;
	v -= j;
	q -= k;
	const double w0 = (1 - u) * (1 - v) * (1 - q), w1 = (1 - u) * (1 - v) * q,
			w2 = (1 - u) * v * (1 - q), w3 = (1 - u) * v * q,
			w4 = u * (1 - v) * (1 - q), w5 = u * (1 - v) * q,
			w6 = u * v * (1 - q), w7 = u * v * q;
	const int b = (i * TDV_NE1 + j) * TDV_NE1 + k;
	const int di = TDV_NE1 * TDV_NE1, dj = TDV_NE1;
#pragma unroll
	for (int d = 0; d < 3; d++) {
		const double *F = tab + d * TDV_NE3 + b;
		f[d] = (w0 * F[0] + w1 * F[1] + w2 * F[dj] + w3 * F[dj + 1]
				+ w4 * F[di] + w5 * F[di + 1] + w6 * F[di + dj]
				+ w7 * F[di + dj + 1]) * sg[d];
	}
}

/* The node-size rule of create_node_from_particle (tree_build.c:546):
 * Domain.Size / (int)(1 << lvl), with the 32-bit shift of x86-64 (count
 * mod 32, 1<<31 = INT_MIN). */
__device__ __forceinline__ double int_shift_size(double size, int lvl)
{
	const int p = (int)(1u << (lvl & 31));
	return size / (double)p;
}

#endif
