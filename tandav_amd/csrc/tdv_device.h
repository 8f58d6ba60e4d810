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
 * The reference runs the undo down to bit 1 on the 64-bit words; step s
 * changes only bits below s and the Gray encode combines bits only
 * downwards, so bits 0-20, which the 42-level key (bits 21-62 of each word)
 * drops, never reach the others.  The words come in shifted right by 21
 * (their bits 21-63 as bits 0-42) and the undo runs steps 42..1 on them:
 * the same key in two thirds of the steps, 31 of them on the low 32 bits
 * only. */
__host__ __device__ inline void hilbert_transpose(uint64_t &X0, uint64_t &X1,
		uint64_t &X2)
{
#pragma unroll
	for (int s = 42; s >= 1; s--) {       /* unrolled: constant masks */
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

/* bit j of the low 21 bits of x to bit 3j (the usual shift-and-mask
 * dilation of Morton codes: five steps instead of 21 single-bit moves) */
__host__ __device__ inline uint64_t dilate3_21(uint64_t x)
{
	x &= 0x1fffffull;
	x = (x | x << 32) & 0x1f00000000ffffull;
	x = (x | x << 16) & 0x1f0000ff0000ffull;
	x = (x | x << 8) & 0x100f00f00f00f00full;
	x = (x | x << 4) & 0x10c30c30c30c30c3ull;
	x = (x | x << 2) & 0x1249249249249249ull;
	return x;
}

/* The standard or reversed 128-bit key of one position (peano.c:122-276):
 * the same code in k_peano_keys and on the host (tdv_peano_keys_host).
 * After the transform, the key is the three coordinates' bits 62..21
 * interleaved, X0 the most significant of each level's triplet.  With
 * u_c = X_c >> 21 (its low 42 bits), interleaving the low and the high
 * 21 bits of the three u_c gives 63-bit words L and H, and the standard
 * key (levels from the top, at bits 127..2) is (H 2^63 + L) 4.  The
 * reversed key (level k at bits 3k..3k+2, level 42 losing its top bit) is
 * the same interleave of the bit-reversed u_c, times 8, modulo 2^128. */
__host__ __device__ inline void peano_key_words(double x, double y, double z,
		double o0, double o1, double o2, double fac, int reversed,
		uint64_t *lo_out, uint64_t *hi_out)
{
	uint64_t u0 = (uint64_t)((x - o0) * fac) >> 21;
	uint64_t u1 = (uint64_t)((y - o1) * fac) >> 21;
	uint64_t u2 = (uint64_t)((z - o2) * fac) >> 21;
	hilbert_transpose(u0, u1, u2);
	if (reversed) {                      /* bit j -> bit 41 - j */
		u0 = __builtin_bitreverse64(u0) >> 22;
		u1 = __builtin_bitreverse64(u1) >> 22;
		u2 = __builtin_bitreverse64(u2) >> 22;
	}
	const uint64_t L = dilate3_21(u0) << 2 | dilate3_21(u1) << 1
			| dilate3_21(u2);
	const uint64_t H = dilate3_21(u0 >> 21) << 2 | dilate3_21(u1 >> 21) << 1
			| dilate3_21(u2 >> 21);
	if (!reversed) {
		*lo_out = L << 2;
		*hi_out = L >> 62 | H << 1;
	} else {
		*lo_out = L << 3;
		*hi_out = L >> 61 | H << 2;
	}
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
