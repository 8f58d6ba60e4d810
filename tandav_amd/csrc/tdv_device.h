/*
 * Device helpers shared by the kernels: 128-bit key words, the Hilbert
 * transform, key-divergence levels.
 */
#ifndef TDV_DEVICE_H
#define TDV_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

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
 * the transform of Peano_Key (peano.c:131-170). */
__device__ __forceinline__ void hilbert_transpose(uint64_t &X0, uint64_t &X1,
		uint64_t &X2)
{
#pragma unroll 4
	for (int s = 63; s >= 1; s--) {
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

/* The node-size rule of create_node_from_particle (tree_build.c:546):
 * Domain.Size / (int)(1 << lvl), with the 32-bit shift of x86-64 (count
 * mod 32, 1<<31 = INT_MIN). */
__device__ __forceinline__ double int_shift_size(double size, int lvl)
{
	const int p = (int)(1u << (lvl & 31));
	return size / (double)p;
}

#endif
