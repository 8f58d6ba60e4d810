/*
 * PH keys, the key sort, gathers and the device-wide scan.
 *
 * Keys: one lane per particle runs Skilling's transform (peano.c:122-195)
 * and writes the standard 128-bit key.  Reversed keys (peano.c:203-276) are
 * the triplet permutation of the standard key, produced after the sort.
 *
 * Sort: LSD radix sort, 8-bit digits, over the high 64 key bits with a
 * 32-bit payload (the original index), 8 passes of histogram -> scan ->
 * stable scatter.  Runs of equal high words are then ordered by the low
 * word; equal full keys raise TDV_ERR_DUP_KEYS (the reference's quicksort
 * is unstable, sort.c:253-341, so only unique keys have a defined order).
 * Traffic per pass: 12 B read (histogram) + 12 B read + 12 B written.
 */
#include "tdv_internal.h"
#include "tdv_device.h"

#include <cstring>
#include <mutex>

#define RS_THREADS 256
#define RS_ITEMS (TDV_SORT_TILE / RS_THREADS)
#define RS_TILE TDV_SORT_TILE
#define RS_BINS 256
#define RS_WAVES (RS_THREADS / 64)

/* ------------------------------------------------------------------------ */
/* keys                                                                      */
/* ------------------------------------------------------------------------ */

__global__ void __launch_bounds__(256) k_peano_keys(size_t n,
		const double *__restrict__ x, const double *__restrict__ y,
		const double *__restrict__ z, double o0, double o1, double o2,
		double fac, uint64_t *__restrict__ out, int reversed)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint64_t lo, hi;
	peano_key_words(x[i], y[i], z[i], o0, o1, o2, fac, reversed, &lo, &hi);
	out[2 * i] = lo;
	out[2 * i + 1] = hi;
}

void tdv_launch_keys(tdv_context *c, size_t n, const double *x,
		const double *y, const double *z, const double o[3], double size,
		uint64_t *out, int reversed)
{
	if (!n)
		return;
	const double fac = (double)(1ull << 63) / size;
	hipLaunchKernelGGL(k_peano_keys, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, x, y, z, o[0], o[1], o[2], fac, out, reversed);
}

/* reversed key from a standard key: level k (1..41) triplet moves from
 * bits [128-3k, 130-3k] to [3k, 3k+2]; level 42 keeps its two low bits at
 * 126..127; level 0 is 000 */
__global__ void k_reverse_keys(size_t n, const uint64_t *__restrict__ in,
		uint64_t *__restrict__ out)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const Key2 k = load_key(in, i);
	uint64_t lo = 0, hi = 0;
#pragma unroll 6
	for (int L = 1; L <= 42; L++) {
		const uint64_t t = key_triplet(k, L);
		const int p = 3 * L;
		if (p >= 64)
			hi |= t << (p - 64);
		else if (p <= 61)
			lo |= t << p;
		else {
			lo |= t << p;
			hi |= t >> (64 - p);
		}
	}
	store_key(out, i, Key2{ lo, hi });
}

void tdv_launch_reverse_keys(tdv_context *c, size_t n, const uint64_t *in,
		uint64_t *out)
{
	if (!n)
		return;
	hipLaunchKernelGGL(k_reverse_keys, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, in, out);
}

/* ------------------------------------------------------------------------ */
/* device-wide exclusive scan of u32 (out has n+1 entries, out[n] = total)   */
/* ------------------------------------------------------------------------ */

#define SC_THREADS 256
#define SC_ITEMS 4
#define SC_TILE (SC_THREADS * SC_ITEMS)

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *tot,
		uint32_t *sh)
{
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	uint32_t inc = v;
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		uint32_t t = __shfl_up(inc, d, 64);
		if (lane >= d)
			inc += t;
	}
	if (lane == 63)
		sh[w] = inc;
	__syncthreads();
	uint32_t wbase = 0, all = 0;
	for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
		if (q < w)
			wbase += sh[q];
		all += sh[q];
	}
	__syncthreads();
	*tot = all;
	return wbase + inc - v;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_reduce(
		const uint32_t *__restrict__ in, size_t n, uint32_t *__restrict__ part)
{
	__shared__ uint32_t sh[SC_THREADS / 64];
	const size_t base = (size_t)blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
	uint32_t s = 0;
	for (int k = 0; k < SC_ITEMS; k++)
		if (base + k < n)
			s += in[base + k];
	uint32_t tot;
	block_excl_scan(s, &tot, sh);
	if (threadIdx.x == 0)
		part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_partials(
		uint32_t *__restrict__ part, size_t np, uint32_t *__restrict__ total)
{
	__shared__ uint32_t sh[SC_THREADS / 64];
	uint32_t carry = 0;
	for (size_t b = 0; b < np; b += SC_THREADS) {
		const size_t i = b + threadIdx.x;
		const uint32_t v = i < np ? part[i] : 0;
		uint32_t tot;
		const uint32_t e = block_excl_scan(v, &tot, sh);
		if (i < np)
			part[i] = carry + e;
		carry += tot;
	}
	if (threadIdx.x == 0)
		*total = carry;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_down(
		const uint32_t *__restrict__ in, size_t n,
		const uint32_t *__restrict__ part, uint32_t *__restrict__ out)
{
	__shared__ uint32_t sh[SC_THREADS / 64];
	const size_t base = (size_t)blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
	uint32_t v[SC_ITEMS];
	uint32_t s = 0;
	for (int k = 0; k < SC_ITEMS; k++) {
		v[k] = base + k < n ? in[base + k] : 0;
		s += v[k];
	}
	uint32_t tot;
	uint32_t e = block_excl_scan(s, &tot, sh) + part[blockIdx.x];
	for (int k = 0; k < SC_ITEMS; k++) {
		if (base + k < n)
			out[base + k] = e;
		e += v[k];
	}
}

/* exclusive scan on stream s with the block partials in part
 * ((n + SC_TILE - 1) / SC_TILE + 1 entries); errors go to c (may be NULL) */
static int scan_u32(tdv_context *c, hipStream_t s, uint32_t *part,
		const uint32_t *in, uint32_t *out, size_t n)
{
	const size_t nb = (n + SC_TILE - 1) / SC_TILE;
	if (nb == 0) {
		TDV_HIP(c, hipMemsetAsync(out, 0, sizeof(uint32_t), s));
		return 0;
	}
	hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SC_THREADS), 0,
			s, in, n, part);
	hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SC_THREADS), 0, s, part,
			nb, out + n);
	hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(SC_THREADS), 0,
			s, in, n, part, out);
	return tdv_check(c, hipGetLastError(), "scan");
}

int tdv_scan_u32(tdv_context *c, const uint32_t *in, uint32_t *out, size_t n)
{
	return scan_u32(c, c->stream, c->scan_part, in, out, n);
}

/* ------------------------------------------------------------------------ */
/* LSD radix sort of u64 keys with a u32 payload                             */
/* ------------------------------------------------------------------------ */

/* key source: pass 0 may read the high words of 128-bit keys (stride 2) */
__device__ __forceinline__ uint64_t rs_load(const uint64_t *k, size_t i,
		int stride, int off)
{
	return k[i * stride + off];
}

/* The tile workgroup b of nb takes.  Workgroups go to the 8 XCDs in turn
 * (b to XCD b mod 8); numbering the tiles so that each XCD takes one
 * contiguous range of them keeps neighbouring tiles' histogram and offset
 * entries (4 bytes apart, digit-major) and the digit runs they write side
 * by side in one XCD's L2, where they had been partial lines of eight
 * L2s, each written back or fetched on its own. */
__device__ __forceinline__ unsigned rs_tile(unsigned b, unsigned nb)
{
	const unsigned q = nb / 8, r = nb % 8, x = b % 8, i = b / 8;
	return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(
		const uint64_t *__restrict__ keys, int stride, int off, size_t n,
		int shift, uint32_t *__restrict__ hist, unsigned nblocks)
{
	__shared__ uint32_t h[RS_BINS];
	h[threadIdx.x] = 0;
	__syncthreads();
	const unsigned tile = rs_tile(blockIdx.x, nblocks);
	const size_t base = (size_t)tile * RS_TILE;
	/* every load issued before the first is used: unconditional loads of
	 * an index clamped into [0, n) (n >= 1), the count only when i < n */
	uint64_t kk[RS_ITEMS];
#pragma unroll
	for (int it = 0; it < RS_ITEMS; it++) {
		const size_t i = base + (size_t)it * RS_THREADS + threadIdx.x;
		kk[it] = rs_load(keys, i < n ? i : n - 1, stride, off);
	}
#pragma unroll
	for (int it = 0; it < RS_ITEMS; it++) {
		const size_t i = base + (size_t)it * RS_THREADS + threadIdx.x;
		if (i < n)
			atomicAdd(&h[(unsigned)(kk[it] >> shift) & 255u], 1u);
	}
	__syncthreads();
	hist[(size_t)threadIdx.x * nblocks + tile] = h[threadIdx.x];
}

/* One LSD pass's scatter of a tile of RS_TILE keys: every key's rank
 * among the tile's keys of its digit (stable: waves, items and lanes in
 * input order), the tile sorted by digit in LDS, then written out in that
 * order, so consecutive threads store consecutive positions of a digit's
 * run (one run of ~16 keys per digit and tile) instead of every lane a
 * key of its own digit somewhere else. */
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(
		const uint64_t *__restrict__ kin, int stride, int off,
		const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
		uint32_t *__restrict__ vout, size_t n, int shift,
		const uint32_t *__restrict__ hscan, unsigned nblocks)
{
	__shared__ uint32_t wcnt[RS_WAVES][RS_BINS];
	__shared__ uint32_t binoff[RS_BINS];
	__shared__ uint32_t dstart[RS_BINS];     /* digit's first slot in tile */
	__shared__ uint32_t wsum[RS_WAVES];
	__shared__ uint64_t skey[RS_TILE];
	__shared__ uint32_t sval[RS_TILE];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	for (int q = 0; q < RS_WAVES; q++)
		wcnt[q][threadIdx.x] = 0;
	const unsigned tile = rs_tile(blockIdx.x, nblocks);
	binoff[threadIdx.x] = hscan[(size_t)threadIdx.x * nblocks + tile];
	__syncthreads();

	const size_t base = (size_t)tile * RS_TILE + (size_t)w * 64 * RS_ITEMS;
	uint64_t kk[RS_ITEMS];
	uint32_t vv[RS_ITEMS];
	uint32_t rk[RS_ITEMS];
	const uint64_t lt = (1ull << lane) - 1;
	/* all of the wave's loads in flight at once: unconditional, at an
	 * index clamped into [0, n) (n >= 1); what lies past n is not used */
#pragma unroll
	for (int it = 0; it < RS_ITEMS; it++) {
		const size_t i = base + (size_t)it * 64 + lane;
		kk[it] = rs_load(kin, i < n ? i : n - 1, stride, off);
	}
	if (vin) {
#pragma unroll
		for (int it = 0; it < RS_ITEMS; it++) {
			const size_t i = base + (size_t)it * 64 + lane;
			vv[it] = vin[i < n ? i : n - 1];
		}
	} else {
#pragma unroll
		for (int it = 0; it < RS_ITEMS; it++)
			vv[it] = (uint32_t)(base + (size_t)it * 64 + lane);
	}
#pragma unroll
	for (int it = 0; it < RS_ITEMS; it++) {
		const size_t i = base + (size_t)it * 64 + lane;
		const bool valid = i < n;
		const unsigned d = (unsigned)(kk[it] >> shift) & 255u;
		uint64_t peers = __ballot(valid);
#pragma unroll
		for (int b = 0; b < 8; b++) {
			const uint64_t bal = __ballot((d >> b) & 1u);
			peers &= ((d >> b) & 1u) ? bal : ~bal;
		}
		const uint32_t below = (uint32_t)__popcll(peers & lt);
		const int leader = peers ? (__ffsll((unsigned long long)peers) - 1) : 0;
		uint32_t old = 0;
		if (valid && below == 0)
			old = atomicAdd(&wcnt[w][d], (uint32_t)__popcll(peers));
		old = __shfl(old, leader, 64);
		rk[it] = old + below;
	}
	__syncthreads();
	/* per digit: the waves' exclusive prefix, and the tile's count, whose
	 * exclusive scan over the digits places each digit's run in the tile */
	uint32_t cnt = 0;
	for (int q = 0; q < RS_WAVES; q++) {
		const uint32_t t = wcnt[q][threadIdx.x];
		wcnt[q][threadIdx.x] = cnt;
		cnt += t;
	}
	uint32_t x = cnt;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t y = __shfl_up(x, o, 64);
		if (lane >= o)
			x += y;
	}
	if (lane == 63)
		wsum[w] = x;
	__syncthreads();
	uint32_t before = 0;
	for (int q = 0; q < w; q++)
		before += wsum[q];
	dstart[threadIdx.x] = before + x - cnt;
	__syncthreads();
#pragma unroll
	for (int it = 0; it < RS_ITEMS; it++) {
		const size_t i = base + (size_t)it * 64 + lane;
		if (i < n) {
			const unsigned d = (unsigned)(kk[it] >> shift) & 255u;
			const uint32_t slot = dstart[d] + wcnt[w][d] + rk[it];
			skey[slot] = kk[it];
			sval[slot] = vv[it];
		}
	}
	__syncthreads();
	const size_t tile0 = (size_t)tile * RS_TILE;
	const int nvalid = n - tile0 < (size_t)RS_TILE ? (int)(n - tile0)
			: RS_TILE;
	for (int j = threadIdx.x; j < nvalid; j += RS_THREADS) {
		const uint64_t k = skey[j];
		const unsigned d = (unsigned)(k >> shift) & 255u;
		const uint32_t pos = binoff[d] + ((uint32_t)j - dstart[d]);
		if (pos < n) {                       /* always, by construction */
			kout[pos] = k;
			vout[pos] = sval[j];
		}
	}
}

/* the buffers a key sort works in: a context's own, or the no-allocation
 * entry point's static scratch */
struct SortWork {
	hipStream_t stream;
	uint64_t *sk[2];             /* n keys each */
	uint32_t *sv[2];             /* n payloads each */
	uint32_t *hist;              /* hist_cap digit counts */
	size_t hist_cap;
	uint32_t *scan_part;         /* scan block partials */
	unsigned *d_err;             /* TDV_EBIT_* flags */
	uint64_t *hi_tmp;            /* n words: the tie fallback's high words */
	uint32_t *p_tmp;             /* n: the tie fallback's first order */
	bool no_sync;                /* captured: no read back, no fallback (the
	                                tie flag stays in d_err for the caller) */
};

static SortWork context_work(tdv_context *c, const uint32_t *perm_out)
{
	SortWork w;
	w.stream = c->stream;
	for (int b = 0; b < 2; b++) {
		w.sk[b] = c->sk[b];
		w.sv[b] = c->sv[b];
	}
	w.hist = c->hist;
	w.hist_cap = c->hist_cap;
	w.scan_part = c->scan_part;
	w.d_err = c->d_err;
	w.hi_tmp = reinterpret_cast<uint64_t *>(c->tmp_f64);
	w.p_tmp = c->perm == perm_out ? reinterpret_cast<uint32_t *>(c->lohi)
			: c->perm;
	w.no_sync = c->graph_mode != 0;
	return w;
}

/* sorts n u64 keys (read with stride/offset, e.g. the high words of 128-bit
 * keys) carrying payload vin (NULL = original index) into w.sk/sv; returns
 * the buffer index holding the result */
static int radix_sort_u64(tdv_context *c, const SortWork &w, size_t n,
		const uint64_t *keys, int stride, int off, const uint32_t *vin,
		int *res)
{
	const unsigned nb = (unsigned)((n + RS_TILE - 1) / RS_TILE);
	const size_t hn = (size_t)nb * RS_BINS;
	if (hn + 1 > w.hist_cap)
		return tdv_fail(c, TDV_ERR_ARG, "radix sort histogram too small");
	const uint64_t *kin = keys;
	const uint32_t *vi = vin;
	int s = stride, o = off, cur = 0;
	for (int pass = 0; pass < 8; pass++) {
		const int shift = 8 * pass;
		hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_THREADS), 0, w.stream,
				kin, s, o, n, shift, w.hist, nb);
		TDV_TRY(scan_u32(c, w.stream, w.scan_part, w.hist, w.hist, hn));
		hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_THREADS), 0,
				w.stream, kin, s, o, vi, w.sk[cur], w.sv[cur], n, shift,
				w.hist, nb);
		kin = w.sk[cur];
		vi = w.sv[cur];
		s = 1;
		o = 0;
		cur ^= 1;
	}
	*res = cur ^ 1;
	return tdv_check(c, hipGetLastError(), "radix sort");
}

/* One stable pass over the low 8 bits of n u64 keys with the keys' indices
 * as payload: work items in key order (the FP32 walk's groups by expected
 * cost), in index order within a key */
int tdv_radix_pass8(tdv_context *c, size_t n, const uint64_t *keys,
		uint64_t *kout, uint32_t *order)
{
	const unsigned nb = (unsigned)((n + RS_TILE - 1) / RS_TILE);
	const size_t hn = (size_t)nb * RS_BINS;
	if (hn + 1 > c->hist_cap)
		return tdv_fail(c, TDV_ERR_ARG, "radix pass histogram too small");
	hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_THREADS), 0, c->stream,
			keys, 1, 0, n, 0, c->hist, nb);
	TDV_TRY(scan_u32(c, c->stream, c->scan_part, c->hist, c->hist, hn));
	hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_THREADS), 0,
			c->stream, keys, 1, 0, nullptr, kout, order, n, 0, c->hist, nb);
	return tdv_check(c, hipGetLastError(), "radix pass");
}

/* order runs of equal high words by the low word; flag duplicate keys */
__global__ void k_fix_ties(size_t n, const uint64_t *__restrict__ shi,
		uint32_t *__restrict__ perm, const uint64_t *__restrict__ keys,
		unsigned *err)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const uint64_t h = shi[i];
	if (i > 0 && shi[i - 1] == h)
		return;                          /* not a run start */
	if (i + 1 >= n || shi[i + 1] != h)
		return;                          /* run of one */
	size_t e = i + 1;
	while (e + 1 < n && shi[e + 1] == h)
		e++;
	if (e - i + 1 > 4096) {
		atomicOr(err, TDV_EBIT_TIES);
		return;
	}
	for (size_t r = i + 1; r <= e; r++) {   /* insertion sort by low word */
		const uint32_t pv = perm[r];
		const uint64_t lv = keys[2 * (size_t)pv];
		size_t t = r;
		while (t > i && keys[2 * (size_t)perm[t - 1]] > lv) {
			perm[t] = perm[t - 1];
			t--;
		}
		perm[t] = pv;
	}
	for (size_t r = i + 1; r <= e; r++)
		if (keys[2 * (size_t)perm[r]] == keys[2 * (size_t)perm[r - 1]])
			atomicOr(err, TDV_EBIT_DUP);
}

/* gather the high words of 128-bit keys by a permutation */
__global__ void k_gather_hi(size_t n, const uint32_t *__restrict__ perm,
		const uint64_t *__restrict__ keys, uint64_t *__restrict__ out)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = keys[2 * (size_t)perm[i] + 1];
}

__global__ void k_check_dups(size_t n, const uint32_t *__restrict__ perm,
		const uint64_t *__restrict__ keys, unsigned *err)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0 || i >= n)
		return;
	const size_t a = perm[i - 1], b = perm[i];
	if (keys[2 * a] == keys[2 * b] && keys[2 * a + 1] == keys[2 * b + 1])
		atomicOr(err, TDV_EBIT_DUP);
}

/* Qsort_Index's order of unique 128-bit keys (2n words) into perm_out:
 * the high words radix-sorted, runs of equal high words ordered by the low
 * word; runs longer than 4096 switch to a full 128-bit LSD sort.  Duplicate
 * keys set TDV_EBIT_DUP in w.d_err.  Reads w.d_err back (a stream
 * synchronisation) to choose the fallback. */
static int sort_keys_work(tdv_context *c, const SortWork &w, size_t n,
		const uint64_t *keys, uint32_t *perm_out)
{
	if (n == 0)
		return 0;
	int res;
	TDV_TRY(radix_sort_u64(c, w, n, keys, 2, 1, nullptr, &res));
	hipLaunchKernelGGL(k_fix_ties, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			w.stream, n, w.sk[res], w.sv[res], keys, w.d_err);
	if (w.no_sync) {
		TDV_HIP(c, hipMemcpyAsync(perm_out, w.sv[res], n * sizeof(uint32_t),
				hipMemcpyDeviceToDevice, w.stream));
		return tdv_check(c, hipGetLastError(), "sort");
	}
	unsigned e = 0;
	TDV_HIP(c, hipMemcpyAsync(&e, w.d_err, sizeof(e), hipMemcpyDeviceToHost,
			w.stream));
	TDV_HIP(c, hipStreamSynchronize(w.stream));
	if (e & TDV_EBIT_TIES) {
		/* long runs of equal high words: full 128-bit LSD sort, low word
		 * first (8 passes), then the high words in that order (8 passes) */
		TDV_HIP(c, hipMemsetAsync(w.d_err, 0, sizeof(unsigned), w.stream));
		int r1;
		TDV_TRY(radix_sort_u64(c, w, n, keys, 2, 0, nullptr, &r1));
		TDV_HIP(c, hipMemcpyAsync(w.p_tmp, w.sv[r1], n * sizeof(uint32_t),
				hipMemcpyDeviceToDevice, w.stream));
		hipLaunchKernelGGL(k_gather_hi, dim3(tdv_blocks(n, 256)), dim3(256), 0,
				w.stream, n, w.p_tmp, keys, w.hi_tmp);
		TDV_TRY(radix_sort_u64(c, w, n, w.hi_tmp, 1, 0, w.p_tmp, &res));
		hipLaunchKernelGGL(k_check_dups, dim3(tdv_blocks(n, 256)), dim3(256), 0,
				w.stream, n, w.sv[res], keys, w.d_err);
	}
	TDV_HIP(c, hipMemcpyAsync(perm_out, w.sv[res], n * sizeof(uint32_t),
			hipMemcpyDeviceToDevice, w.stream));
	return tdv_check(c, hipGetLastError(), "sort");
}

int tdv_sort_device(tdv_context *c, size_t n, const uint64_t *keys,
		uint32_t *perm_out)
{
	return sort_keys_work(c, context_work(c, perm_out), n, keys, perm_out);
}

/* ------------------------------------------------------------------------ */
/* the key sort without allocation: a static scratch array per device      */
/* ------------------------------------------------------------------------ */

/* byte offsets of the work buffers in the scratch array, for
 * TDV_NOALLOC_MAX_KEYS keys (each 256-byte aligned) */
#define NA_N ((size_t)TDV_NOALLOC_MAX_KEYS)
#define NA_ALIGN(x) (((x) + 255) & ~(size_t)255)
#define NA_HIST_CAP (((NA_N + RS_TILE - 1) / RS_TILE) * RS_BINS + 1)
#define NA_PART_CAP (NA_HIST_CAP / SC_TILE + 2)
#define NA_SK0 ((size_t)0)
#define NA_SK1 (NA_SK0 + NA_ALIGN(NA_N * 8))
#define NA_SV0 (NA_SK1 + NA_ALIGN(NA_N * 8))
#define NA_SV1 (NA_SV0 + NA_ALIGN(NA_N * 4))
#define NA_HI (NA_SV1 + NA_ALIGN(NA_N * 4))
#define NA_PT (NA_HI + NA_ALIGN(NA_N * 8))
#define NA_HIST (NA_PT + NA_ALIGN(NA_N * 4))
#define NA_PART (NA_HIST + NA_ALIGN(NA_HIST_CAP * 4))
#define NA_ERR (NA_PART + NA_ALIGN(NA_PART_CAP * 4))
#define NA_BYTES (NA_ERR + 256)

/* one copy per GPU, held until the process ends; every call on a GPU uses
 * that GPU's copy, so calls that overlap on one GPU race */
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

#define NA_MAX_DEVICES 64
static std::mutex na_mutex;
static unsigned char *na_ptr[NA_MAX_DEVICES];

/* the device address of this GPU's scratch (host code never passes the
 * array's own name: on the host it is not the device address) */
static int noalloc_scratch(int dev, unsigned char **out)
{
	std::lock_guard<std::mutex> lock(na_mutex);
	if (!na_ptr[dev]) {
		void *p = nullptr;
		if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess
				|| !p) {
			(void)hipGetLastError();
			return TDV_ERR_HIP;
		}
		na_ptr[dev] = static_cast<unsigned char *>(p);
	}
	*out = na_ptr[dev];
	return 0;
}

extern "C" int tdv_sort_keys_noalloc(size_t n, const uint64_t *keys,
		uint32_t *perm_out, void *stream)
{
	if (n > NA_N)
		return TDV_ERR_ARG;
	if (n == 0)
		return 0;
	if (!keys || !perm_out)
		return TDV_ERR_ARG;
	int dev = -1;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) {
		(void)hipGetLastError();
		return TDV_ERR_HIP;
	}
	unsigned char *base = nullptr;
	TDV_TRY(noalloc_scratch(dev, &base));
	SortWork w;
	w.stream = static_cast<hipStream_t>(stream);
	w.sk[0] = reinterpret_cast<uint64_t *>(base + NA_SK0);
	w.sk[1] = reinterpret_cast<uint64_t *>(base + NA_SK1);
	w.sv[0] = reinterpret_cast<uint32_t *>(base + NA_SV0);
	w.sv[1] = reinterpret_cast<uint32_t *>(base + NA_SV1);
	w.hi_tmp = reinterpret_cast<uint64_t *>(base + NA_HI);
	w.p_tmp = reinterpret_cast<uint32_t *>(base + NA_PT);
	w.hist = reinterpret_cast<uint32_t *>(base + NA_HIST);
	w.hist_cap = NA_HIST_CAP;
	w.scan_part = reinterpret_cast<uint32_t *>(base + NA_PART);
	w.d_err = reinterpret_cast<unsigned *>(base + NA_ERR);
	w.no_sync = false;
	/* the previous call's flags stay in the array: clear them */
	if (hipMemsetAsync(w.d_err, 0, sizeof(unsigned), w.stream) != hipSuccess) {
		(void)hipGetLastError();
		return TDV_ERR_HIP;
	}
	TDV_TRY(sort_keys_work(nullptr, w, n, keys, perm_out));
	unsigned e = 0;
	if (hipMemcpyAsync(&e, w.d_err, sizeof(e), hipMemcpyDeviceToHost, w.stream)
			!= hipSuccess || hipStreamSynchronize(w.stream) != hipSuccess) {
		(void)hipGetLastError();
		return TDV_ERR_HIP;
	}
	if (e & TDV_EBIT_DUP)
		return TDV_ERR_DUP_KEYS;
	if (e & TDV_EBIT_TIES)
		return TDV_ERR_TIES;
	return 0;
}

/* ------------------------------------------------------------------------ */
/* gathers                                                                   */
/* ------------------------------------------------------------------------ */

template <typename T>
__global__ void k_gather(size_t n, const uint32_t *__restrict__ perm,
		const T *__restrict__ in, T *__restrict__ out)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = in[perm[i]];
}

void tdv_launch_gather_f64(tdv_context *c, size_t n, const uint32_t *perm,
		const double *in, double *out)
{
	hipLaunchKernelGGL(k_gather<double>, dim3(tdv_blocks(n, 256)), dim3(256),
			0, c->stream, n, perm, in, out);
}

void tdv_launch_gather_i32(tdv_context *c, size_t n, const uint32_t *perm,
		const int *in, int *out)
{
	hipLaunchKernelGGL(k_gather<int>, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, perm, in, out);
}

void tdv_launch_gather_key(tdv_context *c, size_t n, const uint32_t *perm,
		const uint64_t *in, uint64_t *out)
{
	hipLaunchKernelGGL(k_gather<ulonglong2>, dim3(tdv_blocks(n, 256)),
			dim3(256), 0, c->stream, n, perm,
			reinterpret_cast<const ulonglong2 *>(in),
			reinterpret_cast<ulonglong2 *>(out));
}

__global__ void k_pack_pm(size_t n, const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z,
		const double *__restrict__ m, double4 *__restrict__ pm)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		pm[i] = make_double4(x[i], y[i], z[i], m[i]);
}

void tdv_launch_pack_pm(tdv_context *c)
{
	if (!c->n)
		return;
	hipLaunchKernelGGL(k_pack_pm, dim3(tdv_blocks(c->n, 256)), dim3(256), 0,
			c->stream, c->n, c->pos[0], c->pos[1], c->pos[2], c->mass, c->pm);
}

/* ------------------------------------------------------------------------ */
/* non-periodic domain (domain.c:751-847, select.c:19-80)                    */
/* ------------------------------------------------------------------------ */

/* order-preserving u64 image of a double */
__global__ void k_f64_orderkey(size_t n, const double *__restrict__ x,
		uint64_t *__restrict__ out)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const uint64_t u = (uint64_t)__double_as_longlong(x[i]);
	out[i] = (u >> 63) ? ~u : (u | (1ull << 63));
}

__global__ void k_max_dist(size_t n, const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z, double cx,
		double cy, double cz, unsigned long long *out)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	double v = 0;
	if (i < n)
		v = fmax(fmax(fabs(x[i] - cx), fabs(y[i] - cy)), fabs(z[i] - cz));
	/* non-negative doubles order like their bit patterns */
	unsigned long long b = (unsigned long long)__double_as_longlong(v);
	for (int d = 32; d >= 1; d >>= 1) {
		unsigned long long o = __shfl_xor(b, d, 64);
		b = o > b ? o : b;
	}
	if ((threadIdx.x & 63) == 0)
		atomicMax(out, b);
}

int tdv_domain_nonperiodic(tdv_context *c)
{
	const size_t n = c->n;
	if (n < 3)
		return tdv_fail(c, TDV_ERR_ARG, "non-periodic domain needs 3 particles");
	/* Select(n>>1, n, data) orders data[1..n-1] and returns position n>>1 of
	 * that 1-based range, i.e. the ((n>>1)-1)-th smallest of data[1..n-1] */
	const size_t m = n - 1, k = (n >> 1) - 1;
	uint64_t *tmp = c->key_tmp;
	for (int d = 0; d < 3; d++) {
		hipLaunchKernelGGL(k_f64_orderkey, dim3(tdv_blocks(m, 256)), dim3(256),
				0, c->stream, m, c->pos[d] + 1, tmp);
		int res;
		TDV_TRY(radix_sort_u64(c, context_work(c, nullptr), m, tmp, 1, 0,
				nullptr, &res));
		uint64_t u;
		TDV_HIP(c, hipMemcpyAsync(&u, c->sk[res] + k, 8, hipMemcpyDeviceToHost,
				c->stream));
		TDV_HIP(c, hipStreamSynchronize(c->stream));
		u = (u >> 63) ? (u & ~(1ull << 63)) : ~u;
		double v;
		memcpy(&v, &u, 8);
		c->dom_center[d] = v;
	}
	unsigned long long *dmax = reinterpret_cast<unsigned long long *>(tmp);
	TDV_HIP(c, hipMemsetAsync(dmax, 0, 8, c->stream));
	hipLaunchKernelGGL(k_max_dist, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, c->pos[0], c->pos[1], c->pos[2], c->dom_center[0],
			c->dom_center[1], c->dom_center[2], dmax);
	unsigned long long b;
	TDV_HIP(c, hipMemcpyAsync(&b, dmax, 8, hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	double maxd;
	memcpy(&maxd, &b, 8);
	c->dom_size = 2.001 * maxd;
	for (int d = 0; d < 3; d++)
		c->dom_origin[d] = c->dom_center[d] - 0.5 * c->dom_size;
	return 0;
}
