/*
 * Index sort of 128-bit keys and the permutation gathers.
 * TEST INFRASTRUCTURE (see oracle.h).
 *
 * orc_qsort_index restates Qsort_Index / omp_qsort_index (sort.c:90-116,
 * 214-341): an unstable quicksort of a permutation array with insertion sort
 * below 32 elements, median-of-3 pivots below 64 and median-of-9 above,
 * OpenMP tasks above a spawn threshold.  For unique keys the result is the
 * unique ascending order; the restatement exists so that the CPU baseline
 * pays the same comparison-sort costs as the reference.
 *
 * orc_reorder_* restate Reorder_Array_8/4/Char (aux.c:187-300): the in-place
 * cycle-following gather p_new[i] = p_old[idx[i]].
 */
#include <string.h>
#include <omp.h>
#include "oracle.h"

#define INSERTION_THRESHOLD 32
#define MEDIAN_THRESHOLD 64
#define PARALLEL_THRESHOLD 1024
#define N_PARTITIONS_PER_CPU 32

static inline int key_cmp(const orc_key *k, size_t a, size_t b)
{
	return (k[a] > k[b]) - (k[a] < k[b]);   /* cmp_peanoKeys, peano.c:6-12 */
}

static inline void swap_idx(size_t *a, size_t *b)
{
	size_t t = *a;
	*a = *b;
	*b = t;
}

static size_t *median9(size_t *lo, const orc_key *k, size_t n)
{
	size_t dp = n / 9;
	size_t *a[9] = { lo, lo + dp, lo + 2 * dp, lo + 3 * dp, lo + 4 * dp,
		lo + 5 * dp, lo + 6 * dp, lo + 7 * dp, lo + n - 1 };
	for (int i = 1; i < 9; i++)
		for (int j = i; j > 0 && key_cmp(k, *a[j - 1], *a[j]) > 0; j--)
			swap_idx(a[j], a[j - 1]);
	return a[5];
}

static size_t *median3(size_t *lo, const orc_key *k, size_t n)
{
	size_t *hi = lo + n - 1;
	size_t *mid = lo + (n >> 1);
	if (key_cmp(k, *mid, *lo) < 0)
		swap_idx(mid, lo);
	if (key_cmp(k, *hi, *mid) < 0) {
		swap_idx(hi, mid);
		if (key_cmp(k, *mid, *lo) < 0)
			swap_idx(mid, lo);
	}
	return mid;
}

static void qsort_index_rec(size_t *perm, const orc_key *k, size_t n,
		size_t spawn)
{
	if (n < INSERTION_THRESHOLD) {
		for (size_t r = 1; r < n; r++)
			for (size_t t = r; t > 0 && key_cmp(k, perm[t - 1], perm[t]) > 0; t--)
				swap_idx(&perm[t], &perm[t - 1]);
		return;
	}
	size_t *lo = perm, *hi = perm + n - 1;
	size_t *mid = (n < MEDIAN_THRESHOLD) ? median3(perm, k, n)
			: median9(perm, k, n);
	size_t *left = lo + 1, *right = hi - 1;

	do {
		while (key_cmp(k, *left, *mid) < 0)
			left++;
		while (key_cmp(k, *mid, *right) < 0)
			right--;
		if (left < right) {
			swap_idx(left, right);
			if (mid == left)
				mid = right;
			else if (mid == right)
				mid = left;
			left++;
			right--;
		} else if (left == right) {
			left++;
			right--;
			break;
		}
	} while (left <= right);

	size_t nl = (size_t)(right - lo + 1), nr = (size_t)(hi - left + 1);
	if (nl > 1) {
		if (nl > spawn) {
			#pragma omp task
			qsort_index_rec(lo, k, nl, spawn);
		} else {
			qsort_index_rec(lo, k, nl, spawn);
		}
	}
	if (nr > 1) {
		if (nr > spawn) {
			#pragma omp task
			qsort_index_rec(left, k, nr, spawn);
		} else {
			qsort_index_rec(left, k, nr, spawn);
		}
	}
}

void orc_qsort_index(size_t n, const orc_key *keys, size_t *perm, int nthreads)
{
	if (nthreads < 1)
		nthreads = 1;
	for (size_t i = 0; i < n; i++)
		perm[i] = i;
	if (n < 2)
		return;
	size_t spawn = n / (size_t)nthreads / N_PARTITIONS_PER_CPU * 2;
	if (spawn < PARALLEL_THRESHOLD)
		spawn = PARALLEL_THRESHOLD;
	#pragma omp parallel num_threads(nthreads)
	#pragma omp single
	qsort_index_rec(perm, keys, n, spawn);
}

#define DEFINE_REORDER(NAME, T)                                              \
void NAME(size_t n, void *pv, const size_t *idx_in, size_t *idx)            \
{                                                                            \
	T *p = (T *)pv;                                                          \
	memcpy(idx, idx_in, n * sizeof(*idx));                                   \
	for (size_t i = 0; i < n; i++) {                                         \
		if (idx[i] == i)                                                     \
			continue;                                                        \
		size_t dst = i, src = idx[i];                                        \
		T held = p[i];                                                       \
		for (;;) {                                                           \
			p[dst] = p[src];                                                 \
			idx[dst] = dst;                                                  \
			dst = src;                                                       \
			src = idx[dst];                                                  \
			if (src == i)                                                    \
				break;                                                       \
		}                                                                    \
		p[dst] = held;                                                       \
		idx[dst] = dst;                                                      \
	}                                                                        \
}

DEFINE_REORDER(orc_reorder_8, uint64_t)
DEFINE_REORDER(orc_reorder_4, uint32_t)
DEFINE_REORDER(orc_reorder_16, orc_key)
