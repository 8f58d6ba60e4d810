/*
 * Test infrastructure only (the CPU oracle): Domain_Decomposition with
 * NTarget targets, restated from domain.c for one rank and one thread.
 *
 *   find_min_level         domain.c:300-304
 *   reset_bunchlist        domain.c:269-294
 *   fill_new_bunches       domain.c:462-531 (one thread: every bunch's Cost
 *                          is the float sum of cost_metric in particle order)
 *   cost_metric            domain.c:533-536
 *   find_mean_cost         domain.c:538-556
 *   remove_empty_bunches   domain.c:430-460
 *   distribute             domain.c:574-627
 *   find_global_imbalances domain.c:629-649
 *   mark_bunches_to_split  domain.c:655-676
 *   make_new_bunchlist     domain.c:370-397, split_bunch 405-428
 *   the loop               domain.c:67-115
 *
 * Qsort of fewer than 1024 elements is the C library's qsort
 * (sort.c:39-51); glibc's is a merge sort, so bunches of equal Cost keep
 * their key order in distribute().  The sorts here are explicitly stable
 * to fix that order.  Keys are the short (upper 64-bit) Peano keys of the
 * particles, sorted ascending as Sort_Particles_By_Peano_Key leaves them.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MIN_LEVEL 2                      /* domain.c:3 */
#define N_SHORT_TRIPLETS 21              /* 64-bit shortKey / 3 */

typedef struct {
	uint64_t Key;
	int Target, Level, First_Part;
	uint64_t Npart;
	float Cost;
	int Modify;
} bunch;

static int find_min_level(int ntarget)
{
	return fmax(MIN_LEVEL, log(ntarget) / log(8) + 1);
}

static unsigned int cost_metric(double c)
{
	return fmax(1, c);
}

/* bunches [b0, b0 + nb) in ascending key order take particles
 * [first, first + npart): each particle goes to the first bunch whose key
 * is not below its own */
static void fill_new_bunches(bunch *D, int b0, int nb, const uint64_t *key,
		const double *cost, size_t first, size_t npart)
{
	for (int i = b0; i < b0 + nb; i++) {
		D[i].Npart = 0;
		D[i].Cost = 0;
		D[i].First_Part = INT_MAX;
	}
	int run = b0;
	for (size_t ip = first; ip < first + npart; ip++) {
		while (D[run].Key < key[ip])
			run++;
		D[run].Npart++;
		D[run].Cost += cost_metric(cost[ip]);
		if ((int)ip < D[run].First_Part)
			D[run].First_Part = (int)ip;
	}
}

/* stable insertion sort: descending key, or descending cost */
static void sort_by_key(bunch *D, int n)
{
	for (int i = 1; i < n; i++) {
		bunch t = D[i];
		int j = i - 1;
		while (j >= 0 && D[j].Key < t.Key) {
			D[j + 1] = D[j];
			j--;
		}
		D[j + 1] = t;
	}
}

static void sort_by_cost(bunch *D, int n)
{
	for (int i = 1; i < n; i++) {
		bunch t = D[i];
		int j = i - 1;
		while (j >= 0 && D[j].Cost < t.Cost) {
			D[j + 1] = D[j];
			j--;
		}
		D[j + 1] = t;
	}
}

int orc_domain_targets(const uint64_t *key, const double *cost, size_t n,
		int ntarget, double part_alloc_factor, double imbal_ceil,
		orc_topnode *out, int max_out, int *ntop, int *iterations)
{
	if (ntarget < 1 || n == 0 || n >= INT_MAX)
		return -1;
	int cap = 65536, nbunch;
	bunch *D = calloc((size_t)cap, sizeof(*D));
	float *Cost = calloc((size_t)ntarget, sizeof(float));
	int *Npart = calloc((size_t)ntarget, sizeof(int));
	int *Split_Idx = calloc((size_t)ntarget, sizeof(int));
	int rc = 0;

	/* reset_bunchlist */
	const int level = find_min_level(ntarget);
	nbunch = (int)pow(8, level);
	if (nbunch > cap) {
		rc = -1;
		goto done;
	}
	{
		const int shift = 3 * level;
		const uint64_t base = 0xFFFFFFFFFFFFFFFFULL >> shift;
		for (int i = 0; i < nbunch; i++) {
			D[i].Key = ((uint64_t)i << (64 - shift)) | base;
			D[i].Level = level;
			D[i].First_Part = INT_MAX;
			D[i].Target = -1;
		}
	}
	fill_new_bunches(D, 0, nbunch, key, cost, 0, n);

	/* find_mean_cost */
	double mean_cost = 0;
	for (int i = 0; i < nbunch; i++)
		mean_cost += D[i].Cost;
	mean_cost /= ntarget;
	const double mean_npart = ((double)n) / ntarget;

	int cnt = 0;
	for (;;) {
		/* remove_empty_bunches */
		int m = 0;
		for (int i = 0; i < nbunch; i++)
			if (D[i].Npart != 0)
				D[m++] = D[i];
		nbunch = m;
		if (nbunch < ntarget) {      /* Split_Idx would index past D */
			rc = -2;
			goto done;
		}
		sort_by_key(D, nbunch);

		/* distribute */
		memset(Cost, 0, (size_t)ntarget * sizeof(*Cost));
		memset(Npart, 0, (size_t)ntarget * sizeof(*Npart));
		sort_by_cost(D, nbunch);
		for (int i = 0; i < nbunch; i++) {
			int task = 0;
			double max_delta_mean = 0;
			for (int j = 0; j < ntarget; j++) {
				double delta_mean = mean_cost - Cost[j];
				if (delta_mean > max_delta_mean) {
					task = j;
					max_delta_mean = delta_mean;
				}
			}
			Cost[task] += D[i].Cost;
			Npart[task] += D[i].Npart;
			D[i].Target = -task - 1;
		}
		sort_by_key(D, nbunch);
		for (int i = 0; i < ntarget; i++)
			Split_Idx[i] = i;
		for (int i = 0; i < nbunch; i++) {
			int task = -1 * (D[i].Target + 1);
			int idx = Split_Idx[task];
			if (D[i].Cost > D[idx].Cost)
				Split_Idx[task] = i;
		}

		/* find_global_imbalances */
		double max_mem = -DBL_MAX, max_cost = -DBL_MAX;
		for (int i = 0; i < ntarget; i++) {
			double ci = fabs(Cost[i] - mean_cost) / mean_cost;
			if (ci > max_cost)
				max_cost = ci;
			double ni = fabs(Npart[i] - mean_npart) / mean_npart;
			if (ni > max_mem)
				max_mem = ni;
		}
		if (max_cost < imbal_ceil)
			if (max_mem < part_alloc_factor - 1)
				break;
		if (cnt++ > N_SHORT_TRIPLETS - MIN_LEVEL)
			break;

		/* mark_bunches_to_split */
		for (int i = 0; i < ntarget; i++) {
			int j = Split_Idx[i];
			if (j < 0 || D[j].Level == N_SHORT_TRIPLETS)
				continue;
			D[j].Modify = 1;
		}

		/* make_new_bunchlist */
		const int old = nbunch;
		for (int i = 0; i < old; i++) {
			if (D[i].Modify == 0)
				continue;
			if (nbunch + 8 > cap) {
				rc = -1;
				goto done;
			}
			const int first = nbunch;
			nbunch += 8;
			for (int k = 0; k < 8; k++) {        /* split_bunch */
				bunch *d = &D[first + k];
				d->Level = D[i].Level + 1;
				const int shift = 64 - 3 * d->Level;
				const uint64_t mask = 0x7ULL << shift;
				d->Key = (D[i].Key & ~mask) | ((uint64_t)k << shift);
				d->Npart = 0;
				d->First_Part = INT_MAX;
				d->Target = -1;
				d->Modify = 0;
			}
			fill_new_bunches(D, first, 8, key, cost,
					(size_t)D[i].First_Part, (size_t)D[i].Npart);
			D[i].Npart = 0;
		}
	}
	if (nbunch > max_out) {
		rc = -1;
		goto done;
	}
	for (int i = 0; i < nbunch; i++) {   /* transform_bunches_into_top_nodes */
		memset(&out[i], 0, sizeof(out[i]));
		out[i].Key = D[i].Key;
		out[i].Target = D[i].Target;
		out[i].Level = D[i].Level;
		out[i].First_Part = D[i].First_Part;
		out[i].Npart = (int)D[i].Npart;
	}
	*ntop = nbunch;
	*iterations = cnt;
done:
	free(D);
	free(Cost);
	free(Npart);
	free(Split_Idx);
	return rc;
}
