/*
 * Domain_Decomposition with NTarget > 1 (domain.c:67-145): the bunch
 * refinement loop that places top nodes where the work is.
 *
 * The particle passes run on the device: a bunch's particles are a
 * contiguous range of the key-sorted particles, found by two binary
 * searches of its largest short key (and its predecessor's) over the
 * parent's range, and its Cost is the float sum of cost_metric over that
 * range in particle order (fill_new_bunches, domain.c:462-531, with one
 * thread), one wavefront per bunch.  The loop's bookkeeping over at most a
 * few thousand bunches (remove_empty_bunches, the key and cost sorts,
 * distribute, find_global_imbalances, mark_bunches_to_split,
 * make_new_bunchlist; domain.c:370-460, 571-676) runs on the host, in the
 * reference's types (float bunch and task costs, double means).  The sorts
 * are stable: the reference's Qsort of fewer than 1024 bunches is the C
 * library's qsort (sort.c:39-51), glibc's merge sort.
 */
#include "tdv_internal.h"
#include "tdv_device.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

#define TDV_MIN_LEVEL 2                  /* domain.c:3 */
#define TDV_SHORT_TRIPLETS 21            /* 64-bit short keys */

namespace {

struct BunchFill {                       /* one bunch to fill */
	unsigned long long key, prev;        /* its largest key, its predecessor's */
	long long lo, hi;                    /* the particle range searched */
	int has_prev, pad;
};

struct BunchSum {
	int first, npart;
	float cost;
	int pad;
};

struct Bunch {
	unsigned long long key;
	int target, level, first;
	unsigned long long npart;
	float cost;
	int modify;
};

} // namespace

/* first particle in [lo, hi) whose short key exceeds k */
__device__ __forceinline__ long long upper_bound_short(
		const uint64_t *__restrict__ key, long long lo, long long hi,
		unsigned long long k)
{
	while (lo < hi) {
		const long long mid = (lo + hi) >> 1;
		if (key[2 * mid + 1] <= k)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

#define TDV_FILL_PER_LANE 16              /* particles per lane per batch */

/* fill_new_bunches: a particle joins the first bunch (ascending keys)
 * whose key is not below its own; the bunch's Cost is the float sum of
 * cost_metric (domain.c:533-536: (unsigned) fmax(1, P.Cost)) one particle
 * after the other.  One wavefront per bunch, in
 * batches of 64 x TDV_FILL_PER_LANE particles: the lanes convert a batch
 * (loaded during the previous one) into LDS, issue the next batch's loads
 * and lane 0 adds the batch in particle order while they are in flight. */
__global__ void __launch_bounds__(64) k_bunch_fill(int nb,
		const BunchFill *__restrict__ req, const uint64_t *__restrict__ key,
		const double *__restrict__ cost, BunchSum *__restrict__ out)
{
	constexpr int B = TDV_FILL_PER_LANE;
	__shared__ float buf[64 * B];
	__shared__ long long range[2];
	const int b = blockIdx.x;
	const int lane = threadIdx.x;
	if (b >= nb)
		return;
	if (lane == 0) {
		const BunchFill r = req[b];
		const long long a = r.has_prev
				? upper_bound_short(key, r.lo, r.hi, r.prev) : r.lo;
		range[0] = a;
		range[1] = upper_bound_short(key, a, r.hi, r.key);
	}
	__syncthreads();
	const long long a = range[0], e = range[1];
	float s = 0.f;
	if (a < e) {
		/* loads at indices clamped into [a, e): every batch issues the same
		 * loads; particles past e are not added */
		double raw[B];
#pragma unroll
		for (int u = 0; u < B; u++) {
			const long long k = a + lane + 64 * u;
			raw[u] = cost[k < e ? k : e - 1];
		}
		for (long long base = a; base < e; base += 64 * B) {
			float v[B];
#pragma unroll
			for (int u = 0; u < B; u++)
				v[u] = (float)(unsigned)fmax(1.0, raw[u]);
#pragma unroll
			for (int u = 0; u < B; u++) {
				const long long k = base + 64 * B + lane + 64 * u;
				raw[u] = cost[k < e ? k : e - 1];
			}
#pragma unroll
			for (int u = 0; u < B; u++)
				buf[lane + 64 * u] = v[u];
			__syncthreads();
			if (lane == 0) {
				const int cnt = (int)(e - base < 64 * B ? e - base : 64 * B);
				for (int q = 0; q < cnt; q++)
					s = __fadd_rn(s, buf[q]);
			}
			__syncthreads();
		}
	}
	if (lane == 0) {
		BunchSum o;
		o.first = a < e ? (int)a : INT_MAX;
		o.npart = (int)(e - a);
		o.cost = s;
		o.pad = 0;
		out[b] = o;
	}
}

__global__ void k_first_positions(int nt, const int *__restrict__ first,
		const double *__restrict__ x, const double *__restrict__ y,
		const double *__restrict__ z, double *__restrict__ out)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nt)
		return;
	const int i = first[t];
	out[3 * t] = x[i];
	out[3 * t + 1] = y[i];
	out[3 * t + 2] = z[i];
}

namespace {

/* device scratch for the fills, grown on demand, freed on return */
struct FillScratch {
	tdv_context *c;
	BunchFill *req = nullptr;
	BunchSum *sum = nullptr;
	size_t cap = 0;
	~FillScratch()
	{
		if (req)
			(void)hipFree(req);
		if (sum)
			(void)hipFree(sum);
	}
	int reserve(size_t n)
	{
		if (n <= cap)
			return 0;
		if (req)
			(void)hipFree(req);
		if (sum)
			(void)hipFree(sum);
		req = nullptr;
		sum = nullptr;
		cap = std::max(n, (size_t)512);
		TDV_HIP(c, hipMalloc((void **)&req, cap * sizeof(BunchFill)));
		TDV_HIP(c, hipMalloc((void **)&sum, cap * sizeof(BunchSum)));
		return 0;
	}
};

/* fill the groups of bunches D[g.first .. g.first + 8) (or the whole
 * initial list): each group covers the particle range [lo, lo + n) with
 * its bunches in ascending key order */
struct Group {
	int b0, nb;
	long long lo, n;
};

int fill_groups(tdv_context *c, FillScratch &fs, std::vector<Bunch> &D,
		const std::vector<Group> &groups)
{
	std::vector<BunchFill> req;
	for (const Group &g : groups)
		for (int k = 0; k < g.nb; k++) {
			BunchFill r;
			r.key = D[g.b0 + k].key;
			r.prev = k ? D[g.b0 + k - 1].key : 0;
			r.has_prev = k > 0;
			r.lo = g.lo;
			r.hi = g.lo + g.n;
			r.pad = 0;
			req.push_back(r);
		}
	if (req.empty())
		return 0;
	TDV_TRY(fs.reserve(req.size()));
	TDV_HIP(c, hipMemcpyAsync(fs.req, req.data(), req.size() * sizeof(BunchFill),
			hipMemcpyHostToDevice, c->stream));
	hipLaunchKernelGGL(k_bunch_fill, dim3((unsigned)req.size()), dim3(64), 0,
			c->stream, (int)req.size(), fs.req, c->key, c->cost, fs.sum);
	TDV_HIP(c, hipGetLastError());
	std::vector<BunchSum> sum(req.size());
	TDV_HIP(c, hipMemcpyAsync(sum.data(), fs.sum, sum.size() * sizeof(BunchSum),
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	size_t j = 0;
	for (const Group &g : groups)
		for (int k = 0; k < g.nb; k++, j++) {
			Bunch &b = D[g.b0 + k];
			b.first = sum[j].first;
			b.npart = (unsigned long long)sum[j].npart;
			b.cost = sum[j].cost;
		}
	return 0;
}

/* descending key (compare_bunches_by_key, domain.c:678-684) */
bool by_key(const Bunch &x, const Bunch &y) { return x.key > y.key; }

/* descending cost (compare_bunches_by_cost, domain.c:686-692) */
bool by_cost(const Bunch &x, const Bunch &y) { return x.cost > y.cost; }

} // namespace

int tdv_domain_refine(tdv_context *c, int ntarget, double part_alloc_factor,
		double imbal_ceil, std::vector<TopRecordIn> &out, int *iterations)
{
	const size_t n = c->n;
	FillScratch fs;
	fs.c = c;

	/* reset_bunchlist, find_min_level */
	const int level = (int)fmax(TDV_MIN_LEVEL, log(ntarget) / log(8) + 1);
	if (level > 4)
		return tdv_fail(c, TDV_ERR_ARG, "domain decomposition: too many "
				"targets");
	int nb = (int)pow(8, level);
	std::vector<Bunch> D(nb);
	const int shift = 3 * level;
	const unsigned long long base = 0xFFFFFFFFFFFFFFFFULL >> shift;
	for (int i = 0; i < nb; i++) {
		D[i].key = ((unsigned long long)i << (64 - shift)) | base;
		D[i].level = level;
		D[i].target = -1;
		D[i].first = INT_MAX;
		D[i].npart = 0;
		D[i].cost = 0;
		D[i].modify = 0;
	}
	TDV_TRY(fill_groups(c, fs, D, { Group{ 0, nb, 0, (long long)n } }));

	/* find_mean_cost */
	double mean_cost = 0;
	for (const Bunch &b : D)
		mean_cost += b.cost;
	mean_cost /= ntarget;
	const double mean_npart = ((double)n) / ntarget;

	std::vector<float> cost(ntarget);
	std::vector<int> npart(ntarget), split_idx(ntarget);
	int cnt = 0;
	for (;;) {
		/* remove_empty_bunches */
		D.erase(std::remove_if(D.begin(), D.end(),
				[](const Bunch &b) { return b.npart == 0; }), D.end());
		if ((int)D.size() < ntarget)
			return tdv_fail(c, TDV_ERR_ARG, "domain decomposition: fewer "
					"non-empty bunches than targets");
		std::stable_sort(D.begin(), D.end(), by_key);

		/* distribute */
		std::fill(cost.begin(), cost.end(), 0.f);
		std::fill(npart.begin(), npart.end(), 0);
		std::stable_sort(D.begin(), D.end(), by_cost);
		for (Bunch &b : D) {
			int task = 0;
			double max_delta_mean = 0;
			for (int j = 0; j < ntarget; j++) {
				const double delta_mean = mean_cost - cost[j];
				if (delta_mean > max_delta_mean) {
					task = j;
					max_delta_mean = delta_mean;
				}
			}
			cost[task] += b.cost;
			npart[task] = (int)((unsigned long long)npart[task] + b.npart);
			b.target = -task - 1;
		}
		std::stable_sort(D.begin(), D.end(), by_key);
		for (int i = 0; i < ntarget; i++)
			split_idx[i] = i;
		for (int i = 0; i < (int)D.size(); i++) {
			const int task = -(D[i].target + 1);
			if (D[i].cost > D[split_idx[task]].cost)
				split_idx[task] = i;
		}

		/* find_global_imbalances */
		double max_mem = -DBL_MAX, max_cost = -DBL_MAX;
		for (int i = 0; i < ntarget; i++) {
			max_cost = std::max(max_cost, fabs(cost[i] - mean_cost) / mean_cost);
			max_mem = std::max(max_mem, fabs(npart[i] - mean_npart) / mean_npart);
		}
		if (max_cost < imbal_ceil && max_mem < part_alloc_factor - 1)
			break;
		if (cnt++ > TDV_SHORT_TRIPLETS - TDV_MIN_LEVEL)
			break;

		/* mark_bunches_to_split */
		for (int i = 0; i < ntarget; i++)
			if (D[split_idx[i]].level != TDV_SHORT_TRIPLETS)
				D[split_idx[i]].modify = 1;

		/* make_new_bunchlist, split_bunch */
		const int old = (int)D.size();
		std::vector<Group> groups;
		for (int i = 0; i < old; i++) {
			if (!D[i].modify)
				continue;
			const int first = (int)D.size();
			for (int k = 0; k < 8; k++) {
				Bunch d;
				d.level = D[i].level + 1;
				const int sh = 64 - 3 * d.level;
				const unsigned long long mask = 0x7ULL << sh;
				d.key = (D[i].key & ~mask) | ((unsigned long long)k << sh);
				d.npart = 0;
				d.first = INT_MAX;
				d.target = -1;
				d.cost = 0;
				d.modify = 0;
				D.push_back(d);
			}
			groups.push_back(Group{ first, 8, (long long)D[i].first,
					(long long)D[i].npart });
			D[i].npart = 0;
		}
		if (D.size() > (size_t)TDV_MAX_TOP)
			return tdv_fail(c, TDV_ERR_ARG, "domain decomposition: too many "
					"bunches");
		TDV_TRY(fill_groups(c, fs, D, groups));
	}

	/* transform_bunches_into_top_nodes (domain.c:220-250): the centre of
	 * the level-L cell holding the bunch's first particle */
	const int nt = (int)D.size();
	std::vector<int> first(nt);
	for (int i = 0; i < nt; i++)
		first[i] = D[i].first;
	int *first_dev = nullptr;
	double *pos_dev = nullptr;
	std::vector<double> pos(3 * (size_t)nt);
	int rc = tdv_check(c, hipMalloc((void **)&first_dev, nt * sizeof(int)),
			"hipMalloc");
	if (!rc)
		rc = tdv_check(c, hipMalloc((void **)&pos_dev, 3 * nt * sizeof(double)),
				"hipMalloc");
	if (!rc)
		rc = tdv_check(c, hipMemcpyAsync(first_dev, first.data(),
				nt * sizeof(int), hipMemcpyHostToDevice, c->stream), "copy");
	if (!rc) {
		hipLaunchKernelGGL(k_first_positions, dim3(tdv_blocks(nt, 64)),
				dim3(64), 0, c->stream, nt, first_dev, c->pos[0], c->pos[1],
				c->pos[2], pos_dev);
		rc = tdv_check(c, hipGetLastError(), "k_first_positions");
	}
	if (!rc)
		rc = tdv_check(c, hipMemcpyAsync(pos.data(), pos_dev,
				3 * nt * sizeof(double), hipMemcpyDeviceToHost, c->stream),
				"copy");
	if (!rc)
		rc = tdv_check(c, hipStreamSynchronize(c->stream), "sync");
	if (first_dev)
		(void)hipFree(first_dev);
	if (pos_dev)
		(void)hipFree(pos_dev);
	if (rc)
		return rc;
	out.assign(nt, TopRecordIn{});
	for (int i = 0; i < nt; i++) {
		TopRecordIn &r = out[i];
		r.Key = D[i].key;
		r.Target = D[i].target;
		r.Level = D[i].level;
		r.First_Part = D[i].first;
		r.Npart = (int)D[i].npart;
		const double size = c->dom_size / (double)(1ULL << D[i].level);
		for (int d = 0; d < 3; d++) {
			const double p = pos[3 * i + d] - c->dom_origin[d];
			r.Pos[d] = (float)((floor(p / size) + 0.5) * size
					+ c->dom_origin[d]);
		}
	}
	*iterations = cnt;
	return 0;
}
