/*
 * Top nodes and the octree, built bottom-up from the sorted keys in
 * coalesced particle-parallel passes.
 *
 * Top nodes (domain.c:67-145, 220-313 with NTarget = 1): the non-empty
 * level-2 PH cells in descending key order, each with a float centre.
 *
 * Tree (Gravity/tree_build.c:47-572): the reference inserts particles one
 * by one; this build derives the same node set and the same depth-first
 * array from per-particle quantities (first divergence levels of adjacent
 * keys, the reference's collapse events and the cells they collapse):
 *   k_tree_marks   events and collapsed cells -> min collapse level per
 *                  particle (atomicMin over <= 64 particles per event)
 *   k_tree_counts  nodes owned by each particle: its chain of cells from
 *                  dprev(i) down to its leaf or collapsed bundle
 *   scan           node offsets in (first particle, level) order, which is
 *                  the reference's depth-first order
 *   k_tree_targets subtree sizes and Target offsets of the top nodes
 *   k_tree_nodes   one lane per chain: a forward scan over the chain's
 *                  largest cell gives every chain node its particle count,
 *                  skip link, and the running FP64 sums of m x and m in PH
 *                  order (bit-identical to the reference's incremental
 *                  sums), plus the parent links of the chains it contains
 *   k_tree_centres centres by the reference's sign rule, top-down along
 *                  the parent links from the float top-node centre
 */
#include "tdv_internal.h"
#include "tdv_device.h"
#include <algorithm>

#define NOCOLLAPSE 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* top nodes                                                                 */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ size_t lower_bound_cell(const uint64_t *key,
		size_t n, unsigned cell)
{
	size_t lo = 0, hi = n;
	while (lo < hi) {
		const size_t mid = (lo + hi) >> 1;
		if ((unsigned)(key[2 * mid + 1] >> 58) < cell)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

__global__ void __launch_bounds__(64) k_top_nodes(size_t n,
		const uint64_t *__restrict__ key, const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z, double o0,
		double o1, double o2, double dom_size, TopNodes t)
{
	__shared__ int nonempty[64];
	const unsigned cell = threadIdx.x;
	const size_t f = lower_bound_cell(key, n, cell);
	const size_t e = lower_bound_cell(key, n, cell + 1);
	const int np = (int)(e - f);
	nonempty[cell] = np > 0;
	__syncthreads();
	int rank = 0, total = 0;
	for (int q = 0; q < 64; q++) {
		total += nonempty[q];
		if (q > (int)cell)
			rank += nonempty[q];
	}
	if (cell == 0)
		*t.ntop = total;
	if (np == 0)
		return;
	t.first[rank] = (int)f;
	t.npart[rank] = np;
	t.level[rank] = TDV_TOP_LEVEL;
	t.key[rank] = ((uint64_t)cell << 58) | (~0ull >> 6);
	const double size = dom_size / (double)(1ull << TDV_TOP_LEVEL);
	const double p[3] = { x[f] - o0, y[f] - o1, z[f] - o2 };
	const double o[3] = { o0, o1, o2 };
	for (int d = 0; d < 3; d++)
		t.pos[3 * rank + d] = (float)((floor(p[d] / size) + 0.5) * size + o[d]);
	t.target[rank] = -1;
}

/* per particle: the top node holding it (top nodes are contiguous particle
 * ranges in descending key order, i.e. descending first particle) */
__global__ void k_assign_top(size_t n, TopNodes t)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const int ntop = *t.ntop;
	int lo = 0, hi = ntop - 1;          /* first[] decreases with index */
	while (lo < hi) {
		const int mid = (lo + hi) >> 1;
		if ((size_t)t.first[mid] > i)
			lo = mid + 1;
		else
			hi = mid;
	}
	t.ptop[i] = lo;
}

int tdv_assign_top(tdv_context *c)
{
	hipLaunchKernelGGL(k_assign_top, dim3(tdv_blocks(c->n, 256)), dim3(256), 0,
			c->stream, c->n, c->top);
	return tdv_check(c, hipGetLastError(), "k_assign_top");
}

int tdv_build_top_nodes(tdv_context *c)
{
	hipLaunchKernelGGL(k_top_nodes, dim3(1), dim3(64), 0, c->stream, c->n,
			c->key, c->pos[0], c->pos[1], c->pos[2], c->dom_origin[0],
			c->dom_origin[1], c->dom_origin[2], c->dom_size, c->top);
	TDV_HIP(c, hipGetLastError());
	if (!c->graph_mode) {            /* a captured step keeps it on the device */
		TDV_HIP(c, hipMemcpyAsync(&c->ntop, c->top.ntop, sizeof(int),
				hipMemcpyDeviceToHost, c->stream));
		TDV_HIP(c, hipStreamSynchronize(c->stream));
	}
	return tdv_assign_top(c);
}

/* ------------------------------------------------------------------------ */
/* tree                                                                      */
/* ------------------------------------------------------------------------ */

/* the triplet a node's Bitfield holds (tree.h:25,58): the reversed key's
 * (key_fragment, tree_build.c:580-585), whose level 42 has lost its top
 * bit (peano.c:254-275) */
__device__ __forceinline__ unsigned bitfield_triplet(Key2 k, int lv)
{
	return key_triplet(k, lv) & (lv >= TDV_NT ? 3u : 7u);
}

struct TopRange {
	int t, first, last;     /* top node, its particle range */
	int level;              /* top node level */
	bool tree;              /* top node holds > VECTOR_SIZE particles */
};

__device__ __forceinline__ TopRange top_of(const uint64_t *key, size_t i,
		const TopNodes &T, int vs)
{
	(void)key;
	TopRange r;
	r.t = T.ptop[i];
	r.first = T.first[r.t];
	r.last = r.first + T.npart[r.t] - 1;
	r.level = T.level[r.t];
	r.tree = T.npart[r.t] > vs;
	return r;
}

/* dprev of particle i: D(i-1, i), the top level for the first particle */
__device__ __forceinline__ int dprev_of(const uint64_t *key, int i, int first,
		int top_level)
{
	if (i == first)
		return top_level;
	return div_level(load_key(key, i - 1), load_key(key, i));
}

/* particles in the cell of level lvl that ends at i-1, counted back no
 * further than first (bounded: vs+1 when the cell holds more than vs
 * particles).  The keys are sorted, so the cell's keys are the run ending
 * at i-1 of keys that share key[i-1]'s first lvl levels: a binary search
 * over the last vs+1 keys finds its start in 7 dependent loads, where a
 * scan back took up to 64. */
__device__ __forceinline__ int cell_count_back(const uint64_t *key, int i,
		int first, int lvl, int vs)
{
	const Key2 kl = load_key(key, i - 1);
	int lo = i - vs - 1 > first ? i - vs - 1 : first, hi = i - 1;
	while (lo < hi) {                    /* key[hi] is in the cell */
		const int mid = (lo + hi) >> 1;
		if (div_level(load_key(key, mid), kl) > lvl)
			hi = mid;
		else
			lo = mid + 1;
	}
	return i - hi <= vs ? i - hi : vs + 1;
}

__global__ void __launch_bounds__(256) k_tree_marks(size_t n,
		const uint64_t *__restrict__ key, TopNodes T, int vs,
		uint32_t *__restrict__ minpack, unsigned *err)
{
	const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (gi >= n)
		return;
	const int i = (int)gi;
	const TopRange r = top_of(key, gi, T, vs);
	if (!r.tree || i == r.first)
		return;
	const int dp = dprev_of(key, i, r.first, r.level);
	/* below the key resolution (its key matches the previous one down to
	 * level 41, or entirely): the reference dumps the particle into the
	 * level-41 node it reached, with no node of its own and no collapse
	 * event, leaving last_key and last_parent as they were
	 * (tree_build.c:337-342) */
	if (dp >= TDV_NT)
		return;
	/* the previous particle that was inserted: dumped ones are skipped */
	int h = i - 1;
	while (h > r.first && dprev_of(key, h, r.first, r.level) >= TDV_NT)
		h--;
	const int dpp = dprev_of(key, h, r.first, r.level);
	const bool newb = (h >= r.first + 1) && dp < dpp;
	if (!(newb || i == r.last))
		return;
	int lk = -1, cnt = 0;
	const int c1 = cell_count_back(key, i, r.first, dp, vs);
	if (c1 <= vs) {
		lk = dp;
		cnt = c1;
	} else if (newb) {
		const int c2 = cell_count_back(key, i, r.first, dpp - 1, vs);
		if (c2 <= vs) {
			lk = dpp - 1;
			cnt = c2;
		}
	}
	if (lk < 0)
		return;
	const uint32_t packed = ((uint32_t)lk << 24) | (uint32_t)cnt;
	for (int q = i - cnt; q < i; q++)
		atomicMin(&minpack[q], packed);
}

__global__ void __launch_bounds__(256) k_tree_counts(size_t n,
		const uint64_t *__restrict__ key, TopNodes T, int vs,
		const uint32_t *__restrict__ minpack, uint32_t *__restrict__ cnt,
		int *__restrict__ lohi)
{
	const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (gi >= n)
		return;
	const int i = (int)gi;
	const TopRange r = top_of(key, gi, T, vs);
	if (!r.tree) {
		cnt[i] = 0;
		lohi[i] = 0;
		return;
	}
	const int lo = dprev_of(key, i, r.first, r.level);
	if (lo >= TDV_NT) {              /* dumped: no node of its own */
		cnt[i] = 0;
		lohi[i] = 0;
		return;
	}
	const int dn = (i == r.last) ? -1 : div_level(load_key(key, i),
			load_key(key, i + 1));
	const int lf = lo > dn ? lo : dn;
	const uint32_t mp = minpack[i];
	const int ml = (mp == NOCOLLAPSE) ? 99 : (int)(mp >> 24);
	const int hi = lf < ml ? lf : ml;
	const int c = hi >= lo ? hi - lo + 1 : 0;
	cnt[i] = (uint32_t)c;
	lohi[i] = lo | (hi << 8) | ((ml <= lf) ? (1 << 16) : 0);
}

/* the node ranges of the top nodes' subtrees; more nodes than `cap` (a
 * captured step's arrays) set the node count to -1, which every later
 * kernel of the step reads as "do nothing", and raise TDV_EBIT_RANGE */
__global__ void __launch_bounds__(64) k_tree_targets(TopNodes T, int vs,
		const uint32_t *__restrict__ off, long long cap, unsigned *err)
{
	if (threadIdx.x != 0)
		return;
	const int ntop = *T.ntop;
	long long run = 0;
	for (int t = 0; t < ntop; t++) {
		long long sz = 0;
		if (T.npart[t] > vs)
			sz = (long long)(off[T.first[t] + T.npart[t]]
					- off[T.first[t]]) + 1;
		T.target[t] = (int)run;      /* reserve_tree_memory(nNeeded) */
		T.term[t] = (int)(run + sz - 1);
		run += sz;
	}
	if (run > cap) {
		atomicOr(err, TDV_EBIT_RANGE);
		*T.nnodes = -1;
	} else {
		*T.nnodes = (int)run;
	}
}

/* node records cleared before the build fills them (the count on the
 * device) */
__global__ void k_clear_nodes(size_t bound, const TopNodes T,
		int4 *__restrict__ nc, double4 *__restrict__ dp)
{
	const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j < bound && (long long)j < (long long)*T.nnodes) {
		nc[j] = make_int4(0, 0, 0, 0);
		dp[j] = make_double4(0, 0, 0, 0);
	}
}

__global__ void k_fill_u32_val(size_t n, uint32_t *__restrict__ p, uint32_t v)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

__device__ __forceinline__ int node_index(const TopNodes &T, int t, int first,
		const uint32_t *off, int p, int last)
{
	if (p > last)
		return T.term[t];
	return T.target[t] + (int)(off[p] - off[first]);
}

__global__ void __launch_bounds__(256) k_tree_nodes(size_t n,
		const uint64_t *__restrict__ key, const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z,
		const double *__restrict__ m, TopNodes T, int vs,
		const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
		const int *__restrict__ lohi, double4 *__restrict__ na,
		int4 *__restrict__ nc, int *__restrict__ up,
		int *__restrict__ tree_parent, int2 *__restrict__ long_list,
		unsigned long_cap, unsigned *__restrict__ long_count, unsigned *err)
{
	const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (gi >= n)
		return;
	const int i = (int)gi;
	if (cnt[i] == 0 || *T.nnodes < 0)
		return;
	const TopRange r = top_of(key, gi, T, vs);
	const int lo = lohi[i] & 0xFF, hi = (lohi[i] >> 8) & 0xFF;
	const int base = node_index(T, r.t, r.first, off, i, r.last);
	const Key2 ki = load_key(key, i);
	{
		/* the chain's scan covers the level-lo cell starting at i: its last
		 * particle is the last k with div_level(key_i, key_k) > lo (keys are
		 * sorted, so that level only falls with k) */
		int a = i, b = r.last;
		while (a < b) {
			const int mid = a + ((b - a + 1) >> 1);
			if (div_level(ki, load_key(key, mid)) > lo)
				a = mid;
			else
				b = mid - 1;
		}
		if (a - i + 1 > TDV_LONG_CHAIN) {    /* k_tree_nodes_long's */
			const unsigned slot = atomicAdd(long_count, 1u);
			if (slot < long_cap)
				long_list[slot] = make_int2(i, a);
			else
				atomicOr(err, TDV_EBIT_RANGE);
			return;
		}
	}

	double sx = 0, sy = 0, sz = 0, sm = 0;
	int cur = hi;
	int k = i;
	int runmin = 99;
	Key2 kprev = ki;
	for (;;) {
		const double mk = m[k];
		const double tx = x[k] * mk, ty = y[k] * mk, tz = z[k] * mk;
		sx = sx + tx;
		sy = sy + ty;
		sz = sz + tz;
		sm = sm + mk;
		const int kn = k + 1;
		int dn = -1;
		Key2 kk = kprev;
		if (kn <= r.last) {
			kk = load_key(key, kn);
			dn = div_level(kprev, kk);
		}
		if (dn <= cur) {                     /* cells of levels >= dn end at k */
			const int lowest = dn > lo ? dn : lo;
			const int next = node_index(T, r.t, r.first, off, kn, r.last);
			for (int lv = lowest; lv <= cur; lv++) {
				const int j = base + (lv - lo);
				const bool leaf = lv == hi;
				int bits = lv | ((int)bitfield_triplet(ki, lv) << 6)
						| (1 << TDV_BIT_LOCAL);
				const bool root = j == T.target[r.t];
				if (root)
					bits |= 1 << TDV_BIT_TOP;
				const int dnext = leaf ? (-i - 1) : (next - j);
				na[j] = make_double4(sx / sm, sy / sm, sz / sm, sm);
				nc[j] = make_int4(dnext, kn - i, bits, i);
				/* the DUp of a chain head comes from its parent's scan */
				if (root)
					up[j] = r.t;
				else if (lv > lo)
					up[j] = 1;
				if (leaf)
					for (int q = i; q < kn; q++)
						tree_parent[q] = j;
				else if (lv == TDV_NT - 1)   /* the dumped particles' node */
					for (int q = i + 1; q < kn; q++)
						tree_parent[q] = j;
			}
			cur = lowest - 1;
			if (cur < lo)
				break;
		}
		/* kn heads a chain whose parent (level dn-1) is in this chain when
		 * no particle between i and kn closes that level */
		if (kn <= r.last && runmin >= dn && cnt[kn] > 0) {
			const int jk = node_index(T, r.t, r.first, off, kn, r.last);
			const int jp = base + (dn - 1 - lo);
			up[jk] = jk - jp;
		}
		if (dn < runmin)
			runmin = dn;
		kprev = kk;
		k = kn;
	}
}

/* The chains whose scan covers more than TDV_LONG_CHAIN particles (the
 * roots and upper nodes of large top nodes), one wavefront per chain, taken
 * from a work queue.  The result is k_tree_nodes's, bit for bit: the lanes
 * load 64 particles at a time (coalesced), form m x, m y, m z and the
 * divergence level of each particle with the next, and set the parent links
 * of the chains that start inside this one from a running minimum of those
 * levels (a prefix minimum over the lanes); the FP64 sums then run over the
 * particles in PH order from LDS, every lane keeping the same sums, and the
 * cells that end at a particle (its level <= the deepest open level, found
 * by a ballot) take the sums at that particle. */
/* min of v and the value DPP control `ctrl` moves into this lane; lanes
 * that receive nothing (outside the row, or rows off in `rows`) see 99,
 * above every level */
template <int ctrl, int rows>
__device__ __forceinline__ int dpp_min(int v)
{
	const int o = __builtin_amdgcn_update_dpp(99, v, ctrl, rows, 0xf, false);
	return o < v ? o : v;
}

/* inclusive prefix minimum over the wavefront's 64 lanes in registers:
 * row_shr 1, 2, 4, 8 within each row of 16, then row_bcast 15 (rows 1 and
 * 3 take the last lane of the row before) and row_bcast 31 (rows 2 and 3
 * take lane 31) */
__device__ __forceinline__ int wave_prefix_min(int v)
{
	v = dpp_min<0x111, 0xf>(v);
	v = dpp_min<0x112, 0xf>(v);
	v = dpp_min<0x114, 0xf>(v);
	v = dpp_min<0x118, 0xf>(v);
	v = dpp_min<0x142, 0xa>(v);
	v = dpp_min<0x143, 0xc>(v);
	return v;
}

__global__ void __launch_bounds__(64) k_tree_nodes_long(
		const uint64_t *__restrict__ key, const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z,
		const double *__restrict__ m, TopNodes T, int vs,
		const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
		const int *__restrict__ lohi, double4 *__restrict__ na,
		int4 *__restrict__ nc, int *__restrict__ up,
		int *__restrict__ tree_parent, const int2 *__restrict__ long_list,
		const unsigned *__restrict__ long_count, unsigned long_cap,
		unsigned *__restrict__ queue)
{
	__shared__ double4 sh[64];
	__shared__ int shdn[64];
	const int lane = threadIdx.x;
	/* an overflow is flagged by k_tree_nodes; read no further than the list */
	if (*T.nnodes < 0)
		return;
	const unsigned nl = *long_count;
	const int nlong = (int)(nl < long_cap ? nl : long_cap);
	for (;;) {
		/* wave-uniform ticket, as in the walk's queue */
		const unsigned tk = atomicAdd(queue, lane == 0 ? 1u : 0u);
		const int item = __builtin_amdgcn_readfirstlane((int)tk);
		if (item >= nlong)
			break;
		const int2 le = long_list[item];
		const int i = le.x, e = le.y;
		const TopRange r = top_of(key, (size_t)i, T, vs);
		const int lo = lohi[i] & 0xFF, hi = (lohi[i] >> 8) & 0xFF;
		const int base = node_index(T, r.t, r.first, off, i, r.last);
		const Key2 ki = load_key(key, i);
		/* lane 4 q + c sums component c (m x, m y, m z, m): one add per
		 * particle advances all four sums, in PH order each */
		const int comp = lane & 3;
		double acc = 0;
		int cur = hi, runmin = 99;
		/* a block's particles and keys are loaded while the previous
		 * block is summed */
		double nm = 0, nx = 0, ny = 0, nz = 0;
		Key2 nk0 = { 0, 0 }, nk1 = { 0, 0 };
		{
			const int k = i + lane;
			if (k <= e) {
				nm = m[k];
				nx = x[k];
				ny = y[k];
				nz = z[k];
				nk0 = load_key(key, k);
				if (k + 1 <= r.last)
					nk1 = load_key(key, k + 1);
			}
		}
		for (int b0 = i; b0 <= e && cur >= lo; b0 += 64) {
			const int k = b0 + lane;
			const bool v = k <= e;
			const double cm = nm, cx = nx, cy = ny, cz = nz;
			const Key2 ck0 = nk0, ck1 = nk1;
			{
				const int kn = k + 64;
				if (kn <= e) {
					nm = m[kn];
					nx = x[kn];
					ny = y[kn];
					nz = z[kn];
					nk0 = load_key(key, kn);
					if (kn + 1 <= r.last)
						nk1 = load_key(key, kn + 1);
				}
			}
			double tx = 0, ty = 0, tz = 0, mk = 0;
			int dn = 99;
			if (v) {
				mk = cm;
				tx = cx * mk;
				ty = cy * mk;
				tz = cz * mk;
				dn = k + 1 <= r.last ? div_level(ck0, ck1) : -1;
			}
			/* runmin before particle k: min of the levels of i .. k-1 */
			const int incl = wave_prefix_min(dn);
			/* wave_shr 1: lane l takes lane l - 1's, lane 0 gets 99 */
			const int excl = __builtin_amdgcn_update_dpp(99, incl, 0x138, 0xf,
					0xf, false);
			const int rm = runmin < excl ? runmin : excl;
			/* not at e: the chain's last particle closes it before the
			 * single-lane scan looks for a chain starting after it */
			if (v && k < e && k + 1 <= r.last && rm >= dn && cnt[k + 1] > 0) {
				const int jk = node_index(T, r.t, r.first, off, k + 1, r.last);
				const int jp = base + (dn - 1 - lo);
				up[jk] = jk - jp;
			}
			const int bmin = __builtin_amdgcn_readlane(incl, 63);
			runmin = runmin < bmin ? runmin : bmin;
			sh[lane] = make_double4(tx, ty, tz, mk);
			shdn[lane] = dn;
			__syncthreads();
			const int nb = e - b0 + 1 < 64 ? e - b0 + 1 : 64;
			int q0 = 0;
			while (q0 < nb) {
				/* the next particle at which open cells end */
				const unsigned long long ends = __ballot(lane >= q0
						&& lane < nb && dn <= cur);
				const int qe = ends ? __ffsll((long long)ends) - 1 : nb - 1;
				const double *col = reinterpret_cast<const double *>(sh) + comp;
				int q = q0;
				for (; q + 8 <= qe + 1; q += 8) {
					double t[8];
#pragma unroll
					for (int u = 0; u < 8; u++)
						t[u] = col[4 * (q + u)];
#pragma unroll
					for (int u = 0; u < 8; u++)
						acc = acc + t[u];
				}
				for (; q <= qe; q++)
					acc = acc + col[4 * q];
				q0 = qe + 1;
				if (!ends)
					break;
				const double sx = __shfl(acc, 0, 64), sy = __shfl(acc, 1, 64),
						sz = __shfl(acc, 2, 64), sm = __shfl(acc, 3, 64);
				const int kn = b0 + qe + 1;
				const int dnq = shdn[qe];
				const int lowest = dnq > lo ? dnq : lo;
				const int next = node_index(T, r.t, r.first, off, kn, r.last);
				for (int lv = lowest + lane; lv <= cur; lv += 64) {
					const int j = base + (lv - lo);
					const bool leaf = lv == hi;
					int bits = lv | ((int)bitfield_triplet(ki, lv) << 6)
							| (1 << TDV_BIT_LOCAL);
					const bool root = j == T.target[r.t];
					if (root)
						bits |= 1 << TDV_BIT_TOP;
					const int dnext = leaf ? (-i - 1) : (next - j);
					na[j] = make_double4(sx / sm, sy / sm, sz / sm, sm);
					nc[j] = make_int4(dnext, kn - i, bits, i);
					if (root)
						up[j] = r.t;
					else if (lv > lo)
						up[j] = 1;
					if (!leaf && lv == TDV_NT - 1)   /* dumped particles */
						for (int q = i + 1; q < kn; q++)
							tree_parent[q] = j;
				}
				if (cur == hi)            /* the leaf: its particles */
					for (int q = i + lane; q < kn; q += 64)
						tree_parent[q] = base + (hi - lo);
				cur = lowest - 1;
				if (cur < lo)
					break;
			}
			__syncthreads();
		}
	}
}

/* The FP32 walk's build: the same nodes, links and parent pointers as
 * k_tree_nodes/_long, with no scan along the chains.  Every node of a chain
 * is the cell of its level that starts at the chain's head particle i, so
 * its end is a binary search over the sorted keys (the last particle whose
 * key shares the cell with key_i); the head's own parent (a level-(lo-1)
 * cell, owned by the particle that starts it) is found the same way
 * backwards.  The centres of mass come afterwards, bottom-up over the
 * levels (k_com_level): a sum in tree order instead of the reference's
 * insertion order, which the FP64 walk keeps bit for bit and the FP32 walk
 * does not need. */

/* the last particle in [a, b] whose key shares the level-lv cell with ki
 * (particle a does) */
__device__ __forceinline__ int cell_end(const uint64_t *key, Key2 ki, int a,
		int b, int lv)
{
	while (a < b) {
		const int mid = a + ((b - a + 1) >> 1);
		if (div_level(ki, load_key(key, mid)) > lv)
			a = mid;
		else
			b = mid - 1;
	}
	return a;
}

/* the first particle in [a, b] whose key shares the level-lv cell with kb
 * (particle b does) */
__device__ __forceinline__ int cell_start(const uint64_t *key, Key2 kb, int a,
		int b, int lv)
{
	while (a < b) {
		const int mid = a + ((b - a) >> 1);
		if (div_level(load_key(key, mid), kb) > lv)
			b = mid;
		else
			a = mid + 1;
	}
	return b;
}

__global__ void __launch_bounds__(256) k_tree_links(size_t n,
		const uint64_t *__restrict__ key, TopNodes T, int vs,
		const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
		const int *__restrict__ lohi, int4 *__restrict__ nc,
		int *__restrict__ up, int *__restrict__ tree_parent,
		unsigned *err)
{
	const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (gi >= n)
		return;
	const int i = (int)gi;
	if (cnt[i] == 0 || *T.nnodes < 0)
		return;
	const TopRange r = top_of(key, gi, T, vs);
	const int lo = lohi[i] & 0xFF, hi = (lohi[i] >> 8) & 0xFF;
	const int base = node_index(T, r.t, r.first, off, i, r.last);
	const Key2 ki = load_key(key, i);
	int e = r.last;
	for (int lv = lo; lv <= hi; lv++) {
		e = cell_end(key, ki, i, e, lv);
		const int kn = e + 1;
		const int j = base + (lv - lo);
		const bool leaf = lv == hi;
		int bits = lv | ((int)bitfield_triplet(ki, lv) << 6)
				| (1 << TDV_BIT_LOCAL);
		const bool root = j == T.target[r.t];
		if (root)
			bits |= 1 << TDV_BIT_TOP;
		const int next = node_index(T, r.t, r.first, off, kn, r.last);
		nc[j] = make_int4(leaf ? (-i - 1) : (next - j), kn - i, bits, i);
		if (root) {
			up[j] = r.t;
		} else if (lv > lo) {
			up[j] = 1;
		} else {
			/* the parent: the level-(lo-1) cell around i, a node of the
			 * chain its first particle heads */
			const int p0 = cell_start(key, ki, r.first, i, lo - 1);
			const int lp = lohi[p0] & 0xFF, hp = (lohi[p0] >> 8) & 0xFF;
			if (cnt[p0] == 0 || lp > lo - 1 || hp < lo - 1) {
				atomicOr(err, TDV_EBIT_RANGE);
				return;
			}
			const int jp = node_index(T, r.t, r.first, off, p0, r.last)
					+ (lo - 1 - lp);
			up[j] = j - jp;
		}
		if (leaf)
			for (int q = i; q < kn; q++)
				tree_parent[q] = j;
		else if (lv == TDV_NT - 1)           /* the dumped particles' node */
			for (int q = i + 1; q < kn; q++)
				tree_parent[q] = j;
	}
}

/* the particles of [f, f + count) that lane `sub` of eight takes: f + sub,
 * f + sub + 8, ... into m, m x, m y, m z */
__device__ __forceinline__ void com_add_strided(const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z,
		const double *__restrict__ m, int f, int count, int sub, double *s)
{
	for (int p = f + sub; p < f + count; p += 8) {
		const double mp = m[p];
		s[0] += mp;
		s[1] += mp * x[p];
		s[2] += mp * y[p];
		s[3] += mp * z[p];
	}
}

/* Centres of mass and masses of one level's nodes, eight lanes per node,
 * from the children (one level deeper, done by the launch before) and the
 * particles no child covers; a fixed grid strides over the level's list. */
__global__ void __launch_bounds__(256) k_com_level(int lv,
		const unsigned *__restrict__ lcnt, const int *__restrict__ loff,
		const int *__restrict__ list, const int4 *__restrict__ nc,
		const double *__restrict__ x, const double *__restrict__ y,
		const double *__restrict__ z, const double *__restrict__ m,
		double4 *__restrict__ na)
{
	const int count = (int)lcnt[lv];
	const int base = loff[lv];
	for (int t0 = (blockIdx.x * blockDim.x) >> 3; t0 < count;
			t0 += (gridDim.x * blockDim.x) >> 3) {
		const int t = t0 + (threadIdx.x >> 3), sub = threadIdx.x & 7;
		const bool on = t < count;
		double s[4] = { 0, 0, 0, 0 };
		int j = 0;
		if (on) {
			j = list[base + t];
			const int4 c = nc[j];
			if (c.x > 1) {
				int covered = c.w, k = 0;
				const int end = j + c.x;
				for (int ch = j + 1; ch < end; k++) {
					const int4 cc = nc[ch];
					com_add_strided(x, y, z, m, covered, cc.w - covered, sub, s);
					if ((k & 7) == sub) {
						const double4 a = na[ch];
						s[0] += a.w;
						s[1] += a.w * a.x;
						s[2] += a.w * a.y;
						s[3] += a.w * a.z;
					}
					covered = cc.w + cc.y;
					ch = cc.x > 0 ? ch + cc.x : ch + 1;
				}
				com_add_strided(x, y, z, m, covered, c.w + c.y - covered, sub,
						s);
			} else {
				com_add_strided(x, y, z, m, c.w, c.y, sub, s);
			}
		}
		for (int k = 0; k < 4; k++)
			for (int d = 4; d >= 1; d >>= 1)
				s[k] += __shfl_xor(s[k], d, 8);
		if (on && sub == 0)
			na[j] = make_double4(s[1] / s[0], s[2] / s[0], s[3] / s[0], s[0]);
	}
}

__global__ void k_tree_long_reset(unsigned *__restrict__ long_count,
		unsigned *__restrict__ queue)
{
	if (threadIdx.x == 0) {
		long_count[0] = 0;
		queue[0] = 0;
	}
}

__global__ void k_tree_small_top(const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z,
		const double *__restrict__ m, TopNodes T, int vs,
		const double4 *__restrict__ na, int *__restrict__ tree_parent)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= *T.ntop || *T.nnodes < 0)
		return;
	const int f = T.first[t], np = T.npart[t];
	double cx, cy, cz, mm;
	if (np > vs) {
		const double4 a = na[T.target[t]];
		cx = a.x;
		cy = a.y;
		cz = a.z;
		mm = a.w;
	} else {          /* finalise_subtree on a discarded subtree */
		double sx = 0, sy = 0, sz = 0;
		mm = 0;
		for (int q = f; q < f + np; q++) {
			sx = sx + x[q] * m[q];
			sy = sy + y[q] * m[q];
			sz = sz + z[q] * m[q];
			mm = mm + m[q];
			tree_parent[q] = -t - 1;
		}
		cx = sx / mm;
		cy = sy / mm;
		cz = sz / mm;
	}
	T.mass[t] = (float)mm;
	T.com[3 * t] = (float)cx;
	T.com[3 * t + 1] = (float)cy;
	T.com[3 * t + 2] = (float)cz;
	for (int d = 0; d < 3; d++)          /* no tree kicks yet */
		T.dp[3 * t + d] = 0.f;
	T.kicked[t] = 0;
}

__global__ void k_tree_terms(TopNodes T, int vs, double4 *__restrict__ na,
		double4 *__restrict__ nb, int4 *__restrict__ nc, int *__restrict__ up)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= *T.ntop || *T.nnodes < 0 || T.npart[t] <= vs)
		return;
	const int j = T.term[t];
	na[j] = make_double4(0, 0, 0, 1);   /* terminator: Mass = 1, TOP bit */
	nb[j] = make_double4(0, 0, 0, 0);
	nc[j] = make_int4(0, 0, 1 << TDV_BIT_TOP, -1);
	up[j] = 0;
}

__global__ void __launch_bounds__(256) k_tree_centres(size_t bound,
		const double *__restrict__ x, const double *__restrict__ y,
		const double *__restrict__ z, TopNodes T, double dom_size,
		const int4 *__restrict__ nc, const int *__restrict__ up,
		double4 *__restrict__ nb, size_t npart, unsigned *err)
{
	const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= bound || (long long)j >= (long long)*T.nnodes)
		return;
	const int ntop = *T.ntop;
	const int4 me = nc[j];
	if (me.w < 0)
		return;                          /* terminator */
	int path[TDV_NT + 2];
	int np = 0;
	int q = (int)j;
	int bits = me.z;
	while (!((bits >> TDV_BIT_TOP) & 1)) {
		const int u = up[q];
		if (np >= TDV_NT + 1 || u <= 0 || u > q) {
			atomicOr(err, TDV_EBIT_RANGE);   /* broken parent links */
			return;
		}
		path[np++] = q;
		q -= u;
		bits = nc[q].z;
	}
	const int t = up[q];                 /* the root's DUp is its top node */
	if (t < 0 || t >= ntop) {
		atomicOr(err, TDV_EBIT_RANGE);
		return;
	}
	double cx = T.pos[3 * t], cy = T.pos[3 * t + 1], cz = T.pos[3 * t + 2];
	for (int s = np - 1; s >= 0; s--) {
		const int nd = path[s];
		const int4 cn = nc[nd];
		const int lv = cn.z & 63;
		const int f = cn.w;
		if (f < 0 || (size_t)f >= npart) {
			atomicOr(err, TDV_EBIT_RANGE);
			return;
		}
		const double h = int_shift_size(dom_size, lv);
		cx = cx + (x[f] > cx ? 1 : -1) * h * 0.5;
		cy = cy + (y[f] > cy ? 1 : -1) * h * 0.5;
		cz = cz + (z[f] > cz ? 1 : -1) * h * 0.5;
	}
	const int lv = me.z & 63;
	nb[j] = make_double4(cx, cy, cz, dom_size / (double)(1ull << lv));
}

/* The build.  Eagerly it reads the node count back to size the node
 * arrays; in a captured step (c->graph_mode) the arrays keep the capacity
 * of the last eager build and every kernel takes the counts from device
 * memory (a count beyond the capacity is flagged, see k_tree_targets). */
int tdv_build_tree(tdv_context *c)
{
	const size_t n = c->n;
	const int vs = c->par.vector_size;
	const bool graph = c->graph_mode != 0;
	c->lvl_ok = 0;                       /* the FP32 walk's level lists */
	hipLaunchKernelGGL(k_fill_u32_val, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, c->minpack, 0xFFFFFFFFu);
	hipLaunchKernelGGL(k_tree_marks, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, c->key, c->top, vs, c->minpack, c->d_err);
	hipLaunchKernelGGL(k_tree_counts, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, c->key, c->top, vs, c->minpack,
			reinterpret_cast<uint32_t *>(c->cnt), c->lohi);
	TDV_TRY(tdv_scan_u32(c, reinterpret_cast<uint32_t *>(c->cnt), c->off, n));
	hipLaunchKernelGGL(k_tree_targets, dim3(1), dim3(64), 0, c->stream, c->top,
			vs, c->off, graph ? (long long)c->node_cap : (1ll << 62),
			c->d_err);
	TDV_HIP(c, hipGetLastError());
	size_t bound = c->node_cap;          /* nodes the kernels may touch */
	if (!graph) {
		int nn = 0;
		TDV_HIP(c, hipMemcpyAsync(&nn, c->top.nnodes, sizeof(int),
				hipMemcpyDeviceToHost, c->stream));
		TDV_HIP(c, hipStreamSynchronize(c->stream));
		TDV_TRY(tdv_check_err_flags(c));
		TDV_TRY(tdv_ensure_nodes(c, (size_t)nn));
		c->nnodes = (size_t)nn;
		bound = (size_t)nn;
	}
	if (bound > 0)
		hipLaunchKernelGGL(k_clear_nodes, dim3(tdv_blocks(bound, 256)),
				dim3(256), 0, c->stream, bound, c->top, c->node_c, c->node_dp);
	/* at most n / (TDV_LONG_CHAIN + 1) disjoint long cells per level */
	const size_t long_need = (n / (TDV_LONG_CHAIN + 1) + 1) * (TDV_NT + 2);
	if (c->long_cap < long_need) {
		if (graph)
			return tdv_fail(c, TDV_ERR_ARG, "capture: build lists missing");
		if (c->long_list)
			(void)hipFree(c->long_list);
		c->long_list = nullptr;
		c->long_cap = 0;
		TDV_HIP(c, hipMalloc((void **)&c->long_list, long_need * sizeof(int2)));
		c->long_cap = long_need;
	}
	const int tgrid = graph ? c->top_cap : c->ntop;
	if (c->par.precision == TDV_WALK_FP32) {
		/* links by binary searches, centres of mass bottom-up */
		hipLaunchKernelGGL(k_tree_links, dim3(tdv_blocks(n, 256)), dim3(256),
				0, c->stream, n, c->key, c->top, vs,
				reinterpret_cast<uint32_t *>(c->cnt), c->off, c->lohi,
				c->node_c, c->node_up, c->tree_parent, c->d_err);
		hipLaunchKernelGGL(k_tree_terms, dim3(tdv_blocks((size_t)tgrid, 64)),
				dim3(64), 0, c->stream, c->top, vs, c->node_a, c->node_b,
				c->node_c, c->node_up);
		if (bound > 0) {
			TDV_TRY(tdv_level_lists(c, bound));
			const unsigned grid = (unsigned)std::min(
					(size_t)tdv_blocks(8 * bound, 256), (size_t)2048);
			for (int lv = TDV_NT; lv >= 0; lv--)
				hipLaunchKernelGGL(k_com_level, dim3(grid), dim3(256), 0,
						c->stream, lv, c->q_arrived,
						reinterpret_cast<const int *>(c->q_arrived + 128),
						c->lvl_nodes, c->node_c, c->pos[0], c->pos[1],
						c->pos[2], c->mass, c->node_a);
		}
	} else {
		hipLaunchKernelGGL(k_tree_long_reset, dim3(1), dim3(64), 0,
				c->stream, c->long_count, c->long_count + 1);
		hipLaunchKernelGGL(k_tree_nodes, dim3(tdv_blocks(n, 256)), dim3(256),
				0, c->stream, n, c->key, c->pos[0], c->pos[1], c->pos[2],
				c->mass, c->top, vs, reinterpret_cast<uint32_t *>(c->cnt),
				c->off, c->lohi, c->node_a, c->node_c, c->node_up,
				c->tree_parent, c->long_list, (unsigned)c->long_cap,
				c->long_count, c->d_err);
		hipLaunchKernelGGL(k_tree_nodes_long, dim3(TDV_LONG_GRID), dim3(64),
				0, c->stream, c->key, c->pos[0], c->pos[1], c->pos[2],
				c->mass, c->top, vs, reinterpret_cast<uint32_t *>(c->cnt),
				c->off, c->lohi, c->node_a, c->node_c, c->node_up,
				c->tree_parent, c->long_list, c->long_count,
				(unsigned)c->long_cap, c->long_count + 1);
		hipLaunchKernelGGL(k_tree_terms, dim3(tdv_blocks((size_t)tgrid, 64)),
				dim3(64), 0, c->stream, c->top, vs, c->node_a, c->node_b,
				c->node_c, c->node_up);
	}
	hipLaunchKernelGGL(k_tree_small_top, dim3(tdv_blocks((size_t)tgrid, 64)),
			dim3(64), 0, c->stream, c->pos[0], c->pos[1], c->pos[2], c->mass,
			c->top, vs, c->node_a, c->tree_parent);
	if (bound > 0)
		hipLaunchKernelGGL(k_tree_centres, dim3(tdv_blocks(bound, 256)),
				dim3(256), 0, c->stream, bound, c->pos[0], c->pos[1],
				c->pos[2], c->top, c->dom_size, c->node_c, c->node_up,
				c->node_b, n, c->d_err);
	TDV_HIP(c, hipGetLastError());
	c->have_tree = 1;
	return 0;
}
