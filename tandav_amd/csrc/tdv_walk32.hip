/*
 * FP32 group walk.  One wavefront walks the tree for a PH-contiguous group
 * of 64 active particles (one per lane) with a single shared cursor: a node
 * is opened when any lane's criterion (the relative criterion of
 * Gravity/tree_accel.c:257 or the BH criterion of :339, plus the "inside"
 * test of :264-281) asks for it, otherwise every lane takes the node's
 * monopole.  Control flow is uniform across the wavefront, node records are
 * fetched with scalar loads, and every lane does useful work on every
 * interaction.  Particle bundles (the collapsed leaves, up to 64 particles)
 * are tested like nodes: the reference's walk takes every bundle particle
 * by particle once it reaches one (tree_accel.c:219-240), this walk takes a
 * bundle that no lane opens whole, by its multipoles, which in a
 * near-uniform box at the relative criterion replaces nine in ten
 * particle interactions (fp32_bundle_min).
 *
 * Separations come from 32-bit fixed-point coordinates (2^32 per box on
 * periodic runs, so the integer difference is already the nearest image;
 * 2^31 per domain otherwise): the subtraction is exact and only the
 * difference is rounded to FP32, so nearby particles do not lose digits to
 * cancellation.  Masses are pre-scaled so the kernel works in fixed-point
 * length units.  Lanes accumulate in FP32 and carry the partial sums into
 * FP64 accumulators every fp32_carry interactions: in a near-uniform box |a|
 * is a small residue of large terms, and the rounding of long FP32 partial
 * sums, not the tree, would set the error floor.
 *
 * Periodic runs add the Ewald correction in the same traversal: a node that
 * is taken as a monopole adds Newton + correction at its centre of mass
 * (the periodic force of a point mass); when a node is opened but every
 * lane's Ewald criterion (see ewald_ok) accepts its centre of mass, its
 * correction is added once there and its subtree then adds Newton terms
 * only.  Only bundles that straddle the half box or fail that criterion
 * fall back to per-particle corrections.  The correction comes from the
 * reference's 65^3 table (Gravity/periodic.c:35-113) in FP32, L2-resident.
 */
#include "tdv_internal.h"
#include "tdv_device.h"
#include <algorithm>

struct Walk32Params {
	int ntop, vs, periodic, use_bh;
	float theta_bh, alpha_rel;
	float eps2, inv_eps2, k_inner;   /* fixed units: eps^2, 1/eps^2,
	                                    1/(16 eps^3) */
	float eps_abs;                   /* |eps|, fixed units */
	float ewald_size;                /* correction at nodes <= this size */
	float ewald_acc;                 /* factor on the criterion for that */
	float half;                      /* half box in fixed units */
	float box_f;                     /* box in fixed units */
	float dom_f;                     /* domain size in fixed units */
	float u_scale;                   /* fixed units -> table grid units */
	float inv_box_f;                 /* 1 / box_f */
	float inv_s2;                    /* 1 / S^2: quadrupoles -> table units */
	float taylor2;                   /* S/6 per unit box -> (1/2) S per grid
	                                    unit in E's units */
	float inv_s2_box4;               /* 1 / (S^2 Box^4) */
	int carry_every;
	int quad;                        /* accepted nodes add their quadrupole */
	int crit3;                       /* open if M s^3 > r^5 |a| alpha / G */
	int bundle_min;                  /* bundles of >= this many particles
	                                    are tested like nodes */
};

/* The relative opening test of Gravity/tree_accel.c:257, M s^2 > r^4 crit
 * (crit = |a_prev| alpha / G in the kernel's units), or its next-order form
 * M s^3 > r^5 crit, which bounds the octupole error left once quadrupoles
 * are added.  Written as M s^2 (s/r) so no product leaves the FP32 range. */
__device__ __forceinline__ bool rel_open(float m, float s, float r2,
		float crit, const Walk32Params &w)
{
	const float lhs = w.crit3 ? m * s * s * (s * __frsqrt_rn(r2)) : m * s * s;
	return lhs > (r2 * crit) * r2;
}

/* ---- preparation: fixed-point particle / node records -------------------- */

/* fixed-point coordinate modulo 2^32: floor((x - o) S) through a signed
 * 64-bit integer (a position that drifted below the origin gives a
 * negative value, which the conversion to unsigned wraps, as the walk's
 * modular differences expect) */
__device__ __forceinline__ unsigned fixp(double x, double o, double S)
{
	const double u = floor((x - o) * S);
	return (unsigned)(long long)u;
}


/* The centre of the level-lv cell that holds fixed-point coordinate u, in
 * a frame of 2^bits units per domain (32 for the periodic box, 31
 * otherwise); a cell smaller than one unit is centred on u.  The walk takes
 * node centres from here, not from the reference's Tree_Node Pos, whose
 * Node_Size shift wraps below level 30 (Gravity/tree.h; kept bit for bit by
 * the FP64 build), so deep nodes' Pos can lie far from their particles. */
__device__ __forceinline__ unsigned cell_centre(unsigned u, int lv, int bits)
{
	const int sh = bits - lv;
	if (sh <= 0)
		return u;
	return ((u >> sh) << sh) + (1u << (sh - 1));
}

/* One node's (or bundle's) walk record, 64 bytes, fetched with one
 * scalar load:
 *   a = {CoM (fixed units), m'}
 *   b = {centre - CoM (fixed units, FP32 bits), size (fixed units)}
 *   c = {DNext (or -(first particle) - 1 for a bundle), Npart, Txx, Tyy}
 *   q = {Txy, Txz, Tyz, 0}
 * (T the scaled traceless quadrupole, zero when the walk takes monopoles
 * only); a lane takes the offset to the centre with one add per axis. */
struct __align__(64) Node32 {
	int4 a, b, c, q;
};

/* Every node's cell centre in fixed units, from its first particle (the
 * opening test's centre), when the records are made for a tree; a
 * substep's records keep them */
__global__ void k_prep32_centres(size_t bound, const int *__restrict__ nn,
		const int4 *__restrict__ nc, const double4 *__restrict__ pm,
		double o0, double o1, double o2, double S, int bits,
		int4 *__restrict__ ctr)
{
	const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= bound || (long long)j >= (long long)*nn)
		return;
	const int4 c = nc[j];
	if (c.w < 0) {                       /* terminator */
		ctr[j] = make_int4(0, 0, 0, 0);
		return;
	}
	const double4 q = pm[c.w];
	const int lv = c.z & 63;
	ctr[j] = make_int4((int)cell_centre(fixp(q.x, o0, S), lv, bits),
			(int)cell_centre(fixp(q.y, o1, S), lv, bits),
			(int)cell_centre(fixp(q.z, o2, S), lv, bits), 0);
}

/* the records from the tree (node_c, FP64 CoM and size) and the
 * quadrupoles; a32 keeps {CoM, m'} */
__global__ void k_prep32_nodes(size_t bound, const int *__restrict__ nn,
		const double4 *__restrict__ na,
		const double4 *__restrict__ nb, const int4 *__restrict__ nc,
		const float4 *__restrict__ q32, int quad, const int4 *__restrict__ ctr,
		double o0, double o1, double o2, double S, double mscale,
		int4 *__restrict__ a32, Node32 *__restrict__ rec)
{
	const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= bound || (long long)j >= (long long)*nn)
		return;
	const double4 a = na[j], b = nb[j];
	const int4 c = nc[j];
	const int4 A = make_int4((int)fixp(a.x, o0, S), (int)fixp(a.y, o1, S),
			(int)fixp(a.z, o2, S), __float_as_int((float)(a.w * mscale)));
	a32[j] = A;
	const float s = (float)(b.w * S);
	Node32 r;
	r.a = A;
	const int4 C = ctr[j];
	r.b = make_int4(
			__float_as_int((float)(int)((unsigned)C.x - (unsigned)A.x)),
			__float_as_int((float)(int)((unsigned)C.y - (unsigned)A.y)),
			__float_as_int((float)(int)((unsigned)C.z - (unsigned)A.z)),
			__float_as_int(s));
	float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
	if (quad) {
		q0 = q32[2 * j];
		q1 = q32[2 * j + 1];
	}
	r.c = make_int4(c.x, c.y, __float_as_int(q0.x), __float_as_int(q0.y));
	r.q = make_int4(__float_as_int(q0.z), __float_as_int(q0.w),
			__float_as_int(q1.x), 0);
	rec[j] = r;
}

__global__ void k_prep32_top(const TopNodes T, double o0, double o1,
		double o2, double S, double mscale, int4 *__restrict__ t32)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= *T.ntop)
		return;
	const double m = (double)T.mass[t] * mscale;
	t32[t] = make_int4((int)fixp(T.com[3 * t], o0, S),
			(int)fixp(T.com[3 * t + 1], o1, S), (int)fixp(T.com[3 * t + 2], o2, S),
			__float_as_int((float)m));
}

/* The walk's Ewald records, one per table node, 64 bytes: E, its first
 * derivatives J (FP32) and its second derivatives S / 6 (FP16 pairs):
 *   {Ex, Ey, Ez, Jxx}, {Jyy, Jzz, Jxy, Jxz},
 *   {Jyz, (Sxxx, Syyy), (Szzz, Sxxy), (Sxxz, Sxyy)}, {(Syyz, Sxzz), (Syzz, Sxyz), 0, 0}.
 * E = -F (the table's sign rule: Ewald_Correction returns -F(|d|) on the
 * positive octant) in units where m' E is an acceleration (m' = G m S^2);
 * J per table grid unit in the same units; S per unit box (it enters the
 * node's quadrupole term, and, rescaled, the expansion's second-order term,
 * where FP16's 5e-4 relative precision is far below the terms' own size). */
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float pack_h2(double a, double b)
{
	half2_t h;
	h.x = (_Float16)a;
	h.y = (_Float16)b;
	return __builtin_bit_cast(float, h);
}

__global__ void k_prep32_ewald_taylor(const double *__restrict__ raw,
		const double *__restrict__ s10, const double *__restrict__ j6,
		double e_scale, double j_scale, float4 *__restrict__ t)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= TDV_NE3)
		return;
	double S[10], J[6];
	for (int q = 0; q < 10; q++)
		S[q] = s10[(size_t)q * TDV_NE3 + i] / 6.0;
	for (int q = 0; q < 6; q++)
		J[q] = j6[(size_t)q * TDV_NE3 + i] * j_scale;
	t[4 * i] = make_float4((float)(-raw[i] * e_scale),
			(float)(-raw[(size_t)TDV_NE3 + i] * e_scale),
			(float)(-raw[2 * (size_t)TDV_NE3 + i] * e_scale), (float)J[0]);
	t[4 * i + 1] = make_float4((float)J[1], (float)J[2], (float)J[3],
			(float)J[4]);
	t[4 * i + 2] = make_float4((float)J[5], pack_h2(S[0], S[1]),
			pack_h2(S[2], S[3]), pack_h2(S[4], S[5]));
	t[4 * i + 3] = make_float4(pack_h2(S[6], S[7]), pack_h2(S[8], S[9]), 0.f,
			0.f);
}

/* Every particle's fixed-point record (position, mass') and, in one pass
 * (after the nodes' records), its offset from its bundle's centre of mass
 * in fixed units (an exact integer difference, rounded once to FP32) with
 * its mass', so a group taking the bundle in the CoM's frame adds one FP32
 * add per component.  (An offset from the bundle's cell centre, which a
 * tree update does not move, lost an order of magnitude in accuracy on the
 * Hernquist sphere: the centre lies farther from the particles.)  Particles
 * of small top nodes (no leaf) are not read from the offsets. */
__global__ void k_prep32_particles(size_t n, const double4 *__restrict__ pm,
		double o0, double o1, double o2, double S, double mscale,
		const int *__restrict__ leaf, const int4 *__restrict__ a32,
		int4 *__restrict__ p32, float4 *__restrict__ p32o)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const double4 q = pm[i];
	const int4 p = make_int4((int)fixp(q.x, o0, S), (int)fixp(q.y, o1, S),
			(int)fixp(q.z, o2, S), __float_as_int((float)(q.w * mscale)));
	p32[i] = p;
	const int j = leaf[i];
	float4 o = make_float4(0.f, 0.f, 0.f, __int_as_float(p.w));
	if (j >= 0) {
		const int4 c = a32[j];
		o.x = (float)(int)((unsigned)p.x - (unsigned)c.x);
		o.y = (float)(int)((unsigned)p.y - (unsigned)c.y);
		o.z = (float)(int)((unsigned)p.z - (unsigned)c.z);
	}
	p32o[i] = o;
}

/* Quadrupoles of every node, bundle and top node, for the walk: the
 * traceless moment T = 3 S - tr(S) I of the second moment S = sum m y y^T,
 * y = x - CoM, in FP64 about the node's FP64 centre of mass (a top node's
 * about its float centre of mass, the one the walk uses), stored scaled (G
 * S^2 masses, fixed lengths) as {Txx, Tyy, Txy, Txz}, {Tyz, 0, 0, 0}; Tzz =
 * -Txx - Tyy.
 *
 * Bottom-up in O(nodes + particles), one launch per tree level from the
 * deepest: every node covers the contiguous particle range [first, first +
 * Npart) and its children (one level deeper: the nodes from j + 1 below
 * j + DNext, siblings linked by DNext, bundles by + 1) cover sub-ranges of
 * it in PH order.  A node without child nodes takes S from its particles;
 * any other node S = sum_c (S_c + m_c d_c d_c^T), d_c = CoM_c - CoM, plus
 * the particles no child covers. */
__device__ __forceinline__ void moments_add_particles(
		const double4 *__restrict__ pm, int first, int count, double cx,
		double cy, double cz, double *m6)
{
	for (int p = first; p < first + count; p++) {
		const double4 q = pm[p];
		const double dx = q.x - cx, dy = q.y - cy, dz = q.z - cz;
		const double mx = q.w * dx, my = q.w * dy;
		m6[0] = fma(mx, dx, m6[0]);
		m6[1] = fma(my, dy, m6[1]);
		m6[2] = fma(q.w * dz, dz, m6[2]);
		m6[3] = fma(mx, dy, m6[3]);
		m6[4] = fma(mx, dz, m6[4]);
		m6[5] = fma(my, dz, m6[5]);
	}
}

__device__ __forceinline__ void moments_store(const double *m6, double scale,
		float4 *__restrict__ out)
{
	const double tr = m6[0] + m6[1] + m6[2];
	out[0] = make_float4((float)(scale * (3 * m6[0] - tr)),
			(float)(scale * (3 * m6[1] - tr)), (float)(scale * 3 * m6[3]),
			(float)(scale * 3 * m6[4]));
	out[1] = make_float4((float)(scale * 3 * m6[5]), 0.f, 0.f, 0.f);
}

/* the next sibling of child ch (DNext for nodes, + 1 for bundles) */
__device__ __forceinline__ int next_sibling(const int4 *__restrict__ nc,
		int ch)
{
	const int d = nc[ch].x;
	return d > 0 ? ch + d : ch + 1;
}

/* lane `sub` of a node's eight takes particles first + sub, first + sub + 8,
 * ... of the range */
__device__ __forceinline__ void moments_add_strided(
		const double4 *__restrict__ pm, int first, int count, int sub,
		double cx, double cy, double cz, double *m6)
{
	for (int p = first + sub; p < first + count; p += 8) {
		const double4 q = pm[p];
		const double dx = q.x - cx, dy = q.y - cy, dz = q.z - cz;
		const double mx = q.w * dx, my = q.w * dy;
		m6[0] = fma(mx, dx, m6[0]);
		m6[1] = fma(my, dy, m6[1]);
		m6[2] = fma(q.w * dz, dz, m6[2]);
		m6[3] = fma(mx, dy, m6[3]);
		m6[4] = fma(mx, dz, m6[4]);
		m6[5] = fma(my, dz, m6[5]);
	}
}

/* Eight lanes per node: each takes every eighth particle of the ranges no
 * child covers and every eighth child term, then the eight partial sums
 * meet in a fixed butterfly (the same order on every run; every lane ends
 * with the same sums, addition being commutative). */
__global__ void __launch_bounds__(256) k_moments_level(int lv,
		const unsigned *__restrict__ lcnt, const int *__restrict__ loff,
		const int *__restrict__ list, const int4 *__restrict__ nc,
		const double4 *__restrict__ na, const double4 *__restrict__ pm,
		double *__restrict__ s6, double scale, float4 *__restrict__ q)
{
	/* the level's nodes, count and list offset on the device; a fixed grid
	 * strides over them */
	const int count = (int)lcnt[lv];
	const int base = loff[lv];
	for (int t0 = (blockIdx.x * blockDim.x) >> 3; t0 < count;
			t0 += (gridDim.x * blockDim.x) >> 3) {
	const int t = t0 + (threadIdx.x >> 3), sub = threadIdx.x & 7;
	const bool on = t < count;          /* the same for a node's 8 lanes */
	double m6[6] = { 0, 0, 0, 0, 0, 0 };
	int j = 0;
	if (on) {
		j = list[base + t];
		const int4 c = nc[j];
		const double4 a = na[j];
		if (c.x > 1) {
			int covered = c.w, k = 0;
			const int end = j + c.x;
			for (int ch = j + 1; ch < end; ch = next_sibling(nc, ch), k++) {
				const int4 cc = nc[ch];
				moments_add_strided(pm, covered, cc.w - covered, sub, a.x, a.y,
						a.z, m6);
				if ((k & 7) == sub) {
					const double4 ca = na[ch];
					const double dx = ca.x - a.x, dy = ca.y - a.y,
							dz = ca.z - a.z;
					const double *sc = s6 + 6 * (size_t)ch;
					const double mx = ca.w * dx, my = ca.w * dy;
					m6[0] += sc[0] + mx * dx;
					m6[1] += sc[1] + my * dy;
					m6[2] += sc[2] + ca.w * dz * dz;
					m6[3] += sc[3] + mx * dy;
					m6[4] += sc[4] + mx * dz;
					m6[5] += sc[5] + my * dz;
				}
				covered = cc.w + cc.y;
			}
			moments_add_strided(pm, covered, c.w + c.y - covered, sub, a.x,
					a.y, a.z, m6);
		} else {
			moments_add_strided(pm, c.w, c.y, sub, a.x, a.y, a.z, m6);
		}
	}
	for (int k = 0; k < 6; k++)
		for (int d = 4; d >= 1; d >>= 1)
			m6[k] += __shfl_xor(m6[k], d, 8);
	if (on && sub == 0) {
		for (int k = 0; k < 6; k++)
			s6[6 * (size_t)j + k] = m6[k];
		moments_store(m6, scale, q + 2 * j);
	}
	}
}

/* exclusive offsets of the per-level counts, and the fill counters
 * cleared (one lane) */
__global__ void k_level_offsets(const unsigned *__restrict__ cnt,
		int *__restrict__ off, unsigned *__restrict__ fill)
{
	if (threadIdx.x != 0)
		return;
	int run = 0;
	for (int lv = 0; lv < 64; lv++) {
		off[lv] = run;
		run += (int)cnt[lv];
		fill[lv] = 0;
	}
}

/* per-level node lists (terminators left out); counts are gathered per
 * block in LDS first, so the device-wide counters take one atomic per level
 * and block */
__global__ void __launch_bounds__(256) k_level_count(int bound,
		const int *__restrict__ nnd, const int4 *__restrict__ nc,
		unsigned *__restrict__ cnt)
{
	__shared__ unsigned h[64];
	const int nn = min(bound, *nnd);
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (threadIdx.x < 64)
		h[threadIdx.x] = 0;
	__syncthreads();
	if (j < nn) {
		const int4 c = nc[j];
		if (c.w >= 0)
			atomicAdd(&h[c.z & 63], 1u);
	}
	__syncthreads();
	if (threadIdx.x < 64 && h[threadIdx.x])
		atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

__global__ void __launch_bounds__(256) k_level_fill(int bound,
		const int *__restrict__ nnd, const int4 *__restrict__ nc,
		const int *__restrict__ off, unsigned *__restrict__ fill,
		int *__restrict__ list)
{
	__shared__ unsigned h[64], base[64];
	const int nn = min(bound, *nnd);
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (threadIdx.x < 64)
		h[threadIdx.x] = 0;
	__syncthreads();
	int lv = -1;
	unsigned r = 0;
	if (j < nn) {
		const int4 c = nc[j];
		if (c.w >= 0) {
			lv = c.z & 63;
			r = atomicAdd(&h[lv], 1u);
		}
	}
	__syncthreads();
	if (threadIdx.x < 64)
		base[threadIdx.x] = h[threadIdx.x]
				? atomicAdd(&fill[threadIdx.x], h[threadIdx.x]) : 0u;
	__syncthreads();
	if (lv >= 0)
		list[off[lv] + (int)(base[lv] + r)] = j;
}

/* top nodes: from the subtree's root, shifted to the float centre of mass
 * (S about a point P = S_cm + M (c - P)(c - P)^T), or from the particles of
 * a top node without subtree */
__global__ void k_moments_top(const TopNodes T, int vs,
		const double4 *__restrict__ na, const double4 *__restrict__ pm,
		const double *__restrict__ s6, double scale, float4 *__restrict__ q)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= *T.ntop || *T.nnodes < 0)
		return;
	const double cx = T.com[3 * t], cy = T.com[3 * t + 1],
			cz = T.com[3 * t + 2];
	double m6[6] = { 0, 0, 0, 0, 0, 0 };
	if (T.npart[t] > vs) {
		const int r = T.target[t];
		const double4 a = na[r];
		const double dx = a.x - cx, dy = a.y - cy, dz = a.z - cz;
		const double mx = a.w * dx, my = a.w * dy;
		const double *sr = s6 + 6 * (size_t)r;
		m6[0] = sr[0] + mx * dx;
		m6[1] = sr[1] + my * dy;
		m6[2] = sr[2] + a.w * dz * dz;
		m6[3] = sr[3] + mx * dy;
		m6[4] = sr[4] + mx * dz;
		m6[5] = sr[5] + my * dz;
	} else {
		moments_add_particles(pm, T.first[t], T.npart[t], cx, cy, cz, m6);
	}
	moments_store(m6, scale, q + 2 * t);
}

/* ---- the walk ---------------------------------------------------------- */

struct Acc32 {
	float x, y, z;
	double X, Y, Z;
	int since;
	unsigned long long cost, ewald;  /* interactions, Ewald corrections */
	unsigned nodes, ewq;             /* of which multipole nodes, and
	                                    corrections with quadrupoles */
};

__device__ __forceinline__ void carry(Acc32 &a)
{
	a.X += (double)a.x;
	a.Y += (double)a.y;
	a.Z += (double)a.z;
	a.x = a.y = a.z = 0.f;
	a.since = 0;
}

/* after n interactions: carry when the partial sums have taken `every` */
__device__ __forceinline__ void tickn(Acc32 &a, int n, int every)
{
	a.since += n;
	if (a.since >= every)
		carry(a);
}

__device__ __forceinline__ void tick(Acc32 &a, int every)
{
	if (++a.since >= every)
		carry(a);
}

/* K1 kernel (Gravity/tree_accel.c:362-394) in fixed units; m = G m S^2 */
__device__ __forceinline__ void newton32(Acc32 &a, float m, float dx,
		float dy, float dz, const Walk32Params &w)
{
	/* every lane takes every interaction of the group's list, so the count
	 * is wave-uniform (it includes the particle itself, which adds nothing) */
	a.cost++;
	const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
	const float ri = __frsqrt_rn(r2);        /* inf at r = 0: not used */
	float fac = m * ri * (ri * ri);
	/* inside the softening length (rare; the particle itself, at r = 0,
	 * lands here too and adds fac * 0): only when some lane needs it */
	if (__any(r2 < w.eps2)) {
		const float u2 = r2 * w.inv_eps2;
		const float in = m * (175.f - u2 * (294.f - u2 * 135.f)) * w.k_inner;
		fac = r2 < w.eps2 ? in : fac;
	}
	a.x = fmaf(fac, dx, a.x);
	a.y = fmaf(fac, dy, a.y);
	a.z = fmaf(fac, dz, a.z);
}

/* Quadrupole term of a node seen at d (particle -> node CoM):
 * a = G [-T d / r^5 + 5/2 (d.T d) d / r^7], written with u = d / r so that
 * no intermediate leaves the FP32 range: a = r^-4 (5/2 (u.T u) u - T u).
 * Skipped inside the softening length, where the expansion does not hold. */
__device__ __forceinline__ void quad32(Acc32 &a, const float4 qa, float qyz,
		float dx, float dy, float dz, const Walk32Params &w)
{
	const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
	if (r2 < w.eps2 || r2 == 0.f)
		return;
	const float ri = __frsqrt_rn(r2);
	const float ux = dx * ri, uy = dy * ri, uz = dz * ri;
	const float txx = qa.x, tyy = qa.y, tzz = -(qa.x + qa.y), txy = qa.z,
			txz = qa.w, tyz = qyz;
	const float tux = fmaf(txx, ux, fmaf(txy, uy, txz * uz));
	const float tuy = fmaf(txy, ux, fmaf(tyy, uy, tyz * uz));
	const float tuz = fmaf(txz, ux, fmaf(tyz, uy, tzz * uz));
	const float c = 2.5f * fmaf(ux, tux, fmaf(uy, tuy, uz * tuz));
	const float ri2 = ri * ri, ri4 = ri2 * ri2;
	a.x = fmaf(ri4, fmaf(c, ux, -tux), a.x);
	a.y = fmaf(ri4, fmaf(c, uy, -tuy), a.y);
	a.z = fmaf(ri4, fmaf(c, uz, -tuz), a.z);
}

/* The Ewald table node nearest to a separation d (fixed units) on the
 * positive octant, the offset from it in grid units (|offset| <= 1/2 per
 * axis) and the reflection signs (the correction is odd in each axis,
 * sign(0) = +1; Gravity/periodic.c:35-113). */
struct EwaldNode {
	int b;                           /* node (i, j, k) */
	float du, dv, dq;
	float sx, sy, sz;
};

__device__ __forceinline__ EwaldNode ewald_node(float dx, float dy, float dz,
		const Walk32Params &w)
{
	EwaldNode c;
	c.sx = dx < 0.f ? -1.f : 1.f;
	c.sy = dy < 0.f ? -1.f : 1.f;
	c.sz = dz < 0.f ? -1.f : 1.f;
	const float u = fabsf(dx) * w.u_scale, v = fabsf(dy) * w.u_scale,
			q = fabsf(dz) * w.u_scale;
	const int i = min((int)(u + 0.5f), TDV_NE), j = min((int)(v + 0.5f), TDV_NE),
			k = min((int)(q + 0.5f), TDV_NE);
	c.du = u - (float)i;
	c.dv = v - (float)j;
	c.dq = q - (float)k;
	c.b = (i * TDV_NE1 + j) * TDV_NE1 + k;
	return c;
}

__device__ __forceinline__ float h2lo(float v)
{
	return (float)__builtin_bit_cast(half2_t, v).x;
}

__device__ __forceinline__ float h2hi(float v)
{
	return (float)__builtin_bit_cast(half2_t, v).y;
}

/* The correction on the positive octant from the nearest node to second
 * order, E + J d + (1/2) S d d (d the offset in grid units; S fully
 * symmetric, the gradient's second derivatives), and the node's S / 6. */
struct EwaldLocal {
	float ex, ey, ez;
	float Sxxx, Syyy, Szzz, Sxxy, Sxxz, Sxyy, Syyz, Sxzz, Syzz, Sxyz;
};

__device__ __forceinline__ EwaldLocal ewald_local(const float4 *__restrict__ t,
		const EwaldNode &c, const Walk32Params &w)
{
	const float4 *r = t + 4 * c.b;
	const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
	EwaldLocal e;
	e.Sxxx = h2lo(r2.y);
	e.Syyy = h2hi(r2.y);
	e.Szzz = h2lo(r2.z);
	e.Sxxy = h2hi(r2.z);
	e.Sxxz = h2lo(r2.w);
	e.Sxyy = h2hi(r2.w);
	e.Syyz = h2lo(r3.x);
	e.Sxzz = h2hi(r3.x);
	e.Syzz = h2lo(r3.y);
	e.Sxyz = h2hi(r3.y);
	const float Jxx = r0.w, Jyy = r1.x, Jzz = r1.y, Jxy = r1.z, Jxz = r1.w,
			Jyz = r2.x;
	const float uu = c.du * c.du, vv = c.dv * c.dv, qq = c.dq * c.dq;
	const float uv = 2.f * c.du * c.dv, uq = 2.f * c.du * c.dq,
			vq = 2.f * c.dv * c.dq;
	/* w.taylor2 = (1/2) 6 S/6 per unit box -> per grid unit, E's units */
	const float hx = fmaf(e.Sxxx, uu, fmaf(e.Sxyy, vv, fmaf(e.Sxzz, qq,
			fmaf(e.Sxxy, uv, fmaf(e.Sxxz, uq, e.Sxyz * vq)))));
	const float hy = fmaf(e.Sxxy, uu, fmaf(e.Syyy, vv, fmaf(e.Syzz, qq,
			fmaf(e.Sxyy, uv, fmaf(e.Sxyz, uq, e.Syyz * vq)))));
	const float hz = fmaf(e.Sxxz, uu, fmaf(e.Syyz, vv, fmaf(e.Szzz, qq,
			fmaf(e.Sxyz, uv, fmaf(e.Sxzz, uq, e.Syzz * vq)))));
	e.ex = fmaf(w.taylor2, hx, fmaf(Jxx, c.du, fmaf(Jxy, c.dv,
			fmaf(Jxz, c.dq, r0.x))));
	e.ey = fmaf(w.taylor2, hy, fmaf(Jxy, c.du, fmaf(Jyy, c.dv,
			fmaf(Jyz, c.dq, r0.y))));
	e.ez = fmaf(w.taylor2, hz, fmaf(Jxz, c.du, fmaf(Jyz, c.dv,
			fmaf(Jzz, c.dq, r0.z))));
	return e;
}

/* Ewald_Correction (Gravity/periodic.c:35-113) of a point mass m' */
__device__ __forceinline__ void ewald16(Acc32 &a, const float4 *__restrict__ t,
		float m, float dx, float dy, float dz, const Walk32Params &w)
{
	const EwaldNode c = ewald_node(dx, dy, dz, w);
	const EwaldLocal e = ewald_local(t, c, w);
	a.x = fmaf(m * c.sx, e.ex, a.x);
	a.y = fmaf(m * c.sy, e.ey, a.y);
	a.z = fmaf(m * c.sz, e.ez, a.z);
	a.ewald++;
}

/* The correction of a node of mass m' and traceless second moment T'
 * {Txx, Tyy, Txy, Txz}, Tyz about its centre of mass at d, to second order:
 * m' E(d) + Q:S(d) with Q:S = T:S/6 (S = d d E is symmetric in its three
 * indices and traceless in the last two, so the trace part of the second
 * moment adds nothing), S taken at the nearest node.  The reflection signs
 * (S_abc is odd in each axis that appears an odd number of times) are
 * folded into the coefficients. */
__device__ __forceinline__ void ewald16q(Acc32 &a, const float4 *__restrict__ t,
		float m, const float4 qa, float qyz, float dx, float dy, float dz,
		const Walk32Params &w)
{
	const EwaldNode c = ewald_node(dx, dy, dz, w);
	const EwaldLocal e = ewald_local(t, c, w);
	/* T' (masses G S^2, lengths S) to table units: / (S^4 Box^4), in two
	 * steps so no constant is subnormal */
	const float cxx = qa.x * w.inv_s2 * w.inv_s2_box4,
			cyy = qa.y * w.inv_s2 * w.inv_s2_box4;
	const float czz = -(cxx + cyy);
	const float t2 = 2.f * w.inv_s2 * w.inv_s2_box4;
	const float cxy = qa.z * t2 * (c.sx * c.sy), cxz = qa.w * t2 * (c.sx * c.sz),
			cyz = qyz * t2 * (c.sy * c.sz);
	float vx = m * e.ex, vy = m * e.ey, vz = m * e.ez;
	vx = fmaf(cxx, e.Sxxx, fmaf(cyy, e.Sxyy, fmaf(czz, e.Sxzz, vx)));
	vx = fmaf(cxy, e.Sxxy, fmaf(cxz, e.Sxxz, fmaf(cyz, e.Sxyz, vx)));
	vy = fmaf(cxx, e.Sxxy, fmaf(cyy, e.Syyy, fmaf(czz, e.Syzz, vy)));
	vy = fmaf(cxy, e.Sxyy, fmaf(cxz, e.Sxyz, fmaf(cyz, e.Syyz, vy)));
	vz = fmaf(cxx, e.Sxxz, fmaf(cyy, e.Syyz, fmaf(czz, e.Szzz, vz)));
	vz = fmaf(cxy, e.Sxyz, fmaf(cxz, e.Sxzz, fmaf(cyz, e.Syzz, vz)));
	a.x = fmaf(c.sx, vx, a.x);
	a.y = fmaf(c.sy, vy, a.y);
	a.z = fmaf(c.sz, vz, a.z);
	a.ewald++;
	a.ewq++;
}

__device__ __forceinline__ float fdiff(int a, int b)
{
	return (float)(int)((unsigned)a - (unsigned)b);
}

/* the node's straddle zone: some of its particles have another nearest
 * image than its centre (Gravity/tree_periodic.c:280-308) */
__device__ __forceinline__ bool straddles(float sx, float sy, float sz,
		float s, const Walk32Params &w)
{
	const float lim = 0.5f * (w.box_f - s);
	return fabsf(sx) > lim || fabsf(sy) > lim || fabsf(sz) > lim;
}

/* Can a node's Ewald correction be taken at its centre of mass (to second
 * order, ewald16q)?  What is left is the third-order term, M s^3 |E'''|,
 * with |E'''| <= K / box^5, K = 60 + 24 (box / d_img)^5: the correction's
 * own curvature grows about linearly from the origin (|E''| ~ 58 |d| /
 * box^4 measured on the table) and near the half-box faces like the Newton
 * force of the nearest other image, at d_img >= box - max |d_k|.  It must
 * stay below ewald_accuracy times the relative criterion's bound; the BH
 * walk of the prepare step only caps the size.  Nodes whose subtree then
 * takes Newton terms per lane must not straddle the half-box planes (their
 * particles must share the centre's periodic image); bundles take their
 * Newton terms in the centre's frame and may straddle. */
__device__ __forceinline__ bool ewald_ok(float m, float s, float rx, float ry,
		float rz, float sx, float sy, float sz, float crit, bool may_straddle,
		const Walk32Params &w)
{
	if (s > w.ewald_size)
		return false;
	if (!may_straddle && straddles(sx, sy, sz, s, w))
		return false;
	if (w.use_bh)
		return true;
	const float dimg = 1.f - fmaxf(fabsf(rx), fmaxf(fabsf(ry), fabsf(rz)))
			* w.inv_box_f;
	const float di2 = 1.f / (dimg * dimg);
	const float K = fmaf(24.f * di2 * di2, 1.f / dimg, 60.f);
	const float sb = s * w.inv_box_f;
	return m * (sb * sb * sb) * K * (w.inv_box_f * w.inv_box_f)
			<= crit * w.ewald_acc;
}

/* The walk's decision at an internal node (tree_accel.c:249-262): the BH
 * criterion, or the relative one or the particle within 0.6 s of the
 * node's centre; (rx, ry, rz) to the CoM, (sx, sy, sz) to the centre.  A
 * node whose particles may lie inside the softening length (r < |eps| + s)
 * is opened too: there the force is not Newtonian and the node's multipole
 * expansion does not hold. */
__device__ __forceinline__ bool node_opens(float m, float s, float r2,
		float sx, float sy, float sz, float crit, const Walk32Params &w)
{
	const float es = w.eps_abs + s;
	if (r2 < es * es)
		return true;
	if (w.use_bh)
		return s * s > r2 * w.theta_bh;
	return rel_open(m, s, r2, crit, w) || (fabsf(sx) < 0.6f * s
			&& fabsf(sy) < 0.6f * s && fabsf(sz) < 0.6f * s);
}

/* node_opens given the node's inverse distance ri (the same decisions) */
__device__ __forceinline__ bool node_opens_ri(float m, float s, float r2,
		float ri, float sx, float sy, float sz, float crit,
		const Walk32Params &w)
{
	const float es = w.eps_abs + s;
	if (r2 < es * es)
		return true;
	if (w.use_bh)
		return s * s > r2 * w.theta_bh;
	const float ms2 = m * s * s;
	const float lhs = w.crit3 ? ms2 * (s * ri) : ms2;
	const float s6 = 0.6f * s;
	return lhs > (r2 * crit) * r2 || (fabsf(sx) < s6 && fabsf(sy) < s6
			&& fabsf(sz) < s6);
}

/* An accepted node (r >= |eps| + s, outside the softening length): its
 * monopole and, with quadrupoles, T's term in one pass,
 * a = r^-2 (m' + (5/2) (u.T u) r^-2) u - r^-4 T u, u = d / r. */
__device__ __forceinline__ void node32(Acc32 &a, float m, float4 qa,
		float qyz, float dx, float dy, float dz, float ri,
		const Walk32Params &w)
{
	a.cost++;
	a.nodes++;
	const float ri2 = ri * ri;
	if (!w.quad) {
		const float fac = m * ri * ri2;
		a.x = fmaf(fac, dx, a.x);
		a.y = fmaf(fac, dy, a.y);
		a.z = fmaf(fac, dz, a.z);
		return;
	}
	const float ux = dx * ri, uy = dy * ri, uz = dz * ri;
	const float txx = qa.x, tyy = qa.y, tzz = -(qa.x + qa.y), txy = qa.z,
			txz = qa.w, tyz = qyz;
	const float tux = fmaf(txx, ux, fmaf(txy, uy, txz * uz));
	const float tuy = fmaf(txy, ux, fmaf(tyy, uy, tyz * uz));
	const float tuz = fmaf(txz, ux, fmaf(tyz, uy, tzz * uz));
	const float c = 2.5f * fmaf(ux, tux, fmaf(uy, tuy, uz * tuz));
	const float ri4 = ri2 * ri2;
	const float g = fmaf(c, ri4, m * ri2);
	a.x = fmaf(-ri4, tux, fmaf(g, ux, a.x));
	a.y = fmaf(-ri4, tuy, fmaf(g, uy, a.y));
	a.z = fmaf(-ri4, tuz, fmaf(g, uz, a.z));
}

/* Whether a lane's particle wants top node j walked (its subtree) rather
 * than taken whole, and the separation; the walk and k_walk32_want
 * evaluate it alike. */
__device__ __forceinline__ bool top_want(const int4 tc, float ts,
		const int4 me, float crit, const Walk32Params &w, float &rx,
		float &ry, float &rz)
{
	const float tm = __int_as_float(tc.w);
	rx = fdiff(tc.x, me.x);
	ry = fdiff(tc.y, me.y);
	rz = fdiff(tc.z, me.z);
	const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
	const bool want = w.use_bh ? (ts * ts > r2 * w.theta_bh)
			: rel_open(tm, ts, r2, crit, w);
	/* the top node holds the particle (or is close): walk it */
	return want || r2 < 3.f * ts * ts;
}

/* One wave walks `group` (64 active particles, one per lane) over part
 * `part` of `nparts`: every subtree the group walks is cut into nparts
 * contiguous node ranges and this wave takes range `part` of each, so
 * several waves share a group whose walk is long (few active groups).  The
 * walk visits nodes in increasing index order, so the walk of a range
 * starts at the first node >= its start that the whole walk would visit:
 * the descent below replays the whole walk's decisions on the range
 * start's ancestors (and their Ewald state) without interacting.  Part 0
 * also takes the top-node interactions.  With nparts == 1 the results go
 * to the particles; otherwise to partial[part * nslot + slot], summed in
 * part order by k_walk32_parts_sum. */
__device__ __forceinline__ void walk32_body(int group, int part, int nparts,
		double4 *__restrict__ partial, int nslot, int nact,
		const int *__restrict__ active, const int4 *__restrict__ p32,
		const double *__restrict__ last_acc, TopNodes T,
		const Node32 *__restrict__ rec, const float4 *__restrict__ tab,
		Walk32Params w, double *__restrict__ ax, double *__restrict__ ay,
		double *__restrict__ az, double *__restrict__ gx,
		double *__restrict__ gy, double *__restrict__ gz,
		double *__restrict__ cost, unsigned long long *counter,
		const int4 *__restrict__ top32, const float4 *__restrict__ tq32,
		const float4 *__restrict__ p32o,
		const unsigned long long *__restrict__ wmask)
{
	if (*T.nnodes < 0)                   /* the build overflowed: no walk */
		return;
	const int ntop = *T.ntop;
	const int gid = group * 64 + (int)threadIdx.x;
	const bool live = gid < nact;
	const int ip = live ? active[gid] : active[nact - 1];
	const int4 me = p32[ip];
	/* relative criterion: open if M s^2 > r^4 |a| alpha / G; with masses
	 * scaled by G S^2 and lengths in fixed units this is m s^2 > r^4 crit */
	const float crit = (float)(last_acc[ip] * (double)w.alpha_rel);
	Acc32 a = { 0.f, 0.f, 0.f, 0.0, 0.0, 0.0, 0, 0ull, 0ull, 0u, 0u };
	const int every = w.carry_every;
	const float inv_parts = 1.f / (float)nparts;

	/* a split walk's parts other than 0 walk only the top nodes the group
	 * wants walked, which k_walk32_want found once for the group */
	const unsigned long long *wm = part != 0 && wmask
			&& ntop <= 64 * TDV_WANT_WORDS
			? wmask + (size_t)group * TDV_WANT_WORDS : nullptr;
	for (int j = 0; j < ntop; j++) {
		if (wm) {
			const unsigned long long bits = wm[j >> 6] >> (j & 63);
			if (!bits) {
				j |= 63;                 /* the next word */
				continue;
			}
			j += (int)__builtin_ctzll(bits);
			if (j >= ntop)
				break;
			if (T.npart[j] <= w.vs)      /* part 0's */
				continue;
		} else {
		const int4 tc = top32[j];
		const float tm = __int_as_float(tc.w);
		/* dom_f / 2^level, exact: one ldexp instead of a division */
		const float ts = ldexpf(w.dom_f, -T.level[j]);
		float rx, ry, rz;
		if (!__any(top_want(tc, ts, me, crit, w, rx, ry, rz))) {
			if (part != 0)
				continue;
			newton32(a, tm, rx, ry, rz, w);
			if (w.quad) {
				const float4 q0 = tq32[2 * j];
				const float qyz = tq32[2 * j + 1].x;
				quad32(a, q0, qyz, rx, ry, rz, w);
				if (w.periodic)
					ewald16q(a, tab, tm, q0, qyz, rx, ry, rz, w);
			} else if (w.periodic) {
				ewald16(a, tab, tm, rx, ry, rz, w);
			}
			tick(a, every);
			continue;
		}
		if (T.npart[j] <= w.vs) {
			if (part != 0)
				continue;
			const int f = T.first[j], l = f + T.npart[j];
			for (int q = f; q < l; q++) {
				const int4 pq = p32[q];
				const float dx = fdiff(pq.x, me.x), dy = fdiff(pq.y, me.y),
						dz = fdiff(pq.z, me.z);
				const float m = __int_as_float(pq.w);
				newton32(a, m, dx, dy, dz, w);
				if (w.periodic)           /* E(0) = 0: the self term adds 0 */
					ewald16(a, tab, m, dx, dy, dz, w);
				tick(a, every);
			}
			continue;
		}
		}
		const int start = T.target[j];
		int lo = start, hi = T.term[j];
		if (nparts > 1) {
			/* cut at floor(len p / nparts) in float: monotone in p, 0 at
			 * p = 0 and below len for p < nparts, so every part computes
			 * the same cuts and the ranges tile the subtree (a 64-bit
			 * integer division here cost ~100 scalar instructions per cut,
			 * twice per top node and part) */
			const int len = hi - start;
			lo = start + (int)((float)len * ((float)part * inv_parts));
			if (part + 1 < nparts)
				hi = start + (int)((float)len * ((float)(part + 1) * inv_parts));
			if (lo >= hi)
				continue;
		}
		int node = start;
		int ew_until = w.periodic ? -1 : 0x7FFFFFFF;  /* correction done below */
		while (node < lo) {
			/* internal nodes have DNext > 0, the root's reaching the
			 * terminator (tree_build.c:466-491) */
			const Node32 R = rec[node];
			const int4 c = R.c;
			if (c.x < 0) {
				node++;
				continue;
			}
			if (c.x == 0) {                  /* only the terminator */
				node = hi;
				break;
			}
			const int skip = node + c.x;
			if (skip <= lo) {            /* left before lo, open or not */
				node = skip;
				continue;
			}
			const int4 A = R.a;
			const int4 B = R.b;
			const float m = __int_as_float(A.w), s = __int_as_float(B.w);
			const float rx = fdiff(A.x, me.x), ry = fdiff(A.y, me.y),
					rz = fdiff(A.z, me.z);
			const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
			const float ri = __frsqrt_rn(r2);
			const float sx = rx + __int_as_float(B.x),
					sy = ry + __int_as_float(B.y),
					sz = rz + __int_as_float(B.z);
			if (!__any(node_opens_ri(m, s, r2, ri, sx, sy, sz, crit, w))) {
				node = skip;             /* taken whole by an earlier part */
				break;
			}
			if (node >= ew_until && __all(ewald_ok(m, s, rx, ry, rz, sx, sy,
					sz, crit, false, w)))
				ew_until = skip;
			node++;
		}
		for (;;) {
			if (node >= hi)
				break;
			const Node32 R = rec[node];
			const int4 c = R.c;
			if (c.x == 0 && node != start)
				break;
			const bool pending = node >= ew_until;
			/* the quadrupole {Txx, Tyy, Txy, Txz}, Tyz */
			const float4 q0 = make_float4(__int_as_float(c.z),
					__int_as_float(c.w), __int_as_float(R.q.x),
					__int_as_float(R.q.y));
			const float qyz = __int_as_float(R.q.z);
			const int4 A = R.a;
			const int4 B = R.b;
			if (c.x < 0) {
				const int f = -c.x - 1, l = f + c.y;
				if (c.y >= w.bundle_min) {
					/* a bundle that no lane's criterion opens is taken
					 * whole, like a node: its multipoles at its centre of
					 * mass instead of c.y particle interactions */
					const float m = __int_as_float(A.w),
							s = __int_as_float(B.w);
					const float rx = fdiff(A.x, me.x), ry = fdiff(A.y, me.y),
							rz = fdiff(A.z, me.z);
					const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
					const float ri = __frsqrt_rn(r2);
					if (!__any(node_opens_ri(m, s, r2, ri,
							rx + __int_as_float(B.x), ry + __int_as_float(B.y),
							rz + __int_as_float(B.z), crit, w))) {
						node32(a, m, q0, qyz, rx, ry, rz, ri, w);
						if (pending) {
							if (w.quad)
								ewald16q(a, tab, m, q0, qyz, rx, ry, rz, w);
							else
								ewald16(a, tab, m, rx, ry, rz, w);
						}
						tick(a, every);
						node++;
						continue;
					}
				}
				/* Newton in the CoM's periodic frame unless the correction
				 * is taken per particle: a correction taken higher up came
				 * from a node that does not straddle the half-box planes,
				 * so its particles share the CoM's image for every lane */
				bool com_frame = !pending;
				if (pending) {
					/* the bundle's correction at its CoM to second order when
					 * every lane allows it, else per particle */
					const float s = __int_as_float(B.w),
							m = __int_as_float(A.w);
					const float rx = fdiff(A.x, me.x), ry = fdiff(A.y, me.y),
							rz = fdiff(A.z, me.z);
					const bool ok = ewald_ok(m, s, rx, ry, rz,
							rx + __int_as_float(B.x), ry + __int_as_float(B.y),
							rz + __int_as_float(B.z), crit, true, w);
					com_frame = __all(ok);
					if (com_frame) {
						if (w.quad)
							ewald16q(a, tab, m, q0, qyz, rx, ry, rz, w);
						else
							ewald16(a, tab, m, rx, ry, rz, w);
					}
				}
				if (com_frame) {
					/* Newton in the CoM's periodic frame, consistent with the
					 * correction taken there: d = (A - me) + (q - A) */
					const float cx = fdiff(A.x, me.x), cy = fdiff(A.y, me.y),
							cz = fdiff(A.z, me.z);
					/* four particles per wide scalar load, the next four
					 * loaded while these are taken (the load's latency
					 * hides behind the arithmetic) */
					int q = f;
					if (q + 4 <= l) {
						float4 n0 = p32o[q], n1 = p32o[q + 1],
								n2 = p32o[q + 2], n3 = p32o[q + 3];
						for (;;) {
							const float4 p0 = n0, p1 = n1, p2 = n2, p3 = n3;
							const bool more = q + 8 <= l;
							if (more) {
								n0 = p32o[q + 4];
								n1 = p32o[q + 5];
								n2 = p32o[q + 6];
								n3 = p32o[q + 7];
							}
							newton32(a, p0.w, cx + p0.x, cy + p0.y, cz + p0.z,
									w);
							newton32(a, p1.w, cx + p1.x, cy + p1.y, cz + p1.z,
									w);
							newton32(a, p2.w, cx + p2.x, cy + p2.y, cz + p2.z,
									w);
							newton32(a, p3.w, cx + p3.x, cy + p3.y, cz + p3.z,
									w);
							tickn(a, 4, every);
							q += 4;
							if (!more)
								break;
						}
					}
					for (; q < l; q++) {
						const float4 p0 = p32o[q];
						newton32(a, p0.w, cx + p0.x, cy + p0.y, cz + p0.z, w);
						tick(a, every);
					}
				} else {
					int4 pn = p32[f];
					for (int q = f; q < l; q++) {
						const int4 pq = pn;
						if (q + 1 < l)
							pn = p32[q + 1];
						const float dx = fdiff(pq.x, me.x),
								dy = fdiff(pq.y, me.y), dz = fdiff(pq.z, me.z);
						const float m = __int_as_float(pq.w);
						newton32(a, m, dx, dy, dz, w);
						if (pending)          /* E(0) = 0 for the self term */
							ewald16(a, tab, m, dx, dy, dz, w);
						tick(a, every);
					}
				}
				node++;
				continue;
			}
			const float m = __int_as_float(A.w), s = __int_as_float(B.w);
			const float rx = fdiff(A.x, me.x), ry = fdiff(A.y, me.y),
					rz = fdiff(A.z, me.z);
			const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
			const float ri = __frsqrt_rn(r2);
			const float sx = rx + __int_as_float(B.x),
					sy = ry + __int_as_float(B.y),
					sz = rz + __int_as_float(B.z);
			if (!__any(node_opens_ri(m, s, r2, ri, sx, sy, sz, crit, w))) {
				node32(a, m, q0, qyz, rx, ry, rz, ri, w);
				if (pending) {
					if (w.quad)
						ewald16q(a, tab, m, q0, qyz, rx, ry, rz, w);
					else
						ewald16(a, tab, m, rx, ry, rz, w);
				}
				tick(a, every);
				node += c.x;
				continue;
			}
			if (pending && __all(ewald_ok(m, s, rx, ry, rz, sx, sy, sz, crit,
					false, w))) {
				if (w.quad)
					ewald16q(a, tab, m, q0, qyz, rx, ry, rz, w);
				else
					ewald16(a, tab, m, rx, ry, rz, w);
				ew_until = node + c.x;
			}
			node++;
		}
	}
	carry(a);
	if (nparts > 1) {
		if (live)
			partial[(size_t)part * nslot + gid] = make_double4(a.X, a.Y, a.Z,
					(double)(a.cost + a.ewald));
	} else if (live) {
		ax[ip] = ax[ip] + a.X;
		ay[ip] = ay[ip] + a.Y;
		az[ip] = az[ip] + a.Z;
		gx[ip] = a.X;
		gy[ip] = a.Y;
		gz[ip] = a.Z;
		cost[ip] = (double)(a.cost + a.ewald);
	}
	unsigned long long v = live ? a.cost : 0ull, e = live ? a.ewald : 0ull,
			nd = live ? a.nodes : 0ull, eq = live ? a.ewq : 0ull;
	for (int d = 32; d >= 1; d >>= 1) {
		v += __shfl_xor(v, d, 64);
		e += __shfl_xor(e, d, 64);
		nd += __shfl_xor(nd, d, 64);
		eq += __shfl_xor(eq, d, 64);
	}
	if (threadIdx.x == 0) {
		if (v)
			atomicAdd(counter, v);
		if (e)
			atomicAdd(counter + 1, e);
		if (nd)
			atomicAdd(counter + 2, nd);
		if (eq)
			atomicAdd(counter + 3, eq);
	}
}

/* The walk with a register budget for W waves per SIMD (0: the
 * compiler's own); the autotuner picks W per workload. */
#define TDV_WALK32_ARGS int nact, const int *__restrict__ active, \
		const int4 *__restrict__ p32, const double *__restrict__ last_acc, \
		TopNodes T, const Node32 *__restrict__ rec, \
		const float4 *__restrict__ tab, \
		Walk32Params w, double *__restrict__ ax, double *__restrict__ ay, \
		double *__restrict__ az, double *__restrict__ gx, \
		double *__restrict__ gy, double *__restrict__ gz, \
		double *__restrict__ cost, unsigned long long *counter, \
		const int4 *__restrict__ top32, const float4 *__restrict__ tq32, \
		const float4 *__restrict__ p32o, const uint32_t *__restrict__ gorder
#define TDV_WALK32_CALL_PARTW(g, part, nparts, partial, nslot, wmask) \
		walk32_body(g, part, nparts, partial, nslot, nact, active, p32, \
		last_acc, T, rec, tab, w, ax, ay, az, gx, gy, gz, cost, \
		counter, top32, tq32, p32o, wmask)
#define TDV_WALK32_CALL_PART(g, part, nparts, partial, nslot) \
		TDV_WALK32_CALL_PARTW(g, part, nparts, partial, nslot, nullptr)
#define TDV_WALK32_CALL(g) TDV_WALK32_CALL_PART(g, 0, 1, nullptr, 0)
/* the i-th group to walk: by descending expected cost when ordered */
#define TDV_GROUP(i) (gorder ? (int)gorder[i] : (i))

__global__ void __launch_bounds__(64) k_walk32(TDV_WALK32_ARGS)
{
	TDV_WALK32_CALL(TDV_GROUP((int)blockIdx.x));
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6)))
k_walk32_w6(TDV_WALK32_ARGS)
{
	TDV_WALK32_CALL(TDV_GROUP((int)blockIdx.x));
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
k_walk32_w8(TDV_WALK32_ARGS)
{
	TDV_WALK32_CALL(TDV_GROUP((int)blockIdx.x));
}

/* Parts per group: 1 when the active groups alone fill the chip, else
 * enough that groups x parts reaches the item budget, at most
 * TDV_SPLIT_AUTO_MAX (split 0) or `split` (>= 2); split 1 never splits.
 * The budget is TDV_SPLIT_ITEMS work items from TDV_SPLIT_WIDE_GROUPS
 * active groups up and half of that below: the sets of a few thousand
 * groups gain from the finer cut (clustered 2^23 box: 3.9% active 3.48 ->
 * 2.69 ms, 11% 6.04 -> 4.99 ms per substep), the sparsest lose on the
 * parts' own overhead (0.15%: 1.18 -> 1.34 ms). */
__host__ __device__ __forceinline__ int walk32_parts(int ngroups, int split)
{
	if (split == 1 || ngroups <= 0)
		return 1;
	const int kmax = split >= 2 ? split : TDV_SPLIT_AUTO_MAX;
	const int k = (ngroups >= TDV_SPLIT_WIDE_GROUPS ? TDV_SPLIT_ITEMS
			: TDV_SPLIT_ITEMS / 2) / ngroups;
	return k < 1 ? 1 : k < kmax ? k : kmax;
}

/* For a split walk: each group's wanted top nodes (bit j of word j / 64:
 * some lane wants top node j's subtree walked), evaluated once per group
 * instead of once per part; the parts other than 0 then go straight to
 * those top nodes.  At most 64 TDV_WANT_WORDS top nodes (more: every part
 * evaluates them itself). */
__global__ void __launch_bounds__(64) k_walk32_want(
		const int *__restrict__ nact_dev, int split,
		const int *__restrict__ active, const int4 *__restrict__ p32,
		const double *__restrict__ last_acc, TopNodes T,
		const int4 *__restrict__ top32, Walk32Params w,
		unsigned long long *__restrict__ wmask)
{
	const int group = blockIdx.x;
	const int nact = *nact_dev;
	const int ngroups = (nact + 63) / 64;
	if (group >= ngroups || walk32_parts(ngroups, split) == 1
			|| *T.nnodes < 0)
		return;
	const int ntop = *T.ntop;
	if (ntop > 64 * TDV_WANT_WORDS)
		return;
	const int gid = group * 64 + (int)threadIdx.x;
	const int ip = gid < nact ? active[gid] : active[nact - 1];
	const int4 me = p32[ip];
	const float crit = (float)(last_acc[ip] * (double)w.alpha_rel);
	for (int q = 0; q < TDV_WANT_WORDS && q * 64 < ntop; q++) {
		unsigned long long bits = 0;
		for (int b = 0; b < 64 && q * 64 + b < ntop; b++) {
			const int j = q * 64 + b;
			const float ts = ldexpf(w.dom_f, -T.level[j]);
			float rx, ry, rz;
			if (__any(top_want(top32[j], ts, me, crit, w, rx, ry, rz)))
				bits |= 1ull << b;
		}
		if (threadIdx.x == 0)
			wmask[(size_t)group * TDV_WANT_WORDS + q] = bits;
	}
}

/* Persistent form: a fixed grid whose waves take work items (a 64-particle
 * group, or one part of one, see walk32_body) from a work queue (one atomic
 * counter) until it is exhausted, with the active count read from device
 * memory, so one launch serves any active subset and a captured graph can
 * be replayed for every substep.  Items of one group are consecutive, so
 * its parts run side by side.  Every wave leaves once the counter has
 * passed the last item. */
__global__ void __launch_bounds__(64) k_walk32_queue(const int *__restrict__ nact_dev, unsigned *__restrict__ queue,
		int split, double4 *__restrict__ partial,
		const unsigned long long *__restrict__ wmask, TDV_WALK32_ARGS)
{
	nact = *nact_dev;
	const int ngroups = (nact + 63) / 64;
	const int nparts = walk32_parts(ngroups, split);
	const int nitems = ngroups * nparts, nslot = ngroups * 64;
	for (;;) {
		/* The ticket: every lane takes part in the atomic and only lane 0
		 * adds, so lane 0's old value is the group; readfirstlane makes it
		 * an SGPR and the exit a scalar branch.  No branch on the lane
		 * here: the compiler merges a `threadIdx.x == 0` test at the loop
		 * head with the one that ends the body, after which lanes 1-63
		 * loop alone and read the ticket from an inactive lane 0 (zero),
		 * i.e. walk group 0 forever. */
		const unsigned t = atomicAdd(queue, threadIdx.x == 0 ? 1u : 0u);
		const int item = __builtin_amdgcn_readfirstlane((int)t);
		if (item >= nitems)
			break;
		const int gi = item / nparts;
		TDV_WALK32_CALL_PARTW(TDV_GROUP(gi), item - gi * nparts, nparts,
				partial, nslot, wmask);
	}
}

/* The same work items (group, part) with one workgroup each, the active
 * count read from device memory: the hardware's workgroup dispatch is the
 * queue.  A grid sized for the largest case serves any active count (a
 * captured graph), the workgroups past the last item leave at once; with
 * the register budgets of the plain walk. */
#define TDV_WALK32_ITEMS_BODY                                               \
	nact = *nact_dev;                                                      \
	const int ngroups = (nact + 63) / 64;                                  \
	const int nparts = walk32_parts(ngroups, split);                       \
	const int item = blockIdx.x;                                           \
	if (item >= ngroups * nparts)                                          \
		return;                                                            \
	const int gi = item / nparts;                                          \
	TDV_WALK32_CALL_PARTW(TDV_GROUP(gi), item - gi * nparts, nparts,       \
			partial, ngroups * 64, wmask)

__global__ void __launch_bounds__(64) k_walk32_items(
		const int *__restrict__ nact_dev, int split,
		double4 *__restrict__ partial,
		const unsigned long long *__restrict__ wmask, TDV_WALK32_ARGS)
{
	TDV_WALK32_ITEMS_BODY;
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6)))
k_walk32_items_w6(const int *__restrict__ nact_dev, int split,
		double4 *__restrict__ partial,
		const unsigned long long *__restrict__ wmask, TDV_WALK32_ARGS)
{
	TDV_WALK32_ITEMS_BODY;
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
k_walk32_items_w8(const int *__restrict__ nact_dev, int split,
		double4 *__restrict__ partial,
		const unsigned long long *__restrict__ wmask, TDV_WALK32_ARGS)
{
	TDV_WALK32_ITEMS_BODY;
}

/* The parts' sums of a split walk, in part order, to the particles (as
 * walk32_body writes them when it does not split) */
__global__ void k_walk32_parts_sum(const int *__restrict__ nact_dev, int split,
		const int *__restrict__ active, const double4 *__restrict__ partial,
		double *__restrict__ ax, double *__restrict__ ay,
		double *__restrict__ az, double *__restrict__ gx,
		double *__restrict__ gy, double *__restrict__ gz,
		double *__restrict__ cost)
{
	const int nact = *nact_dev;
	const int ngroups = (nact + 63) / 64;
	const int nparts = walk32_parts(ngroups, split);
	if (nparts == 1)
		return;
	const size_t nslot = (size_t)ngroups * 64;
	for (int gid = blockIdx.x * blockDim.x + threadIdx.x; gid < nact;
			gid += gridDim.x * blockDim.x) {
		double4 a = partial[gid];
		/* unrolled so that several parts' loads are in flight; the sum
		 * stays in part order */
#pragma unroll 8
		for (int k = 1; k < nparts; k++) {
			const double4 b = partial[k * nslot + gid];
			a.x += b.x;
			a.y += b.y;
			a.z += b.z;
			a.w += b.w;
		}
		const int ip = active[gid];
		ax[ip] = ax[ip] + a.x;
		ay[ip] = ay[ip] + a.y;
		az[ip] = az[ip] + a.z;
		gx[ip] = a.x;
		gy[ip] = a.y;
		gz[ip] = a.z;
		cost[ip] = a.w;
	}
}

/* the sum's grid: a walk splits only when its groups x 2 <= TDV_SPLIT_ITEMS,
 * so that many particles at most (the kernel strides past the grid) */
static unsigned parts_sum_blocks(size_t n)
{
	const size_t most = (size_t)TDV_SPLIT_ITEMS / 2 * 64;
	return tdv_blocks(n < most ? n : most, 256);
}

/* Each group's expected walk time for the launch order: the Cost of its
 * particles' last walk (interactions), summed by one wave per group, as a
 * key on a log scale, 8 keys per factor 2, 0 for the costliest; 255 for
 * the slots past the active groups (they sort last).  Groups then start
 * longest first and the cheap ones fill the last waves' gaps. */
__global__ void __launch_bounds__(256) k_group_keys(
		const int *__restrict__ nact_dev, int ng_max,
		const int *__restrict__ active, const double *__restrict__ cost,
		uint64_t *__restrict__ key)
{
	const int g = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (g >= ng_max)
		return;
	const int nact = *nact_dev;
	const int i = g * 64 + lane;
	double s = i < nact ? cost[active[i]] : 0.0;
	for (int d = 32; d >= 1; d >>= 1)
		s += __shfl_xor(s, d, 64);
	if (lane == 0) {
		unsigned k = 255u;
		if (g * 64 < nact) {
			const float b = 8.f * __log2f(1.f + (float)s);
			k = 254u - (unsigned)fminf(b, 254.f);
		}
		key[g] = k;
	}
}

__global__ void k_fill_u32(int n, unsigned *__restrict__ p)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0;
}

__global__ void k_walk32_queue_reset(unsigned *__restrict__ queue,
		unsigned long long *__restrict__ counter)
{
	if (threadIdx.x == 0)
		queue[0] = 0;
	if (threadIdx.x < 4)
		counter[threadIdx.x] = 0;
}

/* The per-level node lists, once per build: counts, offsets, fill, all on
 * the device (the node count is read there too, so a captured step can
 * rebuild them). */
int tdv_level_lists(tdv_context *c, size_t bound)
{
	if (c->lvl_cap < c->node_cap) {
		if (c->graph_mode)
			return tdv_fail(c, TDV_ERR_ARG, "capture: level lists missing");
		tdv_graph_drop(c);
		if (c->lvl_nodes)
			(void)hipFree(c->lvl_nodes);
		c->lvl_nodes = nullptr;
		c->lvl_cap = 0;
		TDV_HIP(c, hipMalloc((void **)&c->lvl_nodes,
				c->node_cap * sizeof(int)));
		c->lvl_cap = c->node_cap;
	}
	unsigned *cnt = c->q_arrived, *fill = c->q_arrived + 64;
	int *off = reinterpret_cast<int *>(c->q_arrived + 128);
	hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(128), 0, c->stream, 128,
			c->q_arrived);
	hipLaunchKernelGGL(k_level_count, dim3(tdv_blocks(bound, 256)),
			dim3(256), 0, c->stream, (int)bound, c->top.nnodes, c->node_c,
			cnt);
	hipLaunchKernelGGL(k_level_offsets, dim3(1), dim3(64), 0, c->stream, cnt,
			off, fill);
	hipLaunchKernelGGL(k_level_fill, dim3(tdv_blocks(bound, 256)), dim3(256),
			0, c->stream, (int)bound, c->top.nnodes, c->node_c, off, fill,
			c->lvl_nodes);
	TDV_HIP(c, hipGetLastError());
	c->lvl_ok = 1;
	return 0;
}

/* the softening length in fixed units: negative (the reference's K1
 * length), and at least half a unit, so the particle itself (r = 0)
 * always takes the finite inner kernel */
static double walk32_eps(const tdv_context *c, double S)
{
	const double ef = c->Epsilon * S;
	return ef * ef < 0.5 ? -sqrt(0.5) : ef;
}

/* safe_last_accel (accel.c:28-43) with the active count on the device */
__global__ void k_safe_last_accel_dev(const int *__restrict__ nact_dev,
		const int *__restrict__ active, double *__restrict__ ax,
		double *__restrict__ ay, double *__restrict__ az,
		double *__restrict__ last)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= *nact_dev)
		return;
	const int i = active[g];
	last[i] = sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
	ax[i] = 0;
	ay[i] = 0;
	az[i] = 0;
}

int tdv_walk32(tdv_context *c, int use_bh)
{
	if (!c->nactive && !c->capturing)
		return 0;
	const size_t n = c->n, nn = c->nnodes;
	/* fixed-point frame: 2^32 per periodic box, 2^31 per domain otherwise */
	const double S = c->par.periodic ? 4294967296.0 / c->par.boxsize
			: 2147483648.0 / c->dom_size;
	const double o[3] = { c->par.periodic ? 0.0 : c->dom_origin[0],
		c->par.periodic ? 0.0 : c->dom_origin[1],
		c->par.periodic ? 0.0 : c->dom_origin[2] };
	const double mscale = c->par.gravity * S * S;
	const int bits = c->par.periodic ? 32 : 31;    /* fixed units per domain */
	if (!c->p32) {
		TDV_HIP(c, hipMalloc((void **)&c->p32, c->cap * sizeof(int4)));
		TDV_HIP(c, hipMalloc((void **)&c->p32o, c->cap * sizeof(float4)));
	}
	if (!c->gorder) {                    /* the groups' launch order */
		c->gcap = (c->cap + 63) / 64;
		TDV_HIP(c, hipMalloc((void **)&c->gkey,
				2 * c->gcap * sizeof(uint64_t)));
		TDV_HIP(c, hipMalloc((void **)&c->gorder,
				c->gcap * sizeof(uint32_t)));
	}
	if (!c->partial) {                   /* split walks' partial sums */
		TDV_HIP(c, hipMalloc((void **)&c->partial,
				(size_t)TDV_SPLIT_ITEMS * 64 * sizeof(double4)));
		/* and their groups' wanted top nodes: a walk splits with at most
		 * TDV_SPLIT_ITEMS / 2 groups (two parts each) */
		TDV_HIP(c, hipMalloc((void **)&c->wmask, (size_t)TDV_SPLIT_ITEMS / 2
				* TDV_WANT_WORDS * sizeof(unsigned long long)));
	}
	if (c->n32_cap < c->node_cap) {
		tdv_graph_drop(c);               /* the records move */
		if (c->a32)
			(void)hipFree(c->a32);
		if (c->rec32)
			(void)hipFree(c->rec32);
		if (c->ctr32)
			(void)hipFree(c->ctr32);
		c->ctr32 = nullptr;
		if (c->q32)
			(void)hipFree(c->q32);
		if (c->q_s6)
			(void)hipFree(c->q_s6);
		c->q_s6 = nullptr;
		c->a32 = c->rec32 = nullptr;
		c->q32 = nullptr;
		c->n32_cap = 0;
		TDV_HIP(c, hipMalloc((void **)&c->a32, c->node_cap * sizeof(int4)));
		TDV_HIP(c, hipMalloc((void **)&c->rec32,
				c->node_cap * sizeof(Node32)));
		TDV_HIP(c, hipMalloc((void **)&c->ctr32,
				c->node_cap * sizeof(int4)));
		TDV_HIP(c, hipMalloc((void **)&c->q32,
				2 * c->node_cap * sizeof(float4)));
		TDV_HIP(c, hipMalloc((void **)&c->q_s6,
				6 * c->node_cap * sizeof(double)));
		c->n32_cap = c->node_cap;
	}
	const bool quad = c->par.fp32_order != 1;
	if (c->top32_cap < c->top_cap) {
		tdv_graph_drop(c);
		if (c->top32)
			(void)hipFree(c->top32);
		if (c->tq32)
			(void)hipFree(c->tq32);
		c->top32 = nullptr;
		c->tq32 = nullptr;
		c->top32_cap = 0;
		TDV_HIP(c, hipMalloc((void **)&c->top32, c->top_cap * sizeof(int4)));
		TDV_HIP(c, hipMalloc((void **)&c->tq32,
				2 * (size_t)c->top_cap * sizeof(float4)));
		c->top32_cap = c->top_cap;
	}
	/* quadrupoles: masses G S^2, lengths S -> scale G S^4 */
	const double qscale = mscale * S * S;
	/* nodes the kernels may touch: the count of the last eager build, or
	 * the arrays' capacity in a captured step (the count is read on the
	 * device) */
	const size_t bound = c->graph_mode ? c->node_cap : nn;
	const int tgrid = c->graph_mode ? c->top_cap : c->ntop;
	/* a substep keeps the quadrupoles of the last eager walk on this tree */
	if (quad && !c->substep_prep) {
		if (bound && !c->lvl_ok)
			TDV_TRY(tdv_level_lists(c, bound));
		/* bottom-up: a level's nodes take their children's moments */
		if (bound) {
			const unsigned grid = (unsigned)std::min(
					(size_t)tdv_blocks(8 * bound, 256), (size_t)2048);
			for (int lv = TDV_NT; lv >= 0; lv--)
				hipLaunchKernelGGL(k_moments_level, dim3(grid), dim3(256), 0,
						c->stream, lv, c->q_arrived,
						reinterpret_cast<const int *>(c->q_arrived + 128),
						c->lvl_nodes, c->node_c, c->node_a, c->pm, c->q_s6,
						qscale, c->q32);
		}
		if (tgrid)
			hipLaunchKernelGGL(k_moments_top, dim3(tdv_blocks((size_t)tgrid,
					64)), dim3(64), 0, c->stream, c->top, c->par.vector_size,
					c->node_a, c->pm, c->q_s6, qscale, c->tq32);
	}
	if (tgrid)
		hipLaunchKernelGGL(k_prep32_top, dim3(tdv_blocks((size_t)tgrid, 64)),
				dim3(64), 0, c->stream, c->top, o[0], o[1], o[2], S, mscale,
				c->top32);
	if (bound && !c->substep_prep)
		hipLaunchKernelGGL(k_prep32_centres, dim3(tdv_blocks(bound, 256)),
				dim3(256), 0, c->stream, bound, c->top.nnodes, c->node_c,
				c->pm, o[0], o[1], o[2], S, bits, c->ctr32);
	if (bound)
		hipLaunchKernelGGL(k_prep32_nodes, dim3(tdv_blocks(bound, 256)),
				dim3(256), 0, c->stream, bound, c->top.nnodes, c->node_a,
				c->node_b, c->node_c, c->q32, quad ? 1 : 0, c->ctr32, o[0],
				o[1], o[2], S, mscale, c->a32,
				reinterpret_cast<Node32 *>(c->rec32));
	hipLaunchKernelGGL(k_prep32_particles, dim3(tdv_blocks(n, 256)), dim3(256),
			0, c->stream, n, c->pm, o[0], o[1], o[2], S, mscale,
			c->tree_parent, c->a32, c->p32, c->p32o);
	if (c->par.periodic && !c->ewald16_ok) {
		if (!c->have_ewald)
			return tdv_fail(c, TDV_ERR_ARG, "FP32 walk without an Ewald table");
		TDV_TRY(tdv_ewald_derivs_ready(c));
		if (!c->ewald16)
			TDV_HIP(c, hipMalloc((void **)&c->ewald16,
					4 * (size_t)TDV_NE3 * sizeof(float4)));
		/* E: c->ewald is F / Box^2; m' = G m S^2, so E / S^2.  J: per unit
		 * box -> per grid unit (1 / (2 N_EWALD)), / Box^2 / S^2 */
		const double box = c->par.boxsize;
		hipLaunchKernelGGL(k_prep32_ewald_taylor,
				dim3(tdv_blocks(TDV_NE3, 256)), dim3(256), 0, c->stream,
				c->ewald, c->ewald_s, c->ewald_j, 1.0 / (S * S),
				1.0 / (2.0 * TDV_NE * box * box * S * S), c->ewald16);
		TDV_HIP(c, hipGetLastError());
		c->ewald16_ok = 1;
	}
	Walk32Params w;
	w.ntop = c->ntop;
	w.vs = c->par.vector_size;
	w.periodic = c->par.periodic;
	w.use_bh = use_bh;
	w.theta_bh = (float)c->par.theta_bh;
	w.alpha_rel = (float)c->par.alpha_rel;
	const double ef = walk32_eps(c, S);
	w.eps2 = (float)(ef * ef);
	w.eps_abs = (float)fabs(ef);
	w.inv_eps2 = (float)(1.0 / (ef * ef));
	w.k_inner = (float)(1.0 / (16.0 * ef * ef * ef));
	w.box_f = (float)(c->par.periodic ? 4294967296.0 : 2147483648.0);
	w.dom_f = (float)(c->dom_size * S);
	w.half = 0.5f * w.box_f;
	w.ewald_size = (float)(c->par.ewald_size * c->par.boxsize * S);
	w.ewald_acc = (float)c->par.ewald_accuracy;
	w.u_scale = (float)(2.0 * TDV_NE / 4294967296.0);
	w.inv_box_f = 1.f / w.box_f;
	w.inv_s2 = (float)(1.0 / (S * S));
	{
		/* (1/2) S per grid unit = 3 (S/6) h^2 with h = 1/128 box, in E's
		 * units (E / (Box^2 S^2)) */
		const double box = c->par.periodic ? c->par.boxsize : 1.0;
		const double hu = 1.0 / (2.0 * TDV_NE);
		w.taylor2 = (float)(3.0 * hu * hu / (box * box * S * S));
		w.inv_s2_box4 = (float)(1.0 / (S * S * box * box * box * box));
	}
	w.carry_every = c->par.fp32_carry > 0 ? c->par.fp32_carry : 16;
	w.quad = quad ? 1 : 0;
	w.crit3 = c->par.fp32_criterion == 3;
	w.bundle_min = c->par.fp32_bundle_min == 0 ? 4
			: c->par.fp32_bundle_min < 0 ? 0x7FFFFFFF : c->par.fp32_bundle_min;
	const int split = c->par.fp32_split;
	const int ngroups_h = (int)((c->nactive + 63) / 64);
	const int nparts_h = walk32_parts(ngroups_h, split);
	const bool parts = nparts_h > 1;
	/* launch order of the groups: longest expected walk first, from the
	 * Cost of the last walk, for walks of whole groups that fill the chip
	 * a few to a few hundred times (a regular step, eager or captured as a
	 * whole step), where the last groups to start set the end; a split
	 * walk of few groups and a captured walk for any subset keep the list
	 * order, and so do walks of more groups, whose neighbours in the list
	 * then share more of their tree in the caches than the order saves
	 * (128^3: 88 -> 59 ms, 256^3: 543 -> 524 ms, 512^3 uniform: +0.5%) */
	const uint32_t *gord = nullptr;
	const size_t ng_max = (n + 63) / 64;
	const size_t ng_walk = c->graph_mode ? ng_max : (size_t)ngroups_h;
	if ((c->graph_mode || (!c->capturing && !parts))
			&& ng_walk >= TDV_ORDER_MIN_GROUPS
			&& ng_walk <= TDV_ORDER_MAX_GROUPS) {
		const size_t ng = ng_walk;
		hipLaunchKernelGGL(k_group_keys, dim3(tdv_blocks(ng, 4)), dim3(256),
				0, c->stream, c->d_nactive, (int)ng, c->active, c->cost,
				c->gkey);
		TDV_TRY(tdv_check(c, hipGetLastError(), "k_group_keys"));
		TDV_TRY(tdv_radix_pass8(c, ng, c->gkey, c->gkey + c->gcap,
				c->gorder));
		gord = c->gorder;
	}
	if (c->capturing || parts) {
		/* the groups' wanted top nodes (a captured walk finds them on
		 * every replay; the kernel leaves at once when the active groups
		 * do not split) */
		hipLaunchKernelGGL(k_walk32_want, dim3(c->capturing
				? (unsigned)TDV_SPLIT_ITEMS / 2 : (unsigned)ngroups_h),
				dim3(64), 0, c->stream, c->d_nactive, split, c->active,
				c->p32, c->last_acc, c->top, c->top32, w, c->wmask);
		TDV_TRY(tdv_check(c, hipGetLastError(), "k_walk32_want"));
	}
	if (!c->par.fp32_queue && (c->capturing || parts)) {
		/* one workgroup per item; captured: enough for any active set */
		const size_t items = c->capturing
				? std::max((n + 63) / 64, (size_t)TDV_SPLIT_ITEMS)
				: (size_t)ngroups_h * nparts_h;
		auto kern = c->par.fp32_waves == 6 ? k_walk32_items_w6
				: c->par.fp32_waves == 8 ? k_walk32_items_w8 : k_walk32_items;
		hipLaunchKernelGGL(k_walk32_queue_reset, dim3(1), dim3(64), 0,
				c->stream, c->d_queue, c->d_count);
		hipLaunchKernelGGL(kern, dim3((unsigned)items), dim3(64), 0,
				c->stream, c->d_nactive, split, c->partial, c->wmask, 0,
				c->active,
				c->p32, c->last_acc, c->top,
				reinterpret_cast<const Node32 *>(c->rec32), c->ewald16, w, c->acc[0], c->acc[1], c->acc[2], c->gacc[0],
				c->gacc[1], c->gacc[2], c->cost, c->d_count, c->top32,
				c->tq32, c->p32o, gord);
		TDV_TRY(tdv_check(c, hipGetLastError(), "k_walk32_items"));
		hipLaunchKernelGGL(k_walk32_parts_sum, dim3(parts_sum_blocks(n)),
				dim3(256), 0, c->stream, c->d_nactive, split, c->active,
				c->partial, c->acc[0], c->acc[1], c->acc[2], c->gacc[0],
				c->gacc[1], c->gacc[2], c->cost);
		return tdv_check(c, hipGetLastError(), "k_walk32_parts_sum");
	}
	if (c->par.fp32_queue || c->capturing || parts) {
		/* grid: enough waves to fill every SIMD at the 8-wave budget */
		hipLaunchKernelGGL(k_walk32_queue_reset, dim3(1), dim3(64), 0,
				c->stream, c->d_queue, c->d_count);
		hipLaunchKernelGGL(k_walk32_queue, dim3(TDV_QUEUE_GRID), dim3(64), 0,
				c->stream, c->d_nactive, c->d_queue, split, c->partial,
				c->wmask, 0,
				c->active, c->p32, c->last_acc, c->top,
				reinterpret_cast<const Node32 *>(c->rec32), c->ewald16, w, c->acc[0], c->acc[1], c->acc[2],
				c->gacc[0], c->gacc[1], c->gacc[2], c->cost, c->d_count,
				c->top32, c->tq32, c->p32o, gord);
		TDV_TRY(tdv_check(c, hipGetLastError(), "k_walk32_queue"));
		/* a captured walk serves every active count: always sum (the
		 * kernel returns at once when the walk did not split) */
		if (parts || c->capturing)
			hipLaunchKernelGGL(k_walk32_parts_sum, dim3(parts_sum_blocks(n)),
					dim3(256), 0, c->stream, c->d_nactive, split, c->active,
					c->partial, c->acc[0], c->acc[1], c->acc[2], c->gacc[0],
					c->gacc[1], c->gacc[2], c->cost);
		return tdv_check(c, hipGetLastError(), "k_walk32_parts_sum");
	}
	auto kern = c->par.fp32_waves == 6 ? k_walk32_w6
			: c->par.fp32_waves == 8 ? k_walk32_w8 : k_walk32;
	hipLaunchKernelGGL(kern, dim3(tdv_blocks((size_t)c->nactive, 64)),
			dim3(64), 0, c->stream, c->nactive, c->active, c->p32, c->last_acc,
			c->top, reinterpret_cast<const Node32 *>(c->rec32), c->ewald16, w,
			c->acc[0],
			c->acc[1], c->acc[2], c->gacc[0], c->gacc[1], c->gacc[2], c->cost,
			c->d_count, c->top32, c->tq32, c->p32o, gord);
	return tdv_check(c, hipGetLastError(), "k_walk32");
}

void tdv_launch_safe_last_accel_dev(tdv_context *c)
{
	hipLaunchKernelGGL(k_safe_last_accel_dev, dim3(tdv_blocks(c->n, 256)),
			dim3(256), 0, c->stream, c->d_nactive, c->active, c->acc[0],
			c->acc[1], c->acc[2], c->last_acc);
}
