/*
 * Internal declarations of the MI355X gravity engine: device layouts, the
 * context, kernel launchers.  gfx950 only.
 *
 * HBM layout (all SoA, 256-B aligned allocations):
 *   particles  pos[3], vel[3], acc[3], gacc[3], mass, last_acc, cost (f64),
 *              key (2 x u64 per particle), tree_parent, time_bin, id (32-bit)
 *              plus pm = {x, y, z, m} double4 in PH order for the walk;
 *   tree       node_a = {CoM, Mass} double4, node_b = {centre, size}
 *              double4, node_c = {DNext, Npart, Bitfield, first particle}
 *              int4, node_up = DUp;
 *   top nodes  <= 64 records (level-2 cells, NTarget = 1);
 *   Ewald      Fx, Fy, Fz, Fp (65^3 f64 each), divided by Box^2 / Box.
 */
#ifndef TDV_INTERNAL_H
#define TDV_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "../../include/tandav_gpu.h"

#define TDV_NT 42            /* N_PEANO_TRIPLETS (peano.h:8) */
#define TDV_TOP_LEVEL 2      /* MIN_LEVEL with NTarget = 1 (domain.c:3) */
#define TDV_NE 64            /* N_EWALD (Gravity/periodic.c:9) */
#define TDV_NE1 65
#define TDV_NE3 (TDV_NE1 * TDV_NE1 * TDV_NE1)
#define TDV_BIT_LOCAL 9
#define TDV_BIT_TOP 10
#define TDV_BIT_UPDATED 11
#define TDV_MAX_TOP 65536       /* top-node records a context accepts */
#define TDV_SORT_TILE 6144      /* keys per radix-sort block: 256 threads x
                                   24 (longer runs per digit store better;
                                   the LDS staging fits two blocks per CU) */

/* stage events of a context (stage_ms: keys, sort, reorder, top nodes,
 * build, walk, Ewald walk, kick/drift) */
enum { EV_KEYS0, EV_KEYS1, EV_SORT1, EV_REORDER1, EV_TOP1, EV_BUILD0,
	EV_BUILD1, EV_WALK0, EV_WALK1, EV_EWALD1, EV_KICK0, EV_KICK1, EV_N };

/* error bits raised by kernels into ctx->d_err */
#define TDV_EBIT_DUP 1u
#define TDV_EBIT_DEPTH 2u
#define TDV_EBIT_RANGE 4u
#define TDV_EBIT_TIES 8u

struct TopNodes {            /* device arrays, top_cap entries each */
	int *first, *npart, *level, *target, *term;
	uint64_t *key;
	float *pos;              /* 3 per node */
	float *mass;
	float *com;              /* 3 per node */
	float *dp;               /* 3 per node: TNode.Dp, the tree kicks */
	int *kicked;             /* TNode.Level < 0: kicked since the drift */
	int *ptop;               /* per particle: its top node */
	int *ntop;               /* [1] */
	int *nnodes;             /* [1] total tree nodes */
};

struct tdv_context {
	tdv_params par;
	double Epsilon, Epsilon2, Epsilon3;
	double dom_size, dom_origin[3], dom_center[3];
	hipStream_t stream;
	std::string err;

	size_t n, cap;                   /* particles, allocated particles */
	double *pos[3], *vel[3], *acc[3], *gacc[3], *mass, *last_acc, *cost;
	double *tmp_f64;                 /* reorder double buffer */
	uint64_t *key;                   /* 2 words per particle */
	uint64_t *key_tmp;
	int *tree_parent, *time_bin;
	uint32_t *id;
	uint32_t *it_drift, *it_kick;    /* P.It_Drift_Pos, P.It_Kick_Pos */
	double *kick_dt;                 /* each particle's last kick factor */
	double4 *pm;                     /* {x, y, z, m} for the walk */
	int *active;                     /* active list */
	int nactive;
	int *d_nactive;                  /* nactive on the device (the queue
	                                    walk and captured graphs read it) */
	unsigned *d_queue;               /* work-queue counter of the walk */
	double4 *partial;                /* split walk: TDV_SPLIT_ITEMS x 64 */
	unsigned long long *wmask;       /* split walk: wanted top nodes,
	                                    TDV_WANT_WORDS per group */
	uint64_t *gkey;                  /* FP32 walk: 2 per group, cost keys */
	uint32_t *gorder;                /* groups by descending expected cost */
	size_t gcap;                     /* groups gkey and gorder hold */
	double *q_s6;                    /* nodes' FP64 second moments, 6 each */
	unsigned *q_arrived;             /* per level: counts [0, 64), fill
	                                    counters [64, 128), list offsets
	                                    [128, 192) */
	int *lvl_nodes;                  /* node indices grouped by level */
	size_t lvl_cap;                  /* nodes lvl_nodes holds */
	int lvl_ok;                      /* the lists match the tree */
	int substep_prep;                /* capturing a substep walk: the
	                                    quadrupoles and cell centres kept */
	hipGraph_t graph;                /* captured regular-step walk */
	hipGraphExec_t graph_exec;
	int capturing;
	hipGraph_t step_graph;           /* captured whole step */
	hipGraphExec_t step_exec;
	int graph_mode;                  /* capturing the whole step: fixed
	                                    capacities, counts on the device */
	double step_kick, step_drift;    /* the captured step's factors */

	/* sort scratch */
	uint64_t *sk[2];
	uint32_t *sv[2];
	uint32_t *hist;
	size_t hist_cap;
	uint32_t *scan_part;
	uint32_t *perm;

	/* build scratch (per particle) */
	uint32_t *minpack;
	int2 *long_list;                 /* build: chains scanned by a wave */
	size_t long_cap;
	unsigned *long_count;            /* [0] long chains, [1] their queue */
	int *cnt;
	uint32_t *off;                   /* n+1 exclusive scan of cnt */
	int *lohi;

	/* tree */
	size_t node_cap, nnodes;
	double4 *node_a, *node_b;
	int4 *node_c;                    /* DNext, Npart, Bitfield, first part. */
	int *node_up;                    /* DUp */
	double4 *node_dp;                /* Dp of the tree kicks (tree.h:31) */

	TopNodes top;
	int ntop, top_cap;
	int4 *top32;                     /* FP32 walk: fixed-point top CoM */
	int top32_cap;

	double *ewald;                   /* 4 * TDV_NE3, normalised */
	std::vector<double> ewald_host;  /* the same on the host (per-element
	                                    Ewald_Correction) */
	int have_ewald;

	/* FP32 group walk: fixed-point records */
	int4 *p32, *a32;
	int4 *rec32;                     /* 64-byte node records, 4 int4 each */
	int4 *ctr32;                     /* node cell centres, fixed units */
	float4 *p32o;                    /* offset from the leaf's CoM, mass' */
	size_t n32_cap;
	float4 *q32;                     /* 2 per node: traceless quadrupole */
	float4 *tq32;                    /* 2 per top node */
	double *ewald_s;                 /* 10 * TDV_NE3: second derivatives
	                                    of the correction, per unit box */
	double *ewald_j;                 /* 6 * TDV_NE3: first derivatives
	                                    d_b E_a, xx yy zz xy xz yz */
	float4 *ewald16;                 /* FP32 walk: 4 float4 per grid node,
	                                    {E, J, S} in kernel units */
	int ewald16_ok;                  /* ewald16 matches ewald */

	unsigned *d_err;
	unsigned long long *d_count;     /* [4] interactions: tree, Ewald;
	                                    FP32 walk: of which multipole nodes,
	                                    corrections with quadrupoles */
	unsigned long long count[4];
	int have_domain, have_tree;
	int have_domain_set, have_top;

	hipEvent_t ev[16];
	float stage_ms[8];

	/* integer timeline of the KDK step (tdv_integrate.hip) */
	tdv_timeline tl;
	double tl_step_min, tl_step_max;
	double *tl_spline;               /* 5 x 512: knots, kick y, c, drift y, c */
	int have_timeline;
};

/* ---- launchers (tdv_keys_sort.hip) ---- */
void tdv_launch_keys(tdv_context *c, size_t n, const double *x,
		const double *y, const double *z, const double o[3], double size,
		uint64_t *out, int reversed);
int tdv_sort_device(tdv_context *c, size_t n, const uint64_t *keys,
		uint32_t *perm_out);
void tdv_launch_gather_f64(tdv_context *c, size_t n, const uint32_t *perm,
		const double *in, double *out);
void tdv_launch_gather_i32(tdv_context *c, size_t n, const uint32_t *perm,
		const int *in, int *out);
void tdv_launch_gather_key(tdv_context *c, size_t n, const uint32_t *perm,
		const uint64_t *in, uint64_t *out);
void tdv_launch_reverse_keys(tdv_context *c, size_t n, const uint64_t *in,
		uint64_t *out);
void tdv_launch_pack_pm(tdv_context *c);
int tdv_domain_nonperiodic(tdv_context *c);
int tdv_scan_u32(tdv_context *c, const uint32_t *in, uint32_t *out,
		size_t n);
int tdv_radix_pass8(tdv_context *c, size_t n, const uint64_t *keys,
		uint64_t *kout, uint32_t *order);

/* ---- launchers (tdv_build.hip) ---- */
int tdv_build_top_nodes(tdv_context *c);
int tdv_assign_top(tdv_context *c);
int tdv_build_tree(tdv_context *c);
/* the tree's per-level node lists (lvl_nodes, counts and offsets in
 * q_arrived), on the device; sets lvl_ok */
int tdv_level_lists(tdv_context *c, size_t bound);

/* ---- launchers (tdv_walk64.hip, tdv_ewald.hip) ---- */
int tdv_walk64(tdv_context *c, int use_bh);
int tdv_walk64_ewald(tdv_context *c, int use_bh);
int tdv_walk32(tdv_context *c, int use_bh);
void tdv_launch_safe_last_accel_dev(tdv_context *c);
void tdv_launch_kick_drift(tdv_context *c, double kick, double drift);
void tdv_graph_drop(tdv_context *c);
#define TDV_QUEUE_GRID 8192              /* 256 CUs x 4 SIMDs x 8 waves;
                                            waves beyond those resident
                                            start later and find the queue
                                            drained */
#define TDV_SPLIT_ITEMS 32768            /* work items a split walk aims at
                                            (half of them below ... */
#define TDV_SPLIT_WIDE_GROUPS 4096       /* ... this many active groups) */
#define TDV_LONG_CHAIN 256               /* build: longer chain scans go to
                                            one wavefront each */
#define TDV_LONG_GRID 2048               /* waves of that kernel */
#define TDV_SPLIT_AUTO_MAX 64            /* parts per group, split 0 */
#define TDV_WANT_WORDS 4                 /* a split walk's wanted-top-node
                                            mask: up to 256 top nodes */
#define TDV_ORDER_MIN_GROUPS 2048        /* FP32 walks of this many whole
                                            groups ... */
#define TDV_ORDER_MAX_GROUPS (1 << 20)   /* ... to this many start the
                                            costliest first */
int tdv_direct_sum(tdv_context *c, int nt, const int *idx_dev, double *out_dev);
int tdv_direct_sum_exact(tdv_context *c, int nt, const int *idx_dev,
		double *out_dev);
int tdv_direct_sum_split(tdv_context *c, int nt, const int *idx_dev,
		double *out_dev);
void tdv_launch_safe_last_accel(tdv_context *c);
int tdv_ewald_table_device(tdv_context *c, double *out4_unnormalised);
int tdv_ewald_derivs_device(tdv_context *c, double *out10, double *out6);
int tdv_ewald_derivs_ready(tdv_context *c);
void tdv_launch_ewald_correction(tdv_context *c, size_t n, const double *dr,
		double *f);

/* D[].TNode (domain.h:40-52) as the C-ABI passes it: 64-byte records */
struct TopRecordIn {
	uint64_t Key;
	int Target, Level, First_Part, Npart;
	float Pos[3], Mass, CoM[3], Dp[3];
};

/* Domain_Decomposition's bunch loop with NTarget = ntarget on the sorted
 * particles (tdv_domain.hip): top-node records in descending key order,
 * Target = -task - 1 */
int tdv_domain_refine(tdv_context *c, int ntarget, double part_alloc_factor,
		double imbal_ceil, std::vector<TopRecordIn> &out, int *iterations);

/* ---- helpers (tdv_api.hip) ---- */
int tdv_fail(tdv_context *c, int code, const std::string &msg);
int tdv_check(tdv_context *c, hipError_t e, const char *what);
int tdv_check_err_flags(tdv_context *c);
int tdv_ensure_nodes(tdv_context *c, size_t nodes);

#define TDV_TRY(expr)                                                       \
	do {                                                                    \
		int _r = (expr);                                                    \
		if (_r)                                                             \
			return _r;                                                      \
	} while (0)

#define TDV_HIP(ctx, expr) TDV_TRY(tdv_check((ctx), (expr), #expr))

static inline unsigned tdv_blocks(size_t n, unsigned t)
{
	return (unsigned)((n + t - 1) / t);
}

#endif
