/*
 * tandav_gpu.h -- C ABI of the MI355X gravity engine for Tandav.
 *
 * One context owns one GPU's copy of the particle columns, the top nodes,
 * the octree and the Ewald table, and runs the per-step force path on the
 * device: PH keys -> radix sort -> reorder -> top nodes -> octree build ->
 * tree walk (+ Ewald walk) -> kick/drift.  Every entry point returns 0 on
 * success or a TDV_ERR_* code; tdv_last_error() gives the message.  No C++
 * exception crosses this boundary.  Calls on one context are not thread
 * safe (the reference calls its team-collective functions from an OpenMP
 * team; the drop-in layer funnels them through one thread).
 *
 * Each declaration names the reference function (jdonnert/Tandav) whose
 * work it performs.  "host arrays" are caller-owned host memory; all other
 * state lives in device memory inside the context.
 */
#ifndef TANDAV_GPU_H
#define TANDAV_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	TDV_OK = 0,
	TDV_ERR_ARG = 1,          /* bad argument / call order               */
	TDV_ERR_HIP = 2,          /* HIP runtime error                        */
	TDV_ERR_NOMEM = 3,        /* device allocation failed                 */
	TDV_ERR_DUP_KEYS = 4,     /* tdv_sort_keys: two keys are equal        */
	TDV_ERR_KEY_DEPTH = 5,    /* not raised: particles closer than the key
	                             resolution are dumped as the reference
	                             does (Gravity/tree_build.c:337-342)      */
	TDV_ERR_RANGE = 6,        /* a particle range left the arrays         */
	TDV_ERR_TIES = 7          /* a run of equal 64-bit key prefixes too long */
};

typedef struct tdv_context tdv_context;

/* Config / parameter-file knobs of the force path, at run time. */
typedef struct {
	int periodic;          /* PERIODIC (Config:7)                              */
	double boxsize;        /* Sim.Boxsize[0] (periodic runs)                   */
	double gravity;        /* Const.Gravity (constants.c:3-8)                  */
	double softening;      /* Param.Grav_Softening[1], Plummer equivalent      */
	double theta_bh;       /* TREE_OPEN_PARAM_BH (Config:47)                   */
	double alpha_rel;      /* TREE_OPEN_PARAM_REL (Config:48)                  */
	int vector_size;       /* VECTOR_SIZE (Config:54), must be 64              */
	int compat;            /* 1: small top nodes read particles
	                          [Target, Target+Npart) like
	                          interact_with_topnode_particles
	                          (Gravity/tree_accel.c:183-203); 0: their own
	                          particles                                        */
	int device;            /* HIP device ordinal                               */
	int precision;         /* TDV_WALK_FP64: the reference's per-particle walk
	                          (identical interaction sets); TDV_WALK_FP32: the
	                          group walk on fixed-point separations           */
	double ewald_size;     /* FP32 walk: largest node (box units) whose Ewald
	                          correction is taken at its centre of mass       */
	double ewald_accuracy; /* FP32 walk: factor on the opening parameter for
	                          taking a node's Ewald correction at its centre
	                          of mass (smaller = more accurate)               */
	int fp32_order;        /* FP32 walk: multipole order of accepted nodes,
	                          1 = monopole (the reference's), 2 (or 0) =
	                          monopole + quadrupole                           */
	int fp32_criterion;    /* FP32 walk: relative opening test, 2 (or 0) =
	                          M s^2 > r^4 |a| alpha / G (the reference's,
	                          Gravity/tree_accel.c:257), 3 = M s^3 > r^5 |a|
	                          alpha / G (the error bound of the quadrupole
	                          walk)                                           */
	int fp32_carry;        /* FP32 walk: interactions summed in FP32 before
	                          the partial sums are carried into FP64
	                          (0 = 16)                                        */
	int fp32_waves;        /* FP32 walk: register budget for 6 or 8 waves per
	                          SIMD (0 = the compiler's own allocation)        */
	int fp32_queue;        /* FP32 walk: 1 = persistent work-queue kernel
	                          (fixed grid, groups taken from a counter); 0 =
	                          one wave per group, or per work item for
	                          split and captured walks                       */
	int fp32_split;        /* FP32 walk with few active groups: parts each
	                          group's subtree walks are cut into, one wave per
	                          part (summed in part order); 0 = up to 64 parts,
	                          1 = never split, n >= 2 = up to n (<= 256)      */
	int fp32_bundle_min;   /* FP32 walk: a particle bundle (collapsed leaf)
	                          of at least this many particles is tested with
	                          the opening criterion like a node and taken
	                          whole (monopole + quadrupole at its centre of
	                          mass) when no lane opens it; 0 = 4, -1 = never
	                          (every bundle particle by particle, as the
	                          reference's walk, Gravity/tree_accel.c:219-240) */
} tdv_params;

enum { TDV_WALK_FP64 = 0, TDV_WALK_FP32 = 1 };

/* Setup_Gravity_Tree (Gravity/tree_build.c:124-142) and, for periodic runs,
 * Gravity_Periodic_Init (Gravity/periodic.c:180-233): softening lengths,
 * device buffers and the FP64 Ewald table computed on the GPU. */
int tdv_create(const tdv_params *p, tdv_context **out);

/* Gravity_Tree_Free (Gravity/tree_build.c:144-156) plus everything else. */
void tdv_destroy(tdv_context *ctx);

const char *tdv_last_error(const tdv_context *ctx);

/* Copy the particle columns P.Pos/Vel/Mass/Acc/Last_Acc_Mag/Time_Bin/ID
 * (particles.h:17-37) from host arrays to the device.  Any pointer but
 * pos and mass may be NULL (zero / all-active / identity). */
int tdv_set_particles(tdv_context *ctx, size_t n, const double *pos[3],
		const double *vel[3], const double *mass, const double *acc[3],
		const double *last_acc_mag, const int *time_bin, const uint32_t *id);

/* Copy device columns back to host arrays (NULL skips a column).  key is
 * 2*n uint64 words (lo, hi) holding P.Key as the reference leaves it:
 * reversed keys after a domain decomposition (peano.c:84-97). */
int tdv_get_particles(tdv_context *ctx, double *pos[3], double *vel[3],
		double *mass, double *acc[3], double *grav_acc[3], double *cost,
		double *last_acc_mag, int *tree_parent, uint64_t *key, uint32_t *id,
		int *time_bin);

/* Setup_Domain_Decomposition (domain.c:151-189): computational domain,
 * then Sort_Particles_By_Peano_Key (peano.c:20-48). */
int tdv_setup_domain(tdv_context *ctx);

/* Domain_Decomposition (domain.c:67-145) for one rank, NTarget = 1:
 * domain, PH keys, key sort, reorder of every column, level-2 top nodes in
 * descending key order, reversed keys (Reverse_Peano_Keys, peano.c:84-97),
 * active list (Make_Active_Particle_List, timestep.c:280-299). */
int tdv_domain_decomposition(tdv_context *ctx);

/* Domain_Decomposition (domain.c:67-145) with NTarget = ntarget (1..512;
 * Setup_Domain_Decomposition, domain.c:151-156, takes NTask or NRank): the
 * domain, PH keys, key sort and reorder as tdv_domain_decomposition, then
 * the bunch refinement loop (reset_bunchlist, fill_new_bunches on the
 * particles' Cost, distribute, find_global_imbalances,
 * mark_bunches_to_split, make_new_bunchlist; domain.c:229-313, 370-692)
 * until the cost imbalance is below imbal_ceil (DOMAIN_IMBAL_CEIL, Config:11,
 * 0.05) and the particle imbalance below part_alloc_factor - 1
 * (PartAllocFactor, default 1.1), or after 20 passes.  The bunches become
 * the top nodes (transform_bunches_into_top_nodes, domain.c:220-250).
 * tasks (may be NULL, max_tasks entries): the target each top node went
 * to, in top-node order; ntop, iterations (may be NULL): the top-node count
 * and the loop's pass count.  Fails when fewer non-empty bunches than
 * targets remain (the reference indexes past its bunch list there). */
int tdv_domain_decomposition_targets(tdv_context *ctx, int ntarget,
		double part_alloc_factor, double imbal_ceil, int *tasks, int max_tasks,
		int *ntop, int *iterations);

/* The Domain (domain.h:58-63) the caller computed, for the next
 * tdv_sort_particles. */
int tdv_set_domain(tdv_context *ctx, double size, const double origin[3]);

/* Sort_Particles_By_Peano_Key (peano.c:20-48) with the domain given to
 * tdv_set_domain: keys, key sort, reorder of every column.  Particles with
 * equal keys (within Domain.Size / 2^42 on every axis) keep their input
 * order; the reference's unstable quicksort (sort.c:253-341) puts them in
 * an order of its own.  The tree build dumps them as the reference does. */
int tdv_sort_particles(tdv_context *ctx);

/* D[].TNode (domain.h:40-52) from the caller's own decomposition
 * (transform_bunches_into_top_nodes, domain.c:220-250): ntop 64-byte
 * records in descending key order whose [First_Part, First_Part+Npart)
 * ranges tile the particles; Key, Level, First_Part, Npart and Pos are read.
 * Replaces the NTarget = 1 top nodes of tdv_domain_decomposition. */
int tdv_set_top_nodes(tdv_context *ctx, const void *records, int ntop);

/* Gravity_Tree_Build (Gravity/tree_build.c:47-122) from the sorted keys. */
int tdv_tree_build(tdv_context *ctx);

/* Gravity_Tree_Acceleration (Gravity/tree_accel.c:32-91) followed, when
 * periodic, by Gravity_Tree_Periodic (Gravity/tree_periodic.c:23-72), with
 * the BH (use_bh = 1) or relative criterion, FP64, one lane per active
 * particle. */
int tdv_tree_acceleration(tdv_context *ctx, int use_bh);

/* Compute_Acceleration (accel.c:15-71): |a| into Last_Acc_Mag, build if a
 * decomposition happened, BH walk + relative walk on the prepare step,
 * relative walk otherwise. */
int tdv_compute_acceleration(tdv_context *ctx, int prepare_step);

/* A regular step's FP32 walk (safe_last_accel, accel.c:28-43, the walk
 * records and the walk over the active count held on the device) captured as
 * a HIP graph on the context's stream, for replay on substeps that reuse
 * the tree (hierarchical time steps: Active_Particle_List changes,
 * positions drift through tdv_update_particles, the tree stays).  Needs one
 * eager FP32 walk on the current tree first; a sort, rebuild or new
 * particles drop the graph.  Engine-side; no counterpart in the reference. */
int tdv_capture_walk(tdv_context *ctx);
/* Replay the captured walk (Compute_Acceleration on a regular step without
 * a rebuild). */
int tdv_launch_walk(tdv_context *ctx);

/* The captured walk for substeps that keep the tree (hierarchical steps:
 * Drift_To_Sync_Point moves the active particles, drift.c:12-36, and
 * tdv_tree_update_drift the nodes' centres of mass,
 * Gravity/tree_update.c): the replay refreshes the walk's node and particle
 * records but keeps the nodes' quadrupoles of the last eager walk on this
 * tree, as the reference's tree update keeps every node property but the
 * drifted centre of mass; the per-level quadrupole pass is most of a small
 * substep's cost.  Replayed by tdv_launch_walk; replaces a graph
 * tdv_capture_walk made.  Engine-side; no counterpart in the reference. */
int tdv_capture_substep_walk(tdv_context *ctx);

/* The whole regular step of a periodic FP32 run as one HIP graph:
 * Domain_Decomposition (domain.c:67-145: PH keys, key sort, reorder, the
 * level-2 top nodes), Compute_Acceleration on a regular step (accel.c:15-71:
 * safe_last_accel, octree build, the FP32 walk with its Ewald correction)
 * and, when kick or drift is nonzero, the KDK step of tdv_kick_drift with
 * those factors.  It needs one eager step on the same particles first (the
 * graph keeps the capacities of that step's node arrays and walk records);
 * nothing runs at capture.  Engine-side; no counterpart in the reference. */
int tdv_capture_step(tdv_context *ctx, double kick, double drift);
/* Replay the captured step.  The state afterwards is the eager step's, bit
 * for bit; stage time 5 holds the whole step.  TDV_ERR_TIES (more than 4096
 * equal 64-bit key prefixes: the captured sort has no fallback) and
 * TDV_ERR_RANGE (a tree larger than the captured arrays) ask for an eager
 * step instead; a graph is dropped whenever its buffers may move. */
int tdv_launch_step(tdv_context *ctx);

/* Gravity_Forcetest (Gravity/forcetest.c:24-174): FP64 direct summation
 * over all particles (plus the tabulated Ewald correction of every pair on
 * periodic runs) for nt targets given by index in the current particle
 * order; acc_out = 3*nt doubles. */
int tdv_forcetest(tdv_context *ctx, int nt, const int *idx, double *acc_out);

/* tdv_forcetest with the exact periodic force instead of the 65^3 table's
 * interpolation: the Ewald lattice sum of the Hernquist, Bouchet & Suto
 * form (Gravity/periodic.c:352-403) that the table samples, here split at
 * screening 4 (per pair: the nearest-image Newton term and the real-space
 * images within 1.625 box; over all particles: the structure factors of
 * the wave vectors with h^2 <= 63), converged below FP64 rounding.  The
 * periodic force without the table's interpolation error, for stating the
 * walks' error.  The same as tdv_forcetest on non-periodic runs. */
int tdv_forcetest_exact(tdv_context *ctx, int nt, const int *idx,
		double *acc_out);

/* The same exact force with each pair's correction taken from the
 * reference's own sums (screening 2, |n|, |h| <= 4 per component,
 * Gravity/periodic.c:352-403) evaluated at the pair's separation: about
 * 100 times slower than tdv_forcetest_exact, which it pins. */
int tdv_forcetest_exact_pairs(tdv_context *ctx, int nt, const int *idx,
		double *acc_out);

/* Element-wise batched forms of the pure functions (host arrays). */
/* Peano_Key / Reversed_Peano_Key (peano.c:122-276); keys_out = 2n words. */
int tdv_peano_keys(tdv_context *ctx, size_t n, const double *x,
		const double *y, const double *z, const double origin[3], double size,
		int reversed, uint64_t *keys_out);

/* The same keys computed on the host, by the same code as the device's:
 * no device work, no context, safe to call from any thread at once (the
 * binding's per-element Peano_Key, Reversed_Peano_Key, peano.c:122-276).
 * TDV_ERR_ARG for missing arrays or size <= 0. */
int tdv_peano_keys_host(size_t n, const double *x, const double *y,
		const double *z, const double origin[3], double size, int reversed,
		uint64_t *keys_out);

/* Qsort_Index(perm, keys, n, 16, cmp_peanoKeys) (sort.c:90-116): the
 * permutation that sorts unique 128-bit keys ascending; keys = 2n words. */
int tdv_sort_keys(tdv_context *ctx, size_t n, const uint64_t *keys,
		uint32_t *perm_out);

/* tdv_sort_keys without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc); it has no counterpart in the reference.  keys
 * (2n words, (lo, hi) per key) and perm_out (n entries) are DEVICE
 * pointers; stream is a hipStream_t (NULL = the null stream) on the
 * current device.  The same kernels as tdv_sort_keys work in a static
 * per-device scratch array of about 37 MB (held until the process ends),
 * so n may be at most TDV_NOALLOC_MAX_KEYS (larger n: TDV_ERR_ARG, nothing
 * launched).  Returns after perm_out is written (it synchronises the
 * stream); TDV_ERR_DUP_KEYS if two keys are equal.  Every call on a GPU
 * shares that GPU's scratch: calls that overlap on one GPU race. */
#define TDV_NOALLOC_MAX_KEYS (1u << 20)
int tdv_sort_keys_noalloc(size_t n, const uint64_t *keys, uint32_t *perm_out,
		void *stream);

/* Ewald_Correction (Gravity/periodic.c:35-113) for n separations
 * dr[3n] -> f[3n]. */
int tdv_ewald_correction(tdv_context *ctx, size_t n, const double *dr,
		double *f);

/* The same lookup on the host from the context's host copy of the table,
 * by the same code as the device's: no device work, safe to call from any
 * thread at once while no call changes the table (the binding's
 * per-element Ewald_Correction, Gravity/periodic.c:35-113).  TDV_ERR_ARG
 * without a table. */
int tdv_ewald_correction_host(const tdv_context *ctx, size_t n,
		const double *dr, double *f);

/* read/write_ewald_correction_table (Gravity/periodic.c:303-342): the
 * unnormalised 4 x 65^3 table (Fx, Fy, Fz, Fp) in box units. */
int tdv_get_ewald_table(tdv_context *ctx, double *table4);
int tdv_set_ewald_table(tdv_context *ctx, const double *table4);

/* Second derivatives S_abc = d_b d_c E_a of the Ewald correction E
 * (Ewald_Correction's function, Gravity/periodic.c:35-113, continued past
 * the half-box faces) at the 65^3 table nodes, per unit box: 10 x 65^3
 * doubles, components xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz.  The FP32
 * walk uses them to take a node's correction at its centre of mass to
 * second order; the reference has no counterpart. */
int tdv_get_ewald_derivs(tdv_context *ctx, double *s10);

/* First derivatives J_ab = d_b E_a of the same correction at the same nodes,
 * per unit box: 6 x 65^3 doubles, components xx yy zz xy xz yz (E is a
 * gradient, so J is symmetric).  With the second derivatives they give the
 * FP32 walk the correction near each node to second order; the reference
 * has no counterpart. */
int tdv_get_ewald_jacobian(tdv_context *ctx, double *j6);

/* The FP32 walk's quadrupole records of the last FP32 walk: per tree node
 * (tdv_get_tree's order) and per top node 8 floats, {Txx, Tyy, Txy, Txz,
 * Tyz, 0, 0, 0}, T = sum m (3 y y^T - |y|^2 I) about the node's centre of
 * mass (a top node's float one), masses G m S^2 and lengths in the walk's
 * fixed units (S = 2^32 per box on periodic runs, 2^31 per domain
 * otherwise); node_first gets each node's first particle (-1 for
 * terminators).  Any output may be NULL.  A diagnostic of the FP32 walk; the
 * reference has no counterpart. */
int tdv_get_walk32_quadrupoles(tdv_context *ctx, float *nodes, float *top,
		int *node_first);

/* Domain (domain.h:58-63) after the last decomposition. */
int tdv_get_domain(tdv_context *ctx, double *size, double origin[3],
		double center[3]);

/* D[].TNode (domain.h:40-52), 64-byte records in the reference layout;
 * out = NULL only reports ntop. */
int tdv_get_top_nodes(tdv_context *ctx, void *out, int max, int *ntop);

/* Tree and NNodes (Gravity/tree.h:23-39), 96-byte Tree_Node records;
 * out = NULL only reports nnodes. */
int tdv_get_tree(tdv_context *ctx, void *out, size_t max, size_t *nnodes);

/* Wall times (ms, HIP events on the context's stream) of the stages of the
 * last call: 0 keys, 1 sort, 2 reorder, 3 top nodes, 4 build, 5 walk,
 * 6 Ewald walk, 7 kick/drift. */
int tdv_stage_times(tdv_context *ctx, float ms[8]);

/* Sum of P.Cost over the active particles of the last walk, split into
 * the tree walk (interact, Gravity/tree_accel.c:391) and the Ewald walk
 * (interact_with_ewald_cube, Gravity/tree_periodic.c:437). */
int tdv_interaction_counts(tdv_context *ctx, uint64_t counts[2]);

/* The same two sums and, for the FP32 walk, the parts of them that took
 * more than a point mass: counts[2] of the tree interactions were nodes or
 * bundles taken with their quadrupole, counts[3] of the Ewald corrections
 * carried the quadrupole term (zero for the FP64 walk).  An extension for
 * the flop count of the FP32 walk; no counterpart in the reference. */
int tdv_interaction_kinds(tdv_context *ctx, uint64_t counts[4]);

/* P.Key as 2n uint64 words: the standard keys of the last sort
 * (reversed = 0, what Sort_Particles_By_Peano_Key leaves) or the reversed
 * keys (reversed = 1, what Reverse_Peano_Keys leaves, peano.c:84-97). */
int tdv_get_keys(tdv_context *ctx, uint64_t *keys, int reversed);

/* reorder_collisionless_particles (peano.c:50-82) for P columns the engine
 * does not hold: applies the last sort's permutation to ncols host columns
 * of n elements of elem_bytes[c] (4, 8 or 16) bytes each, in place. */
int tdv_permute_host_columns(tdv_context *ctx, int ncols, void **cols,
		const int *elem_bytes);

/* Positions, masses, accelerations and Last_Acc_Mag the host changed
 * since the last sort (Drift_To_Sync_Point, drift.c; Kick_*_Halfstep,
 * kick.c), in the current particle order; the keys, top nodes and tree
 * stay as they are (a step without Sig.Tree_Update, accel.c:53-54).
 * Any pointer but pos may be NULL (column left as it is). */
int tdv_update_particles(tdv_context *ctx, const double *pos[3],
		const double *mass, const double *acc[3], const double *last_acc_mag);

/* Active_Particle_List / NActive_Particles (timestep.c:280-299): the
 * particles (indices in the current order) the next walk computes.  A sort
 * (tdv_sort_particles, tdv_domain_decomposition) resets it to every
 * particle. */
int tdv_set_active(tdv_context *ctx, int nactive, const int *list);

/* Tree[].CoM/Mass and D[].TNode.CoM/Mass after the host moved them
 * (Gravity_Tree_Update_Kicks / _Drift, Gravity/tree_update.c:12-104), so a
 * walk without a rebuild sees the drifted monopoles.  tree = nnodes 96-byte
 * Tree_Node records, top = ntop 64-byte top-node records. */
int tdv_set_tree_moments(tdv_context *ctx, const void *tree, size_t nnodes,
		const void *top, int ntop);

/* Make_Active_Particle_List (timestep.c:280-299) on the device: the
 * particles with Time_Bin <= max_active_bin, in particle order, become the
 * active list of the next walk (and the device count captured walks read).
 * time_bin, if not NULL, first replaces the context's P.Time_Bin column (n
 * ints, current particle order).  No active particle is an error, as the
 * reference's Assert. */
int tdv_make_active_list(tdv_context *ctx, int max_active_bin,
		const int *time_bin);

/* The FP32 walk's launch options of an existing context (the tdv_params
 * fields fp32_waves, fp32_queue, fp32_split, same ranges): they change the
 * schedule, not the interactions, so an autotuner can time them on one tree.
 * Drops a captured walk.  No counterpart in the reference. */
int tdv_set_fp32_options(tdv_context *ctx, int waves, int queue, int split);

/* The KDK step of a run on one global time step (every particle active and
 * in the same bin; otherwise TDV_ERR_ARG): Kick_First/Second_Halfstep
 * (kick.c:9-70) as v += kick * Acc, then Drift_To_Sync_Point
 * (drift.c:12-64) as x += drift * v, wrapped into the box on periodic runs
 * (Periodic_Constrain_Particles_To_Box, periodic.c:30-53), in FP64 on the
 * device; a zero factor skips its half.  kick and drift are the caller's
 * Particle_Kick_Step / Particle_Drift_Step (comoving.c:30-50).  Hierarchical
 * steps, where each particle kicks and drifts from its own place on the
 * integer timeline, are tdv_kick_halfstep and tdv_drift_to_sync_point.
 * The keys and tree stay: sort again (or move the tree's monopoles) before
 * the next walk.  Timed as stage 7. */
int tdv_kick_drift(tdv_context *ctx, double kick, double drift);

/* The integer timeline of the KDK step: Setup_Time_Integration
 * (timestep.c:130-167; intime_t is 32 bits, globals.h:33, so the run is
 * 2^31 ticks of Time.Step_Min, in ln a on comoving runs) and, on comoving
 * runs, Setup_Comoving's kick and drift factor tables
 * (setup_kick_drift_factors, comoving.c:109-153: 512 knots from 0.95
 * time_begin to 1.05 time_end, natural cubic splines) over the background
 * of cosmology.c:65-74. */
typedef struct {
	int comoving;          /* COMOVING (Config)                               */
	double time_begin;     /* Time.Begin (expansion factor when comoving)     */
	double time_end;       /* Time.End                                        */
	double hubble;         /* Cosmo.Hubble_Constant in code units (comoving)  */
	double omega_0, omega_matter, omega_lambda, omega_rad;  /* Cosmo.Omega_*  */
} tdv_timeline;

int tdv_set_timeline(tdv_context *ctx, const tdv_timeline *tl);

/* setup_kick_drift_factors (comoving.c:109-153) on the host, no context:
 * the 512 knots a[] (a[0] = 0, the knot the reference leaves unset) and the
 * integrals of 1 / (H a^2) (kick[]) and 1 / (H a^3) (drift[]) from
 * 0.95 time_begin, for a comoving timeline. */
int tdv_comoving_tables(const tdv_timeline *tl, double *a, double *kick,
		double *drift);

/* P.It_Drift_Pos and P.It_Kick_Pos (particles.h:20-21), n uint32 each in
 * the current particle order (NULL skips one).  tdv_set_particles sets both
 * to Int_Time.Beg = 0; sorts carry them with the other columns. */
int tdv_set_integer_times(tdv_context *ctx, const uint32_t *it_drift,
		const uint32_t *it_kick);
int tdv_get_integer_times(tdv_context *ctx, uint32_t *it_drift,
		uint32_t *it_kick);

/* Kick_First_Halfstep / Kick_Second_Halfstep (kick.c:9-70) for the active
 * particles: each kicks by half of its own bin from its own It_Kick_Pos,
 * v += Particle_Kick_Step(It_Kick_Pos, It_Kick_Pos + 2^Time_Bin / 2) * Acc,
 * and advances It_Kick_Pos; FP64, the factor per particle on the device
 * (comoving: the kick spline at both ends).  With update_tree (a step
 * without a domain update, !Sig.Domain_Update) the tree takes the kicks too,
 * Gravity_Tree_Update_Kicks (Gravity/tree_update.c:12-57): every tree node
 * on a kicked particle's path below its subtree root takes
 * Dp += m dt a / Mass and the UPDATED bit, its top node the same in float
 * (Level < 0 in tdv_get_top_nodes until the drift); each node's sum runs
 * over its particles in particle order, as the reference's one-thread sums.
 * Timed as stage 7. */
int tdv_kick_halfstep(tdv_context *ctx, int update_tree);

/* Gravity_Tree_Update_Drift (Gravity/tree_update.c:64-104): the UPDATED
 * nodes move CoM += dt Dp and clear Dp and the flag; the kicked top nodes
 * the same in float. */
int tdv_tree_update_drift(tdv_context *ctx, double dt);

/* Drift_To_Sync_Point (drift.c:12-64) for the active particles: each drifts
 * from its own It_Drift_Pos to it_next (Int_Time.Next),
 * x += Particle_Drift_Step(It_Drift_Pos, it_next) * v, It_Drift_Pos =
 * it_next, wrapped into the box on periodic runs
 * (Periodic_Constrain_Particles_To_Box, periodic.c:30-53; particles that
 * do not move stay where they are).  The walk's packed positions follow.
 * Timed as stage 7. */
int tdv_drift_to_sync_point(tdv_context *ctx, uint32_t it_next);

/* Inputs of set_new_particle_timebins (timestep.c:171-201). */
typedef struct {
	double time_int_accuracy;  /* Param.Time_Int_Accuracy                    */
	double softening;          /* Param.Grav_Softening[1]                     */
	double dt_max_global;      /* Dt_Max_Global (held as a float there)       */
	int max_active_bin;        /* Time.Max_Active_Bin                         */
	double expansion_factor;   /* Cosmo.Expansion_Factor (comoving)           */
	double hubble_parameter;   /* Cosmo.Hubble_Parameter (comoving)           */
	double grav_accel_factor;  /* Cosmo.Grav_Accel_Factor (comoving)          */
} tdv_timestep_params;

/* set_new_particle_timebins (timestep.c:171-201) on the device: every
 * particle's P.Time_Bin from its step (cosmological_timestep on |Grav_Acc|,
 * timestep.c:495-526, in the reference's float and double types, capped at
 * dt_max_global, timestep2timebin :380-388), no longer than
 * max(max_active_bin, its bin); *bin_min / *bin_max (may be NULL) get
 * Time_Bin_Min / Time_Bin_Max.  A step below Time.Step_Min, or not finite,
 * is TDV_ERR_RANGE (the reference's Assert, timestep.c:183-187).  The rest
 * of Set_New_Timesteps is scalar (set_system_timestep, max_active_time_bin)
 * and stays with the caller; tdv_make_active_list then gives the active
 * list. */
int tdv_set_new_timebins(tdv_context *ctx, const tdv_timestep_params *p,
		int *bin_min, int *bin_max);

/* Number of particles currently on the device. */
size_t tdv_num_particles(const tdv_context *ctx);

/* struct gadget_header (IO/io.h:23-42): the 256-byte Gadget header. */
typedef struct {
	uint32_t Npart[6];
	double Massarr[6];
	double Time;
	double Redshift;
	int32_t Flag_Sfr;
	int32_t Flag_Feedback;
	uint32_t Nall[6];
	int32_t Flag_Cooling;
	int32_t Num_Files;
	double Boxsize;
	double Omega0;
	double Omega_Lambda;
	double Hubble_Param;
	int32_t Flag_Age;
	int32_t Flag_Metals;
	uint32_t Nall_High_Word[6];
	char fill_bytes[64];           /* 59 in io.h + the struct's padding */
} tdv_snapshot_header;

#define TDV_BLOCK_ACC  1u          /* OUTPUT_TOTAL_ACCELERATION (Config:15)    */
#define TDV_BLOCK_GACC 2u          /* OUTPUT_PARTIAL_ACCELERATIONS (Config:16) */

/* write_file + write_gadget_header (IO/write_snapshot.c:63-219) for one
 * rank and one file: a Gadget format-2 snapshot of the context's particles
 * in their current order, blocks HEAD, POS, VEL, ID (uint32), MASS, then
 * ACC and GACC when `blocks` asks for them (IO/io.h:68-89 with
 * DOUBLE_PRECISION, no LONG_IDS).  ACC holds the x component of P.Acc: the
 * reference's block table gives it one component (io.h:76).  head->Npart
 * must add up to the particle count; Nall and Num_Files are set for one
 * file and the bytes the reference leaves unset are zero. */
int tdv_write_snapshot(tdv_context *ctx, const char *path,
		const tdv_snapshot_header *head, unsigned blocks);

/* Read_Snapshot (IO/read_snapshot.c:19-128) for one rank: the file `path`,
 * or path.0, path.1, ... (find_files, :545-574); HEAD, then POS, VEL, ID
 * (required) and MASS (optional, io.h:74) of each file in order; masses
 * from the header's Massarr when any entry is nonzero
 * (generate_masses_from_header, :421-446).  The particles go to the
 * context as tdv_set_particles puts them; *head (may be NULL) gets the
 * first file's header.  Files written in the other byte order are refused
 * (the reference's swap in safe_fread, :590-607, scrambles arrays). */
int tdv_read_snapshot(tdv_context *ctx, const char *path,
		tdv_snapshot_header *head);

#ifdef __cplusplus
}
#endif

#endif
