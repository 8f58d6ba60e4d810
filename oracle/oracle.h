/*
 * oracle.h -- CPU restatement of Tandav's gravity force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tandav_amd/, include/)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker and as the timed CPU
 * baseline.
 *
 * Each function restates the algorithm of the reference file:line given in
 * its comment (reference = jdonnert/Tandav, built with DOUBLE_PRECISION,
 * PERIODIC optional, GRAVITY_TREE, OUTPUT_PARTIAL_ACCELERATIONS,
 * VECTOR_SIZE 64, NTarget = 1).  The reference itself cannot be compiled in
 * this image (it needs GSL and a Makefile-generated config.h), so the
 * restatement is pinned by closed forms, invariants and brute force in
 * tests/ (see the test files named in each section).
 *
 * Parity status per function (how it is pinned):
 *   orc_peano_key / orc_reversed_peano_key : Hilbert-curve properties
 *       (bijection + face adjacency on 2^L grids), triplet-permutation
 *       identity between the two layouts, brute-force re-derivation.
 *   orc_qsort_index, orc_reorder_*         : sortedness + permutation checks.
 *   orc_domain_*                            : exact order statistic / max.
 *   orc_top_nodes                           : brute-force cell histogram.
 *   orc_tree_build                          : tree invariants (mass, CoM,
 *       leaf sizes, skip links) and an independent parallel re-derivation
 *       of the node set.  The exact collapse sequence is "parity unpinned"
 *       against the compiled reference (not buildable here).
 *   orc_walk (BH / relative / Ewald)        : brute-force direct sum bounds
 *       the error; exact interaction-set identity vs the reference is
 *       "parity unpinned" (reference not buildable here).
 *   orc_ewald_table                         : independent numpy evaluation
 *       of the same finite sums, odd symmetry, F(0)=0, the known value
 *       2.8372975 of the potential at r=0.  ulp-level libm differences
 *       are "parity unpinned".
 */
#ifndef TANDAV_ORACLE_H
#define TANDAV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned __int128 orc_key;   /* peanoKey (globals.h:36) */

/* Tree_Node (Gravity/tree.h:23-32), Float = double */
typedef struct {
	int DNext;
	uint32_t Bitfield;
	int DUp;
	int Npart;
	double Pos[3];
	double Mass;
	double CoM[3];
	double Dp[3];
} orc_node;

/* Top_Tree_Node part of Domain_Node_List (domain.h:40-52) */
typedef struct {
	uint64_t Key;
	int Target;
	int Level;
	int First_Part;
	int Npart;
	float Pos[3];
	float Mass;
	float CoM[3];
	float Dp[3];
} orc_topnode;

/* Run-time form of the Config / parameter-file knobs the path reads. */
typedef struct {
	int periodic;            /* PERIODIC (Config:7)                      */
	double boxsize;          /* Sim.Boxsize[0]                           */
	double gravity;          /* Const.Gravity                            */
	double softening;        /* Param.Grav_Softening[1] (Plummer equiv.) */
	double theta_bh;         /* TREE_OPEN_PARAM_BH (Config:47)           */
	double alpha_rel;        /* TREE_OPEN_PARAM_REL (Config:48)          */
	int vector_size;         /* VECTOR_SIZE (Config:54)                  */
	int nthreads;            /* OpenMP threads for the particle loops    */
} orc_params;

/* Particle columns (particles.h:17-37), caller-owned, length npart. */
typedef struct {
	double *pos[3];
	double *vel[3];
	double *acc[3];
	double *grav_acc[3];
	double *mass;
	double *last_acc_mag;
	double *cost;
	int *tree_parent;
	orc_key *key;
	uint32_t *id;            /* optional (may be NULL) */
	int *time_bin;           /* optional (may be NULL) */
} orc_particles;

typedef struct orc_sim orc_sim;

/* ---- keys (peano.c) ---------------------------------------------------- */
orc_key orc_peano_key(double px, double py, double pz,
		const double origin[3], double size);
orc_key orc_reversed_peano_key(double px, double py, double pz,
		const double origin[3], double size);
void orc_peano_keys(size_t n, const double *x, const double *y,
		const double *z, const double origin[3], double size, orc_key *out);
void orc_reversed_peano_keys(size_t n, const double *x, const double *y,
		const double *z, const double origin[3], double size, orc_key *out);

/* ---- sort + reorder (sort.c, aux.c) ------------------------------------ */
void orc_qsort_index(size_t n, const orc_key *keys, size_t *perm, int nthreads);
void orc_reorder_8(size_t n, void *p, const size_t *idx, size_t *scratch);
void orc_reorder_4(size_t n, void *p, const size_t *idx, size_t *scratch);
void orc_reorder_16(size_t n, void *p, const size_t *idx, size_t *scratch);

/* ---- domain (domain.c, select.c) --------------------------------------- */
double orc_select(int k, int ndata, double *data);
void orc_domain_nonperiodic(size_t n, const double *x, const double *y,
		const double *z, double center[3], double *size, double origin[3]);

/* ---- Ewald table + lookup (Gravity/periodic.c) -------------------------- */
void orc_ewald_table(double *fx, double *fy, double *fz, double *fp,
		int nthreads);

/* ---- the simulation-state restatement ----------------------------------- */
orc_sim *orc_sim_create(const orc_params *par, size_t npart,
		const orc_particles *P);
void orc_sim_free(orc_sim *s);
/* Ewald table in box units, unnormalised, 4 x 65^3 (Fx,Fy,Fz,Fp), as
 * written by write_ewald_correction_table (Gravity/periodic.c:325-342). */
void orc_sim_set_ewald_table(orc_sim *s, const double *table4);
/* active set = particles with time_bin <= max_active_bin (all if < 0),
 * rebuilt after every PH reorder (timestep.c:280-299) */
void orc_sim_set_active_bin(orc_sim *s, int max_active_bin);
int orc_sim_nactive(const orc_sim *s);
const int *orc_sim_active(const orc_sim *s);
void orc_sim_setup_domain(orc_sim *s);          /* Setup_Domain_Decomposition */
void orc_sim_domain_decomposition(orc_sim *s);  /* Domain_Decomposition       */
void orc_sim_set_top_nodes(orc_sim *s, const orc_topnode *rec, int ntop);

/* Domain_Decomposition with NTarget = ntarget (domain.c:67-145) on
 * particles whose short keys (upper 64 bits) are sorted ascending, with
 * their P.Cost: the bunch list in descending key order as top-node records
 * (Key, Target = -task - 1, Level, First_Part, Npart; Pos left 0).
 * Returns 0, -1 on a bad argument or more than max_out bunches, -2 when
 * fewer non-empty bunches than targets remain. */
int orc_domain_targets(const uint64_t *key, const double *cost, size_t n,
		int ntarget, double part_alloc_factor, double imbal_ceil,
		orc_topnode *out, int max_out, int *ntop, int *iterations);
/* The same on a simulation (set_computational_domain, the key sort, the
 * loop on its P.Cost, the bunches as top nodes); tasks: >= 65536 ints. */
int orc_sim_domain_decomposition_targets(orc_sim *s, int ntarget,
		double part_alloc_factor, double imbal_ceil, int *tasks,
		int *iterations);
void orc_sim_tree_build(orc_sim *s);            /* Gravity_Tree_Build         */
void orc_sim_tree_acceleration(orc_sim *s, int use_bh); /* + periodic part */
void orc_sim_compute_acceleration(orc_sim *s, int prepare_step);
void orc_sim_ewald_correction(const orc_sim *s, const double dr[3],
		double f[3]);
void orc_sim_periodic_nearest(const orc_sim *s, double dr[3]);

/* accessors */
void orc_sim_domain(const orc_sim *s, double *size, double origin[3],
		double center[3]);
int orc_sim_ntop(const orc_sim *s);
const orc_topnode *orc_sim_top(const orc_sim *s);
int orc_sim_nnodes(const orc_sim *s);
const orc_node *orc_sim_tree(const orc_sim *s);
int orc_sim_ndumped(const orc_sim *s);
double orc_sim_epsilon(const orc_sim *s);

/* ---- FP64 direct summation (the error oracle) --------------------------- */
/* Softened-monopole direct sum with the path's own pairwise kernel
 * (tree_accel.c:362-394).  Periodic runs add the tabulated Ewald
 * correction of every pair (nearest image + correction, as the tree does
 * for far nodes).  Targets idx[0..nt-1]; writes acc[3*t]. */
void orc_direct_sum(const orc_sim *s, int nt, const int *idx, double *acc);

#ifdef __cplusplus
}
#endif

#endif
