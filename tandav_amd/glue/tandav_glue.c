/*
 * tandav_glue.c -- Tandav's force path on the MI355X engine, as Tandav's own
 * C symbols.
 *
 * This file is the reference-side binding.  Built into Tandav in place of
 * accel.c, peano.c, Gravity/tree_build.c, Gravity/tree_accel.c,
 * Gravity/tree_periodic.c, Gravity/tree_update.c, Gravity/periodic.c and
 * Gravity/forcetest.c, it exports the same functions with the same
 * signatures and leaves the same host-visible state behind (P permuted and
 * P.Key reversed, P.Acc, P.Grav_Acc, P.Cost, P.Last_Acc_Mag, P.Tree_Parent,
 * D[].TNode.Target/Mass/CoM, Tree, NNodes), while keys, sort, reorder, tree
 * build, walks and Ewald corrections run on the GPU through tandav_gpu.h.
 *
 * Threading.  Tandav calls its team functions (Compute_Acceleration,
 * Sort_Particles_By_Peano_Key, Reverse_Peano_Keys, Gravity_Tree_*) from
 * every thread of the run-long OpenMP team (main.c:37).  Each does its
 * device work inside "omp single", whose implicit barrier every thread
 * passes; the threadprivate Sig (signal.h:23) is read on the thread that
 * runs the single and reset on every thread where the reference resets it.
 * The per-element functions (Peano_Key, Reversed_Peano_Key,
 * Ewald_Correction, ...) compute on the host with the engine's own code
 * (tdv_peano_keys_host, tdv_ewald_correction_host): no device round trip,
 * any thread at once.
 *
 * Errors.  A failing engine call reports the engine's message through
 * Assert (macro.h:15, Assert_Info aux.c:49-77), the reference's own abort
 * path; no C++ exception crosses tandav_gpu.h.
 *
 * Build: Tandav with DOUBLE_PRECISION (the engine's columns are FP64) and
 * VECTOR_SIZE 64.  Run-time knobs from the environment: TANDAV_DEVICE (HIP
 * device ordinal, default 0), TANDAV_WALK=fp32 (the FP32 group walk instead
 * of the reference's FP64 walk).
 */
#ifdef TANDAV_GLUE_ENV
#include TANDAV_GLUE_ENV   /* a host program other than Tandav */
#else
#include "includes.h"
#include "accel.h"
#include "domain.h"
#include "peano.h"
#include "timestep.h"
#include "Gravity/tree.h"
#include "Gravity/periodic.h"
#include "Gravity/forcetest.h"
#endif

#include "tandav_gpu.h"

#if !defined(GRAVITY) || !defined(GRAVITY_TREE)
#error "the engine replaces the tree gravity path (GRAVITY + GRAVITY_TREE)"
#endif
#ifdef GRAVITY_POTENTIAL
#error "the engine computes accelerations only (no GRAVITY_POTENTIAL)"
#endif
#ifndef DOUBLE_PRECISION
#error "the engine holds FP64 columns: build Tandav with DOUBLE_PRECISION"
#endif
#if VECTOR_SIZE != 64
#error "the engine's bundles hold at most 64 particles (VECTOR_SIZE 64)"
#endif

/* D[] is handed to the engine as its 64-byte Top_Tree_Node records */
typedef char tdv_glue_D_is_64_bytes[(sizeof(*D) == 64) ? 1 : -1];
typedef char tdv_glue_Tree_is_96_bytes[(sizeof(*Tree) == 96) ? 1 : -1];
typedef char tdv_glue_key_is_16_bytes[(sizeof(peanoKey) == 16) ? 1 : -1];

struct Tree_Node * restrict Tree = NULL;   /* Gravity/tree_build.c:27 */

static tdv_context *Ctx = NULL;
static size_t Max_Nodes = 0;      /* host Tree capacity in nodes */
static bool Moments_Moved = false; /* Tree / D[] CoM drifted on the host */

#define TDV_CHECK(call) do { \
	const int rc_ = (call); \
	Assert(rc_ == TDV_OK, "tandav engine: %s failed (%d): %s", #call, rc_, \
			tdv_last_error(Ctx)); \
} while (0)

/* The engine context, made on first use: parameters come from the same
 * globals and Config macros Setup_Gravity_Tree (Gravity/tree_build.c:
 * 124-142) and Gravity_Periodic_Init (Gravity/periodic.c:180-233) read. */
static tdv_context *context(void)
{
	if (Ctx)
		return Ctx;
	tdv_params p;
	memset(&p, 0, sizeof(p));
#ifdef PERIODIC
	p.periodic = 1;
	p.boxsize = Sim.Boxsize[0];
#endif
	p.gravity = Const.Gravity;
	p.softening = Param.Grav_Softening[1];
	p.theta_bh = TREE_OPEN_PARAM_BH;
	p.alpha_rel = TREE_OPEN_PARAM_REL;
	p.vector_size = VECTOR_SIZE;
	p.compat = 1;
	const char *dev = getenv("TANDAV_DEVICE");
	p.device = dev ? atoi(dev) : 0;
	const char *walk = getenv("TANDAV_WALK");
	p.precision = (walk && !strcmp(walk, "fp32")) ? TDV_WALK_FP32
			: TDV_WALK_FP64;
	p.ewald_size = 0.125;
	p.ewald_accuracy = 0.1;
	if (p.precision == TDV_WALK_FP32) {
		/* the FP32 group walk's own settings, the ones its accuracy is
		 * stated for (p99 |da|/|a| <= 2e-3 against the exact periodic force
		 * on the 128^3 and 256^3 boxes): the next-order relative criterion
		 * M s^3 > r^5 |a| alpha / G at alpha 0.01, quadrupoles, the Ewald
		 * correction at centres of mass to second order (ewald_accuracy 1);
		 * TANDAV_FP32_ALPHA, TANDAV_FP32_CRITERION (2 or 3) and
		 * TANDAV_FP32_EWALD_ACCURACY override them */
		const char *a = getenv("TANDAV_FP32_ALPHA");
		const char *c = getenv("TANDAV_FP32_CRITERION");
		const char *e = getenv("TANDAV_FP32_EWALD_ACCURACY");
		p.alpha_rel = a ? atof(a) : 0.01;
		p.fp32_criterion = c ? atoi(c) : 3;
		p.ewald_accuracy = e ? atof(e) : 1.0;
		p.fp32_order = 2;
		p.compat = 0;
	}
	const int rc = tdv_create(&p, &Ctx);
	Assert(rc == TDV_OK, "tandav engine: tdv_create failed (%d) on device %d",
			rc, p.device);
	return Ctx;
}

/* host columns the engine does not keep, permuted on the device with the
 * sort's permutation (reorder_collisionless_particles, peano.c:50-82) */
static void permute_host_only_columns(tdv_context *c)
{
	void *cols[12];
	int bytes[12];
	int k = 0;
	cols[k] = P.Type;         bytes[k++] = sizeof(*P.Type);
	cols[k] = P.It_Drift_Pos; bytes[k++] = sizeof(*P.It_Drift_Pos);
	cols[k] = P.It_Kick_Pos;  bytes[k++] = sizeof(*P.It_Kick_Pos);
	cols[k] = P.Cost;         bytes[k++] = sizeof(*P.Cost);
	for (int d = 0; d < 3; d++) {
		cols[k] = P.Grav_Acc[d];
		bytes[k++] = sizeof(*P.Grav_Acc[d]);
	}
	cols[k] = P.Tree_Parent;  bytes[k++] = sizeof(*P.Tree_Parent);
	if (sizeof(ID_t) != 4) {
		cols[k] = P.ID;
		bytes[k++] = sizeof(ID_t);
	}
	TDV_CHECK(tdv_permute_host_columns(c, k, cols, bytes));
}

/*
 * Keys
 */

/* peano.c:6-12 */
int cmp_peanoKeys(const void *a, const void *b)
{
	const peanoKey x = *(const peanoKey *)a, y = *(const peanoKey *)b;
	return (x > y) - (x < y);
}

/* Sort_Particles_By_Peano_Key (peano.c:20-48): PH keys of the current
 * positions in the current Domain, the key sort and the reorder of every P
 * column, all on the device; P.Key is left holding the standard keys. */
void Sort_Particles_By_Peano_Key(void)
{
	#pragma omp single
	{
	tdv_context *c = context();
	const size_t n = Task.Npart_Total;
	const double *pos[3] = { P.Pos[0], P.Pos[1], P.Pos[2] };
	const double *vel[3] = { P.Vel[0], P.Vel[1], P.Vel[2] };
	const double *acc[3] = { P.Acc[0], P.Acc[1], P.Acc[2] };
	const uint32_t *id = sizeof(ID_t) == 4 ? (const uint32_t *)P.ID : NULL;
	TDV_CHECK(tdv_set_particles(c, n, pos, vel, P.Mass, acc, P.Last_Acc_Mag,
			P.Time_Bin, id));
	TDV_CHECK(tdv_set_domain(c, Domain.Size, Domain.Origin));
	TDV_CHECK(tdv_sort_particles(c));
	double *pos_w[3] = { P.Pos[0], P.Pos[1], P.Pos[2] };
	double *vel_w[3] = { P.Vel[0], P.Vel[1], P.Vel[2] };
	double *acc_w[3] = { P.Acc[0], P.Acc[1], P.Acc[2] };
	TDV_CHECK(tdv_get_particles(c, pos_w, vel_w, P.Mass, acc_w, NULL, NULL,
			P.Last_Acc_Mag, NULL, NULL, (uint32_t *)(sizeof(ID_t) == 4
			? (void *)P.ID : NULL), P.Time_Bin));
	TDV_CHECK(tdv_get_keys(c, (uint64_t *)P.Key, 0));
	permute_host_only_columns(c);
	} // omp single

	Make_Active_Particle_List();
}

/* Reverse_Peano_Keys (peano.c:84-97): P.Key = reversed keys. */
void Reverse_Peano_Keys(void)
{
	#pragma omp single
	TDV_CHECK(tdv_get_keys(context(), (uint64_t *)P.Key, 1));
}

static peanoKey one_key(const Float px, const Float py, const Float pz,
		const int reversed)
{
	uint64_t w[2] = { 0, 0 };
	const int rc = tdv_peano_keys_host(1, &px, &py, &pz, Domain.Origin,
			Domain.Size, reversed, w);
	Assert(rc == TDV_OK, "tandav engine: tdv_peano_keys_host failed (%d): "
			"Domain.Size %g", rc, Domain.Size);
	return ((peanoKey)w[1] << 64) | w[0];
}

/* Peano_Key (peano.c:122-195), on the host with the engine's transform */
peanoKey Peano_Key(const Float px, const Float py, const Float pz)
{
	return one_key(px, py, pz, 0);
}

/* Reversed_Peano_Key (peano.c:203-276) */
peanoKey Reversed_Peano_Key(const Float px, const Float py, const Float pz)
{
	return one_key(px, py, pz, 1);
}

/* Short_Peano_Key (peano.c:286-289): the top 21 triplets */
shortKey Short_Peano_Key(const Float px, const Float py, const Float pz)
{
	return (shortKey)(Peano_Key(px, py, pz) >> DELTA_PEANO_BITS);
}

/* Reversed_Short_Peano_Key (peano.c:292-297): the reference keeps the low
 * 32 bits of the reversed key (its mask is 0x00000000FFFFFFFF). */
shortKey Reversed_Short_Peano_Key(const Float px, const Float py,
		const Float pz)
{
	return (shortKey)(Reversed_Peano_Key(px, py, pz) & 0xFFFFFFFFu);
}

/* Reverse_Peano_Key (peano.c:278-296), which Tandav marks buggy and never
 * calls: kept value for value.  Step s (s = 0..20, i = 125 - 6 s) ORs the
 * triplet at bits 125-3s.. down by i and the triplet at bits 2+3s.. up by
 * i-4 into the result, which is finally shifted up by one triplet. */
peanoKey Reverse_Peano_Key(const peanoKey pkey)
{
	peanoKey out = 0;
	for (int s = 0; s <= 20; s++) {
		const int i = 125 - 6 * s;
		const peanoKey hi_mask = (peanoKey)0x7 << (125 - 3 * s);
		const peanoKey lo_mask = (peanoKey)0x7 << (2 + 3 * s);
		out |= (pkey & hi_mask) >> i;
		out |= (pkey & lo_mask) << (i - 4);
	}
	return out << 3;
}

/* Test_Peanokey (peano.c:315-373): prints the keys of a corner point and
 * of the 4^3 cell centres of the unit cube, then ends the run. */
void Test_Peanokey(void)
{
	const double d = 0.25;
	double a = 0.9999999999;
	peanoKey s = Peano_Key(a, a, a), r = Reversed_Peano_Key(a, a, a);
	printf("a0=%lg a1=%lg a2=%lg std %016llx%016llx rev %016llx%016llx\n",
			a, a, a, (unsigned long long)(s >> 64), (unsigned long long)s,
			(unsigned long long)(r >> 64), (unsigned long long)r);
	for (int i = 0; i < 4; i++)
	for (int j = 0; j < 4; j++)
	for (int k = 0; k < 4; k++) {
		const double x = (i + 0.5) * d, y = (j + 0.5) * d, z = (k + 0.5) * d;
		s = Peano_Key(x, y, z);
		r = Reversed_Peano_Key(x, y, z);
		printf("%g %g %g std %016llx%016llx rev %016llx%016llx\n", x, y, z,
				(unsigned long long)(s >> 64), (unsigned long long)s,
				(unsigned long long)(r >> 64), (unsigned long long)r);
	}
	exit(0);
}

/*
 * Tree
 */

/* Setup_Gravity_Tree (Gravity/tree_build.c:124-142) */
void Setup_Gravity_Tree(void)
{
	Max_Nodes = 0.3 * Task.Npart_Total;
	if (Max_Nodes < 1)
		Max_Nodes = 1;
	Tree = malloc(Max_Nodes * sizeof(*Tree));
	Assert(Tree != NULL, "no memory for %zu tree nodes", Max_Nodes);
	for (int i = 0; i < NPARTYPE; i++) { /* K1 softening, Q1 sign kept */
		Epsilon[i] = -41.0 / 32.0 * Param.Grav_Softening[i];
		Epsilon2[i] = Epsilon[i] * Epsilon[i];
		Epsilon3[i] = Epsilon[i] * Epsilon[i] * Epsilon[i];
	}
}

/* Gravity_Tree_Free (Gravity/tree_build.c:144-156): the host mirror only;
 * the device tree is rebuilt by the next Gravity_Tree_Build. */
void Gravity_Tree_Free(void)
{
	#pragma omp single
	{
	free(Tree);
	Tree = NULL;
	Max_Nodes = 0;
	} // omp single
}

/* the build proper, on one thread */
static void build_tree(tdv_context *c)
{
	TDV_CHECK(tdv_set_top_nodes(c, D, (int)NTop_Nodes));
	TDV_CHECK(tdv_tree_build(c));
	size_t nn = 0;
	TDV_CHECK(tdv_get_tree(c, NULL, 0, &nn));
	if (nn > Max_Nodes || !Tree) {
		const size_t want = nn > Max_Nodes ? nn : (Max_Nodes ? Max_Nodes : 1);
		struct Tree_Node *t = realloc(Tree, want * sizeof(*Tree));
		Assert(t != NULL, "no memory for %zu tree nodes", want);
		Tree = t;
		Max_Nodes = want;
	}
	TDV_CHECK(tdv_get_tree(c, Tree, Max_Nodes, &nn));
	NNodes = (uint32_t)nn;

	struct Top_Tree_Node *tn = malloc(NTop_Nodes * sizeof(*tn));
	Assert(tn != NULL, "no memory for %u top nodes", NTop_Nodes);
	int ntop = 0;
	TDV_CHECK(tdv_get_top_nodes(c, tn, (int)NTop_Nodes, &ntop));
	for (int i = 0; i < ntop; i++) {
		D[i].TNode.Target = tn[i].Target;
		D[i].TNode.Mass = tn[i].Mass;
		for (int d = 0; d < 3; d++)
			D[i].TNode.CoM[d] = tn[i].CoM[d];
	}
	free(tn);
	TDV_CHECK(tdv_get_particles(c, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
			P.Tree_Parent, NULL, NULL, NULL));
	Moments_Moved = false;
}

/* Gravity_Tree_Build (Gravity/tree_build.c:47-122): the octree below the
 * caller's top nodes D[], from the keys of the last sort. */
void Gravity_Tree_Build(void)
{
	#pragma omp single
	build_tree(context());

	Sig.Tree_Update = false;
}

/* positions may have drifted and P.Acc / P.Last_Acc_Mag are the host's;
 * the active set is the host's Active_Particle_List */
static void upload_step_state(tdv_context *c, const bool last_acc)
{
	const double *pos[3] = { P.Pos[0], P.Pos[1], P.Pos[2] };
	const double *acc[3] = { P.Acc[0], P.Acc[1], P.Acc[2] };
	TDV_CHECK(tdv_update_particles(c, pos, P.Mass, acc,
			last_acc ? P.Last_Acc_Mag : NULL));
	TDV_CHECK(tdv_set_active(c, NActive_Particles, Active_Particle_List));
	if (Moments_Moved) {
		TDV_CHECK(tdv_set_tree_moments(c, Tree, NNodes, D, (int)NTop_Nodes));
		Moments_Moved = false;
	}
}

static void download_forces(tdv_context *c, const bool last_acc)
{
	double *acc[3] = { P.Acc[0], P.Acc[1], P.Acc[2] };
	double *gacc[3] = { P.Grav_Acc[0], P.Grav_Acc[1], P.Grav_Acc[2] };
	TDV_CHECK(tdv_get_particles(c, NULL, NULL, NULL, acc, gacc, P.Cost,
			last_acc ? P.Last_Acc_Mag : NULL, NULL, NULL, NULL, NULL));
}

/* check_total_momentum (Gravity/tree_accel.c:442-469), a host diagnostic
 * over the host's P.Vel (the kicks run on the host): sum m v of the type-1
 * particles into running sums that, as in Tandav, are never reset, and the
 * relative change of their norm since the last call. */
static double Mom_Px = 0, Mom_Py = 0, Mom_Pz = 0, Mom_Last = 0;

static void check_total_momentum(const bool show_change)
{
	const int first = Task.Npart[0], last = Task.Npart[0] + Task.Npart[1];

	#pragma omp for reduction(+: Mom_Px, Mom_Py, Mom_Pz)
	for (int i = first; i < last; i++) {
		Mom_Px += P.Mass[i] * P.Vel[0][i];
		Mom_Py += P.Mass[i] * P.Vel[1][i];
		Mom_Pz += P.Mass[i] * P.Vel[2][i];
	}

	const double ptotal = sqrt(Mom_Px * Mom_Px + Mom_Py * Mom_Py
			+ Mom_Pz * Mom_Pz);
	if (show_change)
		rprintf("Total change in momentum due to gravity: %g \n",
				(ptotal - Mom_Last) / Mom_Last);

	#pragma omp barrier
	#pragma omp single
	Mom_Last = ptotal;
}

/* Gravity_Tree_Acceleration (Gravity/tree_accel.c:32-91), which also runs
 * Gravity_Tree_Periodic (tree_accel.c:84): the walk of the active particles
 * with the BH (Sig.Use_BH_Criterion) or relative criterion, then the Ewald
 * walk on periodic runs. */
void Gravity_Tree_Acceleration(void)
{
	rprintf("Tree acceleration ");
	check_total_momentum(false);

	#pragma omp single
	{
	tdv_context *c = context();
	upload_step_state(c, true);
	TDV_CHECK(tdv_tree_acceleration(c, Sig.Use_BH_Criterion ? 1 : 0));
	download_forces(c, false);
	} // omp single

	rprintf(" done \n");
	check_total_momentum(true);
}

/* Gravity_Tree_Periodic (Gravity/tree_periodic.c:23-72): the engine runs
 * the Ewald walk inside Gravity_Tree_Acceleration, its only caller in
 * Tandav, so there is nothing left to do here. */
void Gravity_Tree_Periodic(void)
{
}

/* Gravity_Tree_Update_Kicks (Gravity/tree_update.c:12-57): the particle's
 * momentum change, per unit node mass, into Dp of every node above it up
 * to its subtree root (marked UPDATED), then into its top node's Dp (marked
 * by a negative Level).  Host mirror of Tree and D[]; called concurrently. */
void Gravity_Tree_Update_Kicks(const int ipart, const double dt)
{
	const Float m_dt = P.Mass[ipart] * dt;
	const Float dp[3] = { m_dt * P.Acc[0][ipart], m_dt * P.Acc[1][ipart],
			m_dt * P.Acc[2][ipart] };
	int node = P.Tree_Parent[ipart], top;
	if (node >= 0) {
		while (!(Tree[node].Bitfield & (1UL << TOP))) {
			for (int d = 0; d < 3; d++) {
				#pragma omp atomic update
				Tree[node].Dp[d] += dp[d] / Tree[node].Mass;
			}
			#pragma omp atomic update
			Tree[node].Bitfield |= 1UL << UPDATED;
			node -= Tree[node].DUp;
		}
		top = Tree[node].DUp;      /* a subtree root's DUp is its top node */
	} else {
		top = -node - 1;
	}
	#pragma omp critical (tandav_topnode_mark)
	if (D[top].TNode.Level > 0)
		D[top].TNode.Level = -D[top].TNode.Level;
	for (int d = 0; d < 3; d++) {
		#pragma omp atomic update
		D[top].TNode.Dp[d] += dp[d] / D[top].TNode.Mass;
	}
}

/* Gravity_Tree_Update_Drift (Gravity/tree_update.c:64-104): move the CoM
 * of every kicked node and top node by dt * Dp and clear the marks; the
 * moved monopoles reach the device before the next walk. */
void Gravity_Tree_Update_Drift(const double dt)
{
	static int moved = 0;

	#pragma omp single
	moved = 0;

	#pragma omp for nowait
	for (int i = 0; i < (int)NNodes; i++) {
		if (!(Tree[i].Bitfield & (1UL << UPDATED)))
			continue;
		for (int d = 0; d < 3; d++) {
			Tree[i].CoM[d] += dt * Tree[i].Dp[d];
			Tree[i].Dp[d] = 0;
		}
		Tree[i].Bitfield &= ~(1UL << UPDATED);
	}

	#pragma omp for reduction(+:moved)
	for (int i = 0; i < (int)NTop_Nodes; i++) {
		if (D[i].TNode.Level >= 0)
			continue;
		for (int d = 0; d < 3; d++) {
			D[i].TNode.CoM[d] += dt * D[i].TNode.Dp[d];
			D[i].TNode.Dp[d] = 0;
		}
		D[i].TNode.Level = -D[i].TNode.Level;
		moved++;
	}

	#pragma omp single
	Moments_Moved = true;

	rprintf("Tree update: Moved %d top nodes \n", moved);
}

/* Tree bitfield accessors (Gravity/tree_accel.c:400-429) on the host
 * mirror. */
Float Node_Size(const int node)
{
	const int lvl = Tree[node].Bitfield & 0x3F;
	return Domain.Size / ((Float)(1ULL << lvl));
}

int Level(const int node)
{
	return Tree[node].Bitfield & 0x3F;
}

bool Node_Is(const enum Tree_Bitfield bit, const int node)
{
	return Tree[node].Bitfield & (1UL << bit);
}

void Node_Set(const enum Tree_Bitfield bit, const int node)
{
	Tree[node].Bitfield |= 1UL << bit;
}

void Node_Clear(const enum Tree_Bitfield bit, const int node)
{
	Tree[node].Bitfield &= ~(1UL << bit);
}

/*
 * Ewald
 */

#ifdef PERIODIC
/* Gravity_Periodic_Init (Gravity/periodic.c:180-233): the engine computes
 * the FP64 Ewald table on the GPU when the context is made; as the
 * reference, the table is read from ./ewald_tables.dat when that file
 * exists (read_ewald_correction_table, :303-323: Fx, Fy, Fz, Fp, each
 * (N_EWALD+1)^3 Float in box units, raw) and written there otherwise
 * (write_ewald_correction_table, :325-342). */
#define EWALD_FILE "./ewald_tables.dat"
#define EWALD_NODES (65 * 65 * 65)

void Gravity_Periodic_Init(void)
{
	tdv_context *c = context();
	double *tab = malloc(4 * (size_t)EWALD_NODES * sizeof(*tab));
	Assert(tab != NULL, "no memory for the Ewald table");
	FILE *fp = fopen(EWALD_FILE, "r");
	if (fp) {
		const size_t got = fread(tab, sizeof(*tab), 4 * (size_t)EWALD_NODES,
				fp);
		fclose(fp);
		Assert(got == 4 * (size_t)EWALD_NODES, "short read of %s (%zu of %zu "
				"values)", EWALD_FILE, got, 4 * (size_t)EWALD_NODES);
		TDV_CHECK(tdv_set_ewald_table(c, tab));
	} else {
		TDV_CHECK(tdv_get_ewald_table(c, tab));
		fp = fopen(EWALD_FILE, "w");
		if (fp) {
			const size_t put = fwrite(tab, sizeof(*tab),
					4 * (size_t)EWALD_NODES, fp);
			fclose(fp);
			Assert(put == 4 * (size_t)EWALD_NODES, "short write of %s",
					EWALD_FILE);
		}
	}
	free(tab);
}

/* Ewald_Correction (Gravity/periodic.c:35-113), one separation looked up
 * on the host in the engine's copy of the table (set by
 * Gravity_Periodic_Init). */
void Ewald_Correction(const Float dr[3], Float f[3])
{
	TDV_CHECK(tdv_ewald_correction_host(context(), 1, dr, f));
}
#endif // PERIODIC

/*
 * Driver
 */

/* Compute_Acceleration (accel.c:15-71): |a| of the last step into
 * Last_Acc_Mag and P.Acc cleared for the active particles, the tree built
 * when Sig.Tree_Update is set, on the prepare step a BH walk followed by a
 * relative walk, else the relative walk, each with its Ewald walk. */
void Compute_Acceleration(void)
{
	#pragma omp single
	upload_step_state(context(), false);

	if (Sig.Tree_Update)
		Gravity_Tree_Build();

	#pragma omp single
	{
	tdv_context *c = context();
	TDV_CHECK(tdv_compute_acceleration(c, Sig.Prepare_Step ? 1 : 0));
	download_forces(c, true);
	} // omp single

	/* Gravity_Tree_Acceleration's output and momentum check for each walk
	 * (BH + relative on the prepare step), in Tandav's order: the walks do
	 * not change velocities, so running the checks after the device work
	 * prints the same lines and values */
	const int nwalks = Sig.Prepare_Step ? 2 : 1;
	for (int w = 0; w < nwalks; w++) {
		rprintf("Tree acceleration ");
		check_total_momentum(false);
		rprintf(" done \n");
		check_total_momentum(true);
	}

#ifdef GRAVITY_FORCETEST
	Gravity_Forcetest();
#endif
}

#ifdef GRAVITY_FORCETEST
#define NTEST 10

/* Gravity_Forcetest (Gravity/forcetest.c:24-174): FP64 direct summation
 * (with the Ewald correction of every pair on periodic runs) for NTEST
 * random active particles, on the device, against P.Acc, printed in the
 * reference's table. */
void Gravity_Forcetest(void)
{
	#pragma omp single
	{
	tdv_context *c = context();
	int idx[NTEST];
	double direct[3 * NTEST];
	for (int i = 0; i < NTEST; i++)
		idx[i] = Active_Particle_List[(int)(erand48(Task.Seed)
				* NActive_Particles)];
	TDV_CHECK(tdv_forcetest(c, NTEST, idx, direct));
	double mean = 0, max = 0;
	int worst = -1;
	rprintf("Direct Gravity, N = %d \n      ID          Errors                 "
			"Tree Accel.                      Direct Accel. \n", NTEST);
	for (int i = 0; i < NTEST; i++) {
		const int ip = idx[i];
		const double *dir = direct + 3 * i;
		double e1 = 0, a_dir = 0, a_tree = 0;
		for (int d = 0; d < 3; d++) {
			const double e = (P.Acc[d][ip] - dir[d]) / dir[d];
			e1 += e * e;
			a_dir += dir[d] * dir[d];
			a_tree += P.Acc[d][ip] * P.Acc[d][ip];
		}
		const double e2 = (sqrt(a_tree) - sqrt(a_dir)) / sqrt(a_dir);
		mean += e2;
		if (e2 > max) {
			max = e2;
			worst = ip;
		}
		printf("   %9d  %05.4f %+05.4f   %+09.3f %+09.3f %+09.3f	"
				"%+09.3f %+09.3f %+09.3f\n", (int)P.ID[ip], sqrt(e1), e2,
				P.Acc[0][ip], P.Acc[1][ip], P.Acc[2][ip], dir[0], dir[1],
				dir[2]);
	}
	rprintf("\nForce test: NActive %d, max error %g @ %d, mean error %g \n\n",
			NActive_Particles, max, worst, mean / NTEST);
	} // omp single
}
#endif // GRAVITY_FORCETEST
