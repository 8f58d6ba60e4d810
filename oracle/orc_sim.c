/*
 * Simulation-state restatement of the force path: domain + top nodes
 * (domain.c, select.c), PH ordering (peano.c), tree build
 * (Gravity/tree_build.c), tree walks (Gravity/tree_accel.c,
 * Gravity/tree_periodic.c), Ewald lookup (Gravity/periodic.c:35-113),
 * periodic wrap (periodic.c:62-80) and the driver (accel.c:15-71).
 * TEST INFRASTRUCTURE (see oracle.h).
 *
 * The restatement runs the reference's decisions in the reference's order
 * and keeps its quirks (negative K1 softening length, small-top-node
 * particle range, float top nodes, sign-placed node centres, squared Ewald
 * BH threshold, any-zero-component Ewald skip).  It is single-rank with
 * NTarget = 1 and every subtree built in the thread buffer (top nodes with
 * fewer than BufferSize/96 particles), which is the configuration whose
 * output does not depend on the OpenMP schedule.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <limits.h>
#include <float.h>
#include <omp.h>
#include "oracle.h"

#define N_PEANO_TRIPLETS 42
#define N_EWALD 64
#define NE1 (N_EWALD + 1)
#define TOP_LEVEL 2          /* MIN_LEVEL, domain.c:3, with NTarget = 1 */
#define BIT_LOCAL 9
#define BIT_TOP 10

struct orc_sim {
	orc_params par;
	double Epsilon, Epsilon2, Epsilon3;
	double dom_size, dom_origin[3], dom_center[3];
	size_t npart;
	orc_particles P;
	int nactive;
	int *active;
	int max_active_bin;       /* < 0: every particle is active */
	int ntop;
	orc_topnode *D;
	orc_node *Tree;
	int nnodes;
	int ndumped;
	double *ewald;            /* Fx,Fy,Fz,Fp, each NE1^3, already / Box^k */
	double box2ewald;
	double boxhalf;
};

static void fatal(const char *msg)
{
	fprintf(stderr, "oracle: %s\n", msg);
	abort();
}

/* ------------------------------------------------------------------------ */
/* parameters, lifetime                                                      */
/* ------------------------------------------------------------------------ */

orc_sim *orc_sim_create(const orc_params *par, size_t npart,
		const orc_particles *P)
{
	orc_sim *s = calloc(1, sizeof(*s));
	s->par = *par;
	if (s->par.nthreads < 1)
		s->par.nthreads = 1;
	s->npart = npart;
	s->P = *P;
	/* Setup_Gravity_Tree, tree_build.c:134-139: K1 length, negative */
	s->Epsilon = -41.0 / 32.0 * par->softening;
	s->Epsilon2 = s->Epsilon * s->Epsilon;
	s->Epsilon3 = s->Epsilon * s->Epsilon * s->Epsilon;
	s->D = calloc(64, sizeof(*s->D));
	s->active = malloc((npart ? npart : 1) * sizeof(int));
	s->max_active_bin = -1;
	s->nactive = (int)npart;
	for (size_t i = 0; i < npart; i++)
		s->active[i] = (int)i;
	if (par->periodic) {
		s->boxhalf = par->boxsize / 2.0;             /* periodic.c:12 */
		s->box2ewald = 2 * N_EWALD / par->boxsize;   /* periodic.c:188 */
	}
	return s;
}

void orc_sim_free(orc_sim *s)
{
	if (!s)
		return;
	free(s->D);
	free(s->active);
	free(s->Tree);
	free(s->ewald);
	free(s);
}

void orc_sim_set_ewald_table(orc_sim *s, const double *t4)
{
	/* Gravity_Periodic_Init, periodic.c:215-228: scale by 1/Box^2, 1/Box */
	const size_t n3 = (size_t)NE1 * NE1 * NE1;
	const double box = s->par.boxsize;
	free(s->ewald);
	s->ewald = malloc(4 * n3 * sizeof(double));
	for (size_t i = 0; i < n3; i++) {
		s->ewald[i] = t4[i] / (box * box);
		s->ewald[n3 + i] = t4[n3 + i] / (box * box);
		s->ewald[2 * n3 + i] = t4[2 * n3 + i] / (box * box);
		s->ewald[3 * n3 + i] = t4[3 * n3 + i] / box;
	}
}

/* Make_Active_Particle_List, timestep.c:280-299 */
static void make_active_list(orc_sim *s)
{
	int n = 0;
	for (size_t i = 0; i < s->npart; i++)
		if (s->max_active_bin < 0 || !s->P.time_bin
				|| s->P.time_bin[i] <= s->max_active_bin)
			s->active[n++] = (int)i;
	s->nactive = n;
}

void orc_sim_set_active_bin(orc_sim *s, int max_active_bin)
{
	s->max_active_bin = max_active_bin;
	make_active_list(s);
}

int orc_sim_nactive(const orc_sim *s) { return s->nactive; }
const int *orc_sim_active(const orc_sim *s) { return s->active; }

/* ------------------------------------------------------------------------ */
/* domain (domain.c:732-847, select.c:19-80)                                 */
/* ------------------------------------------------------------------------ */

static inline void dswap(double *a, double *b)
{
	double t = *a;
	*a = *b;
	*b = t;
}

/* Select, select.c:19-80: Numerical-Recipes quickselect over data[1..n-1]
 * (element 0 takes no part, the 1-based origin of the routine). */
double orc_select(int k, int ndata, double *data)
{
	int l = 1, ir = ndata - 1;
	for (;;) {
		if (ir <= l + 1) {
			if (ir == l + 1 && data[ir] < data[l])
				dswap(&data[l], &data[ir]);
			break;
		}
		int mid = (l + ir) >> 1;
		dswap(&data[mid], &data[l + 1]);
		if (data[l] > data[ir])
			dswap(&data[l], &data[ir]);
		if (data[l + 1] > data[ir])
			dswap(&data[l + 1], &data[ir]);
		if (data[l] > data[l + 1])
			dswap(&data[l], &data[l + 1]);
		int i = l + 1, j = ir;
		double a = data[l + 1];
		for (;;) {
			do
				i++;
			while (data[i] < a);
			do
				j--;
			while (data[j] > a);
			if (j < i)
				break;
			dswap(&data[i], &data[j]);
		}
		data[l + 1] = data[j];
		data[j] = a;
		if (j >= k)
			ir = j - 1;
		if (j <= k)
			l = i;
	}
	return data[k];
}

void orc_domain_nonperiodic(size_t n, const double *x, const double *y,
		const double *z, double center[3], double *size, double origin[3])
{
	const double *c[3] = { x, y, z };
	double *buf = malloc(n * sizeof(double));
	for (int d = 0; d < 3; d++) {   /* find_domain_center, domain.c:772-814 */
		memcpy(buf, c[d], n * sizeof(double));
		center[d] = orc_select((int)(n >> 1), (int)n, buf);
	}
	free(buf);
	double maxd = 0;                /* find_largest_particle_distance */
	for (size_t i = 0; i < n; i++)
		for (int d = 0; d < 3; d++)
			maxd = fmax(maxd, fabs(c[d][i] - center[d]));
	*size = 2.001 * maxd;
	for (int d = 0; d < 3; d++)
		origin[d] = center[d] - 0.5 * *size;
}

static void set_computational_domain(orc_sim *s)
{
	if (s->par.periodic) {            /* domain.c:732-747 */
		s->dom_size = s->par.boxsize;
		for (int d = 0; d < 3; d++) {
			s->dom_center[d] = 0.5 * s->dom_size;
			s->dom_origin[d] = 0;
		}
	} else {
		orc_domain_nonperiodic(s->npart, s->P.pos[0], s->P.pos[1],
				s->P.pos[2], s->dom_center, &s->dom_size, s->dom_origin);
	}
}

/* ------------------------------------------------------------------------ */
/* PH order (peano.c:20-97)                                                  */
/* ------------------------------------------------------------------------ */

static void reorder_all(orc_sim *s, const size_t *idx)
{
	const size_t n = s->npart;
	size_t *scratch = malloc(n * sizeof(size_t));
	orc_particles *P = &s->P;
	for (int d = 0; d < 3; d++) {
		orc_reorder_8(n, P->pos[d], idx, scratch);
		orc_reorder_8(n, P->vel[d], idx, scratch);
		orc_reorder_8(n, P->acc[d], idx, scratch);
		orc_reorder_8(n, P->grav_acc[d], idx, scratch);
	}
	orc_reorder_8(n, P->mass, idx, scratch);
	orc_reorder_8(n, P->last_acc_mag, idx, scratch);
	orc_reorder_8(n, P->cost, idx, scratch);
	orc_reorder_16(n, P->key, idx, scratch);
	orc_reorder_4(n, P->tree_parent, idx, scratch);
	if (P->id)
		orc_reorder_4(n, P->id, idx, scratch);
	if (P->time_bin)
		orc_reorder_4(n, P->time_bin, idx, scratch);
	free(scratch);
}

static void sort_particles_by_peano_key(orc_sim *s)
{
	orc_peano_keys(s->npart, s->P.pos[0], s->P.pos[1], s->P.pos[2],
			s->dom_origin, s->dom_size, s->P.key);
	size_t *idx = malloc(s->npart * sizeof(size_t));
	orc_qsort_index(s->npart, s->P.key, idx, s->par.nthreads);
	reorder_all(s, idx);
	free(idx);
	make_active_list(s);
}

void orc_sim_setup_domain(orc_sim *s)
{
	set_computational_domain(s);      /* Setup_Domain_Decomposition */
	sort_particles_by_peano_key(s);
}

/* ------------------------------------------------------------------------ */
/* top nodes, NTarget = 1 (domain.c:67-145, 220-313, 472-531)                 */
/* ------------------------------------------------------------------------ */

static int cmp_bunch_desc(const void *a, const void *b)
{
	uint64_t x = ((const orc_topnode *)a)->Key;
	uint64_t y = ((const orc_topnode *)b)->Key;
	return (x < y) - (x > y);
}

static void make_top_nodes(orc_sim *s)
{
	/* reset_bunchlist: 8^2 level-2 bunches keyed by their largest short
	 * key; fill_new_bunches: count particles, first particle.  With one
	 * target every bunch lands on task 0 and the imbalance loop stops after
	 * one pass, so the top nodes are the non-empty level-2 cells. */
	orc_topnode b[64];
	memset(b, 0, sizeof(b));
	const int shift = 3 * TOP_LEVEL;
	const uint64_t base = ~(uint64_t)0 >> shift;
	for (int i = 0; i < 64; i++) {
		b[i].Key = ((uint64_t)i << (64 - shift)) | base;
		b[i].Level = TOP_LEVEL;
		b[i].First_Part = INT_MAX;
		b[i].Target = -1;
	}
	int run = 0;
	for (size_t ip = 0; ip < s->npart; ip++) {
		uint64_t sk = (uint64_t)(s->P.key[ip] >> 64);
		while (b[run].Key < sk)
			run++;
		b[run].Npart++;
		if ((int)ip < b[run].First_Part)
			b[run].First_Part = (int)ip;
	}
	int nb = 0;
	for (int i = 0; i < 64; i++)      /* remove_empty_bunches */
		if (b[i].Npart > 0)
			b[nb++] = b[i];
	qsort(b, (size_t)nb, sizeof(b[0]), cmp_bunch_desc);  /* descending */

	const double size = s->dom_size / (double)(1ULL << TOP_LEVEL);
	for (int i = 0; i < nb; i++) {   /* transform_bunches_into_top_nodes */
		int ip = b[i].First_Part;
		for (int d = 0; d < 3; d++) {
			double p = s->P.pos[d][ip] - s->dom_origin[d];
			b[i].Pos[d] = (float)((floor(p / size) + 0.5) * size
					+ s->dom_origin[d]);
		}
	}
	free(s->D);
	s->D = calloc(64, sizeof(*s->D));
	memcpy(s->D, b, (size_t)nb * sizeof(b[0]));
	s->ntop = nb;
}

/* top nodes from a caller's decomposition (transform_bunches_into_top_nodes,
 * domain.c:220-250, with the bunch list given): records in descending key
 * order; Key, Level, First_Part, Npart are read, Pos is recomputed */
void orc_sim_set_top_nodes(orc_sim *s, const orc_topnode *rec, int ntop)
{
	free(s->D);
	s->D = calloc((size_t)ntop, sizeof(*s->D));
	for (int i = 0; i < ntop; i++) {
		orc_topnode b = rec[i];
		const double size = s->dom_size / (double)(1ULL << b.Level);
		const int ip = b.First_Part;
		for (int d = 0; d < 3; d++) {
			double p = s->P.pos[d][ip] - s->dom_origin[d];
			b.Pos[d] = (float)((floor(p / size) + 0.5) * size + s->dom_origin[d]);
		}
		b.Target = -1;
		b.Mass = 0;
		memset(b.CoM, 0, sizeof(b.CoM));
		memset(b.Dp, 0, sizeof(b.Dp));
		s->D[i] = b;
	}
	s->ntop = ntop;
}

void orc_sim_domain_decomposition(orc_sim *s)
{
	set_computational_domain(s);
	sort_particles_by_peano_key(s);
	make_top_nodes(s);
	sort_particles_by_peano_key(s);   /* identity for unique keys */
	orc_reversed_peano_keys(s->npart, s->P.pos[0], s->P.pos[1],
			s->P.pos[2], s->dom_origin, s->dom_size, s->P.key);
}

/* Domain_Decomposition with NTarget = ntarget (domain.c:67-145): the bunch
 * loop (orc_domain.c) on the sorted particles' short keys and P.Cost, its
 * bunches as top nodes; tasks[i] = the target of top node i */
int orc_sim_domain_decomposition_targets(orc_sim *s, int ntarget,
		double part_alloc_factor, double imbal_ceil, int *tasks,
		int *iterations)
{
	set_computational_domain(s);
	sort_particles_by_peano_key(s);
	uint64_t *k = malloc(s->npart * sizeof(*k));
	orc_topnode *rec = malloc(65536 * sizeof(*rec));
	for (size_t i = 0; i < s->npart; i++)
		k[i] = (uint64_t)(s->P.key[i] >> 64);
	int ntop = 0;
	int rc = orc_domain_targets(k, s->P.cost, s->npart, ntarget,
			part_alloc_factor, imbal_ceil, rec, 65536, &ntop, iterations);
	if (rc == 0) {
		for (int i = 0; i < ntop; i++)
			tasks[i] = -(rec[i].Target + 1);
		orc_sim_set_top_nodes(s, rec, ntop);
		sort_particles_by_peano_key(s);   /* identity for unique keys */
		orc_reversed_peano_keys(s->npart, s->P.pos[0], s->P.pos[1],
				s->P.pos[2], s->dom_origin, s->dom_size, s->P.key);
	}
	free(k);
	free(rec);
	return rc;
}

/* ------------------------------------------------------------------------ */
/* tree build (Gravity/tree_build.c:47-572)                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
	orc_sim *s;
	orc_node *t;     /* the zeroed thread buffer */
	int cap;
} builder;

/* create_node_from_particle's node size uses an int shift (tree_build.c:546):
 * on x86-64 the count is taken mod 32 and 1<<31 is INT_MIN. */
static inline double int_shift_size(double size, int lvl)
{
	int32_t p = (int32_t)(1u << (lvl & 31));
	return size / p;
}

static void ensure_cap(builder *b, int need)
{
	if (need < b->cap)
		return;
	int ncap = b->cap * 2 + 64;
	while (ncap <= need)
		ncap *= 2;
	b->t = realloc(b->t, (size_t)ncap * sizeof(orc_node));
	memset(b->t + b->cap, 0, (size_t)(ncap - b->cap) * sizeof(orc_node));
	b->cap = ncap;
}

static inline void add_particle(builder *b, int ip, int node)
{
	orc_node *n = &b->t[node];
	const orc_particles *P = &b->s->P;
	n->CoM[0] += P->pos[0][ip] * P->mass[ip];
	n->CoM[1] += P->pos[1][ip] * P->mass[ip];
	n->CoM[2] += P->pos[2][ip] * P->mass[ip];
	n->Mass += P->mass[ip];
	n->Npart++;
}

static void node_from_particle(builder *b, int ip, int parent, orc_key key,
		int lvl, int node)
{
	ensure_cap(b, node + 1);
	orc_node *t = b->t;
	const orc_particles *P = &b->s->P;
	t[node].DNext = -ip - 1;
	t[node].Bitfield = (uint32_t)lvl | (uint32_t)((key & 7) << 6)
			| (1u << BIT_LOCAL);
	const double size = int_shift_size(b->s->dom_size, lvl);
	for (int d = 0; d < 3; d++) {
		int sgn = -1 + 2 * (P->pos[d][ip] > t[parent].Pos[d]);
		t[node].Pos[d] = t[parent].Pos[d] + sgn * size * 0.5;
	}
	t[node].DUp = node - parent;
	P->tree_parent[ip] = node;
	add_particle(b, ip, node);
}

static void collapse_last_branch(builder *b, int node, int last_parent,
		int ip, int *nn)
{
	orc_node *t = b->t;
	const int vs = b->s->par.vector_size;
	int n;
	if (t[node].Npart <= vs)
		n = node;
	else if (t[last_parent].Npart <= vs)
		n = last_parent;
	else
		return;
	int nzero = *nn - n - 1;
	t[n].DNext = -ip + t[n].Npart - 1;
	*nn -= nzero;
	memset(&t[*nn], 0, (size_t)nzero * sizeof(orc_node));
	int first = -(t[n].DNext + 1);
	for (int j = first; j < first + t[n].Npart; j++)
		b->s->P.tree_parent[j] = n;
}

static int finalise_subtree(builder *b, int top_level, int ti, int nn)
{
	orc_node *t = b->t;
	orc_topnode *T = &b->s->D[ti];
	for (int i = 0; i < nn; i++)
		for (int d = 0; d < 3; d++)
			t[i].CoM[d] /= t[i].Mass;
	T->Mass = (float)t[0].Mass;
	for (int d = 0; d < 3; d++)
		T->CoM[d] = (float)t[0].CoM[d];
	if (t[0].Npart <= b->s->par.vector_size) {
		T->Target = -INT_MAX;
		memset(t, 0, (size_t)nn * sizeof(orc_node));
		return 0;
	}
	ensure_cap(b, nn + 1);
	t = b->t;
	t[nn].Bitfield |= 1u << BIT_TOP;   /* terminator */
	t[nn].Mass = 1;
	nn++;
	t[0].DNext = nn - 1;
	int stack[N_PEANO_TRIPLETS + 1];
	memset(stack, 0, sizeof(stack));
	int lowest = top_level;
	for (int i = 1; i < nn; i++) {
		int lvl = (int)(t[i].Bitfield & 0x3F);
		while (lvl <= lowest) {
			int nd = stack[lowest];
			if (nd > 0)
				t[nd].DNext = i - nd;
			stack[lowest] = 0;
			lowest--;
		}
		if (t[i].DNext == 0) {
			stack[lvl] = i;
			lowest = lvl;
		}
	}
	return nn;
}

static int build_subtree(builder *b, int first, int ti, int top_level)
{
	orc_sim *s = b->s;
	const orc_key *K = s->P.key;
	const int last = first + s->D[ti].Npart - 1;

	/* create_first_node, tree_build.c:379-397 */
	orc_key key0 = K[first] >> (3 * top_level);
	node_from_particle(b, first, 0, key0, top_level, 0);
	for (int d = 0; d < 3; d++)
		b->t[0].Pos[d] = s->D[ti].Pos[d];
	b->t[0].DUp = ti;
	b->t[0].Bitfield |= 1u << BIT_TOP;
	orc_key last_key = key0 >> 3;

	int nn = 1, last_parent = 0;
	for (int ip = first + 1; ip <= last; ip++) {
		orc_key key = K[ip] >> (3 * top_level);
		int node = 0, lvl = top_level, parent = 0;
		int new_branch = 1;
		while (lvl < N_PEANO_TRIPLETS) {
			orc_node *t = b->t;
			if ((int)(key & 7) == (int)((t[node].Bitfield >> 6) & 7)) {
				if (t[node].Npart == 1) {     /* refine */
					t[node].DNext = 0;
					node_from_particle(b, ip - 1, node, last_key, lvl + 1, nn);
					nn++;
					last_key >>= 3;
				}
				add_particle(b, ip, node);
				new_branch = new_branch && (node != last_parent);
				parent = node;
				node++;
				lvl++;
				key >>= 3;
			} else {
				if (t[node].DNext == 0 || node == nn - 1)
					break;
				node += t[node].DNext > 1 ? t[node].DNext : 1;
			}
		}
		if (lvl > N_PEANO_TRIPLETS - 1) {   /* below key resolution */
			s->P.tree_parent[ip] = parent;
			s->ndumped++;
			continue;
		}
		if (new_branch || ip == last)
			collapse_last_branch(b, node, last_parent, ip, &nn);
		if (b->t[node].DNext == 0)
			b->t[node].DNext = nn - node;
		node_from_particle(b, ip, parent, key, lvl, nn);
		nn++;
		last_key = key >> 3;
		last_parent = parent;
	}
	return finalise_subtree(b, top_level, ti, nn);
}

void orc_sim_tree_build(orc_sim *s)
{
	builder b = { s, NULL, 0 };
	int total = 0, cap = 0;
	orc_node *Tree = NULL;
	s->ndumped = 0;
	for (int i = 0; i < s->ntop; i++) {     /* schedule(static,1), 1 thread */
		if (b.t)
			memset(b.t, 0, (size_t)b.cap * sizeof(orc_node));
		int nneed = build_subtree(&b, s->D[i].First_Part, i, s->D[i].Level);
		s->D[i].Target = total;              /* reserve_tree_memory(nNeeded) */
		if (total + nneed > cap) {
			cap = (total + nneed) * 2 + 64;
			Tree = realloc(Tree, (size_t)cap * sizeof(orc_node));
		}
		memcpy(Tree + total, b.t, (size_t)nneed * sizeof(orc_node));
		total += nneed;
		/* set_tree_parent_pointers, tree_build.c:226-242 */
		const int f = s->D[i].First_Part, l = f + s->D[i].Npart;
		for (int ip = f; ip < l; ip++) {
			if (s->D[i].Npart > s->par.vector_size)
				s->P.tree_parent[ip] += s->D[i].Target;
			else
				s->P.tree_parent[ip] = -i - 1;
		}
	}
	free(b.t);
	free(s->Tree);
	s->Tree = Tree;
	s->nnodes = total;
}

/* ------------------------------------------------------------------------ */
/* walks (Gravity/tree_accel.c, Gravity/tree_periodic.c)                      */
/* ------------------------------------------------------------------------ */

typedef struct {
	double pos[3];
	double acc;          /* last |a| */
	double grav[3];
	double cost;
} walker;

void orc_sim_periodic_nearest(const orc_sim *s, double dr[3])
{
	if (!s->par.periodic)
		return;
	const double box = s->par.boxsize, half = s->boxhalf;
	for (int d = 0; d < 3; d++) {
		if (dr[d] > half)
			dr[d] -= box;
		else if (dr[d] < -half)
			dr[d] += box;
	}
}

static inline void interact(const orc_sim *s, walker *w, double mass,
		const double dr[3], double r2)
{
	double fac = s->par.gravity * mass;
	if (r2 < s->Epsilon2) {
		double u2 = r2 / s->Epsilon2;
		fac *= (175 - u2 * (294 - u2 * 135)) / (16 * s->Epsilon3);
	} else {
		double r_inv = 1 / sqrt(r2);
		fac *= r_inv * r_inv * r_inv;
	}
	w->grav[0] += fac * dr[0];
	w->grav[1] += fac * dr[1];
	w->grav[2] += fac * dr[2];
	w->cost += 1;
}

static inline double node_size(const orc_sim *s, int lvl)
{
	return s->dom_size / ((double)(1ULL << lvl));
}

static void pp_range(const orc_sim *s, walker *w, int first, int last)
{
	const orc_particles *P = &s->P;
	if (first < 0 || last > (int)s->npart)
		fatal("particle range outside the particle arrays");
	for (int j = first; j < last; j++) {
		double dr[3] = { P->pos[0][j] - w->pos[0], P->pos[1][j] - w->pos[1],
			P->pos[2][j] - w->pos[2] };
		orc_sim_periodic_nearest(s, dr);
		double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
		if (r2 != 0)
			interact(s, w, P->mass[j], dr, r2);
	}
}

/* interact_with_topnode, tree_accel.c:137-176 */
static int topnode_direct(const orc_sim *s, walker *w, int j, int use_bh)
{
	const orc_topnode *T = &s->D[j];
	const double nSize = s->dom_size / ((double)(1UL << T->Level));
	double dr[3];
	for (int d = 0; d < 3; d++)
		dr[d] = T->Pos[d] - w->pos[d];
	if (fabs(dr[0]) < 0.6 * nSize && fabs(dr[1]) < 0.6 * nSize
			&& fabs(dr[2]) < 0.6 * nSize)
		return 0;
	for (int d = 0; d < 3; d++)
		dr[d] = T->CoM[d] - w->pos[d];
	orc_sim_periodic_nearest(s, dr);
	double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
	double m = T->Mass;
	if (use_bh) {
		if (nSize * nSize > r2 * s->par.theta_bh)
			return 0;
	} else {
		double fac = w->acc / s->par.gravity * s->par.alpha_rel;
		if (m * nSize * nSize > r2 * r2 * fac)
			return 0;
	}
	interact(s, w, m, dr, r2);
	return 1;
}

/* gravity_tree_walk / gravity_tree_walk_BH, tree_accel.c:211-353 */
static void walk_subtree(const orc_sim *s, walker *w, int start, int use_bh)
{
	const orc_node *Tr = s->Tree;
	const double fac = w->acc / s->par.gravity * s->par.alpha_rel;
	int node = start;
	while (Tr[node].DNext != 0 || node == start) {
		if (Tr[node].DNext < 0) {
			int first = -Tr[node].DNext - 1;
			pp_range(s, w, first, first + Tr[node].Npart);
			node++;
			continue;
		}
		double dr[3];
		for (int d = 0; d < 3; d++)
			dr[d] = Tr[node].CoM[d] - w->pos[d];
		orc_sim_periodic_nearest(s, dr);
		double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
		double m = Tr[node].Mass;
		double nSize = node_size(s, (int)(Tr[node].Bitfield & 0x3F));
		if (use_bh) {
			if (nSize * nSize > r2 * s->par.theta_bh) {
				node++;
				continue;
			}
		} else {
			if (m * nSize * nSize > r2 * r2 * fac) {
				node++;
				continue;
			}
			double ds[3];
			for (int d = 0; d < 3; d++)
				ds[d] = Tr[node].Pos[d] - w->pos[d];
			orc_sim_periodic_nearest(s, ds);
			if (fabs(ds[0]) < 0.6 * nSize && fabs(ds[1]) < 0.6 * nSize
					&& fabs(ds[2]) < 0.6 * nSize) {
				node++;
				continue;
			}
		}
		interact(s, w, m, dr, r2);
		node += Tr[node].DNext;
	}
}

/* Ewald_Correction, Gravity/periodic.c:35-113 */
void orc_sim_ewald_correction(const orc_sim *s, const double dr[3],
		double f[3])
{
	const size_t n3 = (size_t)NE1 * NE1 * NE1;
	int sign[3] = { -1, -1, -1 };
	double u[3];
	int c[3];
	for (int d = 0; d < 3; d++) {
		double x = dr[d];
		if (x < 0) {
			x = -x;
			sign[d] = 1;
		}
		u[d] = x * s->box2ewald;
		c[d] = (int)u[d];
		if (c[d] >= N_EWALD)
			c[d] = N_EWALD - 1;
		u[d] -= c[d];
	}
	const double U = u[0], V = u[1], W = u[2];
	const double wt[8] = { (1 - U) * (1 - V) * (1 - W), (1 - U) * (1 - V) * W,
		(1 - U) * V * (1 - W), (1 - U) * V * W, U * (1 - V) * (1 - W),
		U * (1 - V) * W, U * V * (1 - W), U * V * W };
	const int i = c[0], j = c[1], k = c[2];
#define IDX(a, b, cc) (((size_t)(a) * NE1 + (b)) * NE1 + (cc))
	for (int d = 0; d < 3; d++) {
		const double *F = s->ewald + (size_t)d * n3;
		f[d] = wt[0] * F[IDX(i, j, k)] + wt[1] * F[IDX(i, j, k + 1)]
			+ wt[2] * F[IDX(i, j + 1, k)] + wt[3] * F[IDX(i, j + 1, k + 1)]
			+ wt[4] * F[IDX(i + 1, j, k)] + wt[5] * F[IDX(i + 1, j, k + 1)]
			+ wt[6] * F[IDX(i + 1, j + 1, k)]
			+ wt[7] * F[IDX(i + 1, j + 1, k + 1)];
	}
#undef IDX
	for (int d = 0; d < 3; d++)
		f[d] *= sign[d];
}

static inline void ewald_cube(const orc_sim *s, walker *w, const double dr[3],
		double mass)
{
	double f[3];
	orc_sim_ewald_correction(s, dr, f);
	w->grav[0] += s->par.gravity * mass * f[0];
	w->grav[1] += s->par.gravity * mass * f[1];
	w->grav[2] += s->par.gravity * mass * f[2];
	w->cost++;
}

static void ewald_pp_range(const orc_sim *s, walker *w, int first, int last)
{
	const orc_particles *P = &s->P;
	if (first < 0 || last > (int)s->npart)
		fatal("particle range outside the particle arrays");
	for (int j = first; j < last; j++) {
		double dr[3] = { P->pos[0][j] - w->pos[0], P->pos[1][j] - w->pos[1],
			P->pos[2][j] - w->pos[2] };
		orc_sim_periodic_nearest(s, dr);
		if (dr[0] != 0 && dr[1] != 0 && dr[2] != 0)
			ewald_cube(s, w, dr, P->mass[j]);
	}
}

/* interact_with_topnode (periodic), tree_periodic.c:112-177 */
static int topnode_ewald(const orc_sim *s, walker *w, int j, int use_bh)
{
	const orc_topnode *T = &s->D[j];
	const double nsz = s->dom_size / (1UL << T->Level);
	const double box = s->par.boxsize;
	int want_open = 0;
	double dr[3];
	for (int d = 0; d < 3; d++)
		dr[d] = T->CoM[d] - w->pos[d];
	orc_sim_periodic_nearest(s, dr);
	double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
	if (use_bh) {
		if (nsz * nsz > r2 * s->par.theta_bh)
			want_open = 1;
	} else {
		double m = T->Mass;
		double fac = w->acc / s->par.gravity * s->par.alpha_rel;
		if (m * nsz * nsz > r2 * r2 * fac)
			want_open = 1;
	}
	if (fabs(w->pos[0] - T->Pos[0]) < 0.6 * nsz
			&& fabs(w->pos[1] - T->Pos[1]) < 0.6 * nsz
			&& fabs(w->pos[2] - T->Pos[2]) < 0.6 * nsz)
		want_open = 1;
	if (want_open) {
		for (int d = 0; d < 3; d++)
			if (fabs(dr[d]) > 0.5 * (box - nsz))
				return 0;
		if (nsz > 0.2 * box)
			return 0;
	}
	ewald_cube(s, w, dr, T->Mass);
	return 1;
}

/* gravity_tree_walk_ewald(_BH), tree_periodic.c:216-422 */
static void walk_subtree_ewald(const orc_sim *s, walker *w, int start,
		int use_bh)
{
	const orc_node *Tr = s->Tree;
	const double box = s->par.boxsize;
	const double fac = w->acc / s->par.gravity * s->par.alpha_rel;
	const double bh2 = s->par.theta_bh * s->par.theta_bh;
	const int end = start + Tr[start].DNext;
	int node = start;
	while (node != end) {
		if (Tr[node].DNext < 0) {
			int first = -Tr[node].DNext - 1;
			ewald_pp_range(s, w, first, first + Tr[node].Npart);
			node++;
			continue;
		}
		double dr[3], ds[3];
		for (int d = 0; d < 3; d++)
			dr[d] = Tr[node].CoM[d] - w->pos[d];
		orc_sim_periodic_nearest(s, dr);
		double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
		double m = Tr[node].Mass;
		double nsz = node_size(s, (int)(Tr[node].Bitfield & 0x3F));
		for (int d = 0; d < 3; d++)
			ds[d] = Tr[node].Pos[d] - w->pos[d];
		int want_open = 0;
		if (use_bh ? (nsz * nsz > r2 * bh2) : (m * nsz * nsz > r2 * r2 * fac))
			want_open = 1;
		else if (fabs(ds[0]) < 0.6 * nsz && fabs(ds[1]) < 0.6 * nsz
				&& fabs(ds[2]) < 0.6 * nsz)
			want_open = 1;
		orc_sim_periodic_nearest(s, ds);
		if (want_open) {
			double dl = 0.5 * (box - nsz);
			if (fabs(ds[0]) > dl || fabs(ds[1]) > dl || fabs(ds[2]) > dl
					|| nsz > 0.2 * box) {
				node++;
				continue;
			}
		}
		ewald_cube(s, w, dr, m);
		node += Tr[node].DNext;
	}
}

/* Gravity_Tree_Acceleration + Gravity_Tree_Periodic, tree_accel.c:32-129,
 * tree_periodic.c:23-110 */
void orc_sim_tree_acceleration(orc_sim *s, int use_bh)
{
	orc_particles *P = &s->P;
	const int vs = s->par.vector_size;

	#pragma omp parallel for schedule(dynamic, 16) num_threads(s->par.nthreads)
	for (int i = 0; i < s->nactive; i++) {
		const int ip = s->active[i];
		walker w;
		memset(&w, 0, sizeof(w));
		for (int d = 0; d < 3; d++)
			w.pos[d] = P->pos[d][ip];
		w.acc = P->last_acc_mag[ip];
		for (int j = 0; j < s->ntop; j++) {
			if (topnode_direct(s, &w, j, use_bh))
				continue;
			if (s->D[j].Npart <= vs) {
				pp_range(s, &w, s->D[j].Target, s->D[j].Target + s->D[j].Npart);
				continue;
			}
			walk_subtree(s, &w, s->D[j].Target, use_bh);
		}
		for (int d = 0; d < 3; d++) {   /* add_recv_to, tree_accel.c:110-129 */
			P->acc[d][ip] += w.grav[d];
			P->grav_acc[d][ip] = w.grav[d];
		}
		P->cost[ip] = w.cost;
	}

	if (!s->par.periodic)
		return;
	if (!s->ewald)
		fatal("periodic walk without an Ewald table");

	#pragma omp parallel for schedule(dynamic, 16) num_threads(s->par.nthreads)
	for (int i = 0; i < s->nactive; i++) {
		const int ip = s->active[i];
		walker w;
		memset(&w, 0, sizeof(w));
		for (int d = 0; d < 3; d++)
			w.pos[d] = P->pos[d][ip];
		w.acc = P->last_acc_mag[ip];
		for (int j = 0; j < s->ntop; j++) {
			if (topnode_ewald(s, &w, j, use_bh))
				continue;
			if (s->D[j].Npart <= vs) {
				ewald_pp_range(s, &w, s->D[j].Target,
						s->D[j].Target + s->D[j].Npart);
				continue;
			}
			walk_subtree_ewald(s, &w, s->D[j].Target, use_bh);
		}
		for (int d = 0; d < 3; d++) {   /* add_recv_to, tree_periodic.c:91 */
			P->acc[d][ip] += w.grav[d];
			P->grav_acc[d][ip] += w.grav[d];
		}
		P->cost[ip] += w.cost;
	}
}

/* safe_last_accel, accel.c:28-43 */
static void safe_last_accel(orc_sim *s)
{
	orc_particles *P = &s->P;
	for (int i = 0; i < s->nactive; i++) {
		int ip = s->active[i];
		P->last_acc_mag[ip] = sqrt(P->acc[0][ip] * P->acc[0][ip]
				+ P->acc[1][ip] * P->acc[1][ip] + P->acc[2][ip] * P->acc[2][ip]);
		P->acc[0][ip] = P->acc[1][ip] = P->acc[2][ip] = 0;
	}
}

/* Compute_Acceleration + accel_gravity, accel.c:15-71 (tree always rebuilt:
 * the caller ran a domain decomposition, which sets Sig.Tree_Update) */
void orc_sim_compute_acceleration(orc_sim *s, int prepare_step)
{
	safe_last_accel(s);
	orc_sim_tree_build(s);
	if (prepare_step) {
		orc_sim_tree_acceleration(s, 1);
		safe_last_accel(s);
	}
	orc_sim_tree_acceleration(s, 0);
}

/* ------------------------------------------------------------------------ */
/* accessors                                                                 */
/* ------------------------------------------------------------------------ */

void orc_sim_domain(const orc_sim *s, double *size, double origin[3],
		double center[3])
{
	*size = s->dom_size;
	for (int d = 0; d < 3; d++) {
		origin[d] = s->dom_origin[d];
		center[d] = s->dom_center[d];
	}
}

int orc_sim_ntop(const orc_sim *s) { return s->ntop; }
const orc_topnode *orc_sim_top(const orc_sim *s) { return s->D; }
int orc_sim_nnodes(const orc_sim *s) { return s->nnodes; }
const orc_node *orc_sim_tree(const orc_sim *s) { return s->Tree; }
int orc_sim_ndumped(const orc_sim *s) { return s->ndumped; }
double orc_sim_epsilon(const orc_sim *s) { return s->Epsilon; }

/* ------------------------------------------------------------------------ */
/* FP64 direct summation (error oracle; same pairwise kernel as the walk)    */
/* ------------------------------------------------------------------------ */

void orc_direct_sum(const orc_sim *s, int nt, const int *idx, double *acc)
{
	const orc_particles *P = &s->P;
	const int n = (int)s->npart;
	#pragma omp parallel for schedule(dynamic, 1) num_threads(s->par.nthreads)
	for (int t = 0; t < nt; t++) {
		const int i = idx[t];
		walker w;
		memset(&w, 0, sizeof(w));
		for (int d = 0; d < 3; d++)
			w.pos[d] = P->pos[d][i];
		for (int j = 0; j < n; j++) {
			if (j == i)
				continue;
			double dr[3] = { P->pos[0][j] - w.pos[0], P->pos[1][j] - w.pos[1],
				P->pos[2][j] - w.pos[2] };
			orc_sim_periodic_nearest(s, dr);
			double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
			if (r2 == 0)
				continue;
			interact(s, &w, P->mass[j], dr, r2);
			if (s->par.periodic)
				ewald_cube(s, &w, dr, P->mass[j]);
		}
		for (int d = 0; d < 3; d++)
			acc[3 * t + d] = w.grav[d];
	}
}
