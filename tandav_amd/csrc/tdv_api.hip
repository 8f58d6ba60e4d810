/*
 * C ABI of the engine (include/tandav_gpu.h): context, device buffers,
 * host <-> device copies and the step drivers that chain the kernels in the
 * reference's order (domain.c:67-145, accel.c:15-71).
 */
#include "tdv_internal.h"
#include "tdv_device.h"

#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#define TDV_EBIT_MASK (TDV_EBIT_DUP | TDV_EBIT_DEPTH | TDV_EBIT_RANGE | TDV_EBIT_TIES)

/* Tree_Node layout, Gravity/tree.h:23-32 (Float = double) */
struct NodeRecord {
	int DNext;
	uint32_t Bitfield;
	int DUp, Npart;
	double Pos[3], Mass, CoM[3], Dp[3];
};
static_assert(sizeof(NodeRecord) == 96, "Tree_Node is 96 bytes");

/* Top_Tree_Node layout, domain.h:40-52 (records in and out) */


int tdv_fail(tdv_context *c, int code, const std::string &msg)
{
	if (c)
		c->err = msg;
	return code;
}

int tdv_check(tdv_context *c, hipError_t e, const char *what)
{
	if (e == hipSuccess)
		return 0;
	(void)hipGetLastError();
	return tdv_fail(c, e == hipErrorOutOfMemory ? TDV_ERR_NOMEM : TDV_ERR_HIP,
			std::string(what) + ": " + hipGetErrorString(e));
}

int tdv_check_err_flags(tdv_context *c)
{
	unsigned e = 0;
	TDV_HIP(c, hipMemcpyAsync(&e, c->d_err, sizeof(e), hipMemcpyDeviceToHost,
			c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	if (!e)
		return 0;
	TDV_HIP(c, hipMemsetAsync(c->d_err, 0, sizeof(unsigned), c->stream));
	if (e & TDV_EBIT_DUP)
		return tdv_fail(c, TDV_ERR_DUP_KEYS,
				"two particles share a 128-bit Peano-Hilbert key");
	if (e & TDV_EBIT_DEPTH)
		return tdv_fail(c, TDV_ERR_KEY_DEPTH,
				"particles closer than the Peano-Hilbert key resolution");
	if (e & TDV_EBIT_TIES)
		return tdv_fail(c, TDV_ERR_TIES, "run of equal 64-bit key prefixes");
	return tdv_fail(c, TDV_ERR_RANGE,
			"a small top node's particle range leaves the particle arrays");
}

template <typename T>
static int dalloc(tdv_context *c, T **p, size_t count)
{
	*p = nullptr;
	if (count == 0)
		count = 1;
	return tdv_check(c, hipMalloc((void **)p, count * sizeof(T)), "hipMalloc");
}

template <typename T>
static void dfree(T *&p)
{
	if (p)
		(void)hipFree((void *)p);
	p = nullptr;
}

/* a captured walk holds the buffers' addresses: drop it whenever a buffer
 * may move (sort, rebuild, new particles, reallocation) */
void tdv_graph_drop(tdv_context *c)
{
	if (c->graph_exec)
		(void)hipGraphExecDestroy(c->graph_exec);
	if (c->graph)
		(void)hipGraphDestroy(c->graph);
	c->graph_exec = nullptr;
	c->graph = nullptr;
	if (c->step_exec)
		(void)hipGraphExecDestroy(c->step_exec);
	if (c->step_graph)
		(void)hipGraphDestroy(c->step_graph);
	c->step_exec = nullptr;
	c->step_graph = nullptr;
}

/* stage events, left out of a captured step */
static void rec(tdv_context *c, int e)
{
	if (!c->graph_mode)
		(void)hipEventRecord(c->ev[e], c->stream);
}

/* the first and second derivatives of the Ewald correction at the table
 * nodes, computed once per context */
int tdv_ewald_derivs_ready(tdv_context *c)
{
	if (c->ewald_s && c->ewald_j)
		return 0;
	if (!c->ewald_s)
		TDV_TRY(dalloc(c, &c->ewald_s, 10 * (size_t)TDV_NE3));
	if (!c->ewald_j)
		TDV_TRY(dalloc(c, &c->ewald_j, 6 * (size_t)TDV_NE3));
	return tdv_ewald_derivs_device(c, c->ewald_s, c->ewald_j);
}

static int set_nactive(tdv_context *c, int nact)
{
	c->nactive = nact;
	return tdv_check(c, hipMemcpyAsync(c->d_nactive, &c->nactive, sizeof(int),
			hipMemcpyHostToDevice, c->stream), "nactive");
}

static void free_particles(tdv_context *c)
{
	tdv_graph_drop(c);
	for (int d = 0; d < 3; d++) {
		dfree(c->pos[d]);
		dfree(c->vel[d]);
		dfree(c->acc[d]);
		dfree(c->gacc[d]);
	}
	dfree(c->mass);
	dfree(c->last_acc);
	dfree(c->cost);
	dfree(c->tmp_f64);
	dfree(c->key);
	dfree(c->key_tmp);
	dfree(c->tree_parent);
	dfree(c->time_bin);
	dfree(c->id);
	dfree(c->it_drift);
	dfree(c->it_kick);
	dfree(c->kick_dt);
	dfree(c->pm);
	dfree(c->active);
	for (int b = 0; b < 2; b++) {
		dfree(c->sk[b]);
		dfree(c->sv[b]);
	}
	dfree(c->hist);
	dfree(c->scan_part);
	dfree(c->perm);
	dfree(c->minpack);
	dfree(c->cnt);
	dfree(c->off);
	dfree(c->lohi);
	dfree(c->top.ptop);
	dfree(c->p32);                   /* FP32 walk records, sized by cap */
	dfree(c->p32o);
	dfree(c->gkey);
	dfree(c->gorder);
	c->cap = 0;
}

static int alloc_particles(tdv_context *c, size_t n)
{
	free_particles(c);
	for (int d = 0; d < 3; d++) {
		TDV_TRY(dalloc(c, &c->pos[d], n));
		TDV_TRY(dalloc(c, &c->vel[d], n));
		TDV_TRY(dalloc(c, &c->acc[d], n));
		TDV_TRY(dalloc(c, &c->gacc[d], n));
	}
	TDV_TRY(dalloc(c, &c->mass, n));
	TDV_TRY(dalloc(c, &c->last_acc, n));
	TDV_TRY(dalloc(c, &c->cost, n));
	TDV_TRY(dalloc(c, &c->tmp_f64, n));
	TDV_TRY(dalloc(c, &c->key, 2 * n));
	TDV_TRY(dalloc(c, &c->key_tmp, 2 * n));
	TDV_TRY(dalloc(c, &c->tree_parent, n));
	TDV_TRY(dalloc(c, &c->time_bin, n));
	TDV_TRY(dalloc(c, &c->id, n));
	TDV_TRY(dalloc(c, &c->it_drift, n));
	TDV_TRY(dalloc(c, &c->it_kick, n));
	TDV_TRY(dalloc(c, &c->kick_dt, n));
	TDV_TRY(dalloc(c, &c->pm, n));
	TDV_TRY(dalloc(c, &c->active, n));
	for (int b = 0; b < 2; b++) {
		TDV_TRY(dalloc(c, &c->sk[b], n));
		TDV_TRY(dalloc(c, &c->sv[b], n));
	}
	const size_t nb = (n + TDV_SORT_TILE - 1) / TDV_SORT_TILE;
	c->hist_cap = nb * 256 + 1;
	TDV_TRY(dalloc(c, &c->hist, c->hist_cap));
	const size_t sp = (c->hist_cap > n + 1 ? c->hist_cap : n + 1) / 1024 + 2;
	TDV_TRY(dalloc(c, &c->scan_part, sp));
	TDV_TRY(dalloc(c, &c->perm, n));
	TDV_TRY(dalloc(c, &c->minpack, n));
	TDV_TRY(dalloc(c, &c->cnt, n));
	TDV_TRY(dalloc(c, &c->off, n + 1));
	TDV_TRY(dalloc(c, &c->lohi, n));
	TDV_TRY(dalloc(c, &c->top.ptop, n));
	c->cap = n;
	return 0;
}

int tdv_ensure_nodes(tdv_context *c, size_t nodes)
{
	if (nodes <= c->node_cap && c->node_a)
		return 0;
	dfree(c->node_a);
	dfree(c->node_b);
	dfree(c->node_c);
	dfree(c->node_up);
	dfree(c->node_dp);
	/* Max_Nodes grows by TREE_ENLARGEMENT_FACTOR (tree_build.c:6,162) */
	size_t cap = (size_t)(1.2 * (double)nodes) + 1024;
	TDV_TRY(dalloc(c, &c->node_a, cap));
	TDV_TRY(dalloc(c, &c->node_b, cap));
	TDV_TRY(dalloc(c, &c->node_c, cap));
	TDV_TRY(dalloc(c, &c->node_up, cap));
	TDV_TRY(dalloc(c, &c->node_dp, cap));
	c->node_cap = cap;
	return 0;
}

static void free_top(tdv_context *c)
{
	TopNodes &t = c->top;
	dfree(t.first);
	dfree(t.npart);
	dfree(t.level);
	dfree(t.target);
	dfree(t.term);
	dfree(t.key);
	dfree(t.pos);
	dfree(t.mass);
	dfree(t.com);
	dfree(t.dp);
	dfree(t.kicked);
	dfree(t.ntop);
	dfree(t.nnodes);
	c->top_cap = 0;
}

static int alloc_top(tdv_context *c, int cap)
{
	int *ptop = c->top.ptop;        /* per particle, owned by the columns */
	free_top(c);
	c->top.ptop = ptop;
	TopNodes &t = c->top;
	TDV_TRY(dalloc(c, &t.first, cap));
	TDV_TRY(dalloc(c, &t.npart, cap));
	TDV_TRY(dalloc(c, &t.level, cap));
	TDV_TRY(dalloc(c, &t.target, cap));
	TDV_TRY(dalloc(c, &t.term, cap));
	TDV_TRY(dalloc(c, &t.key, cap));
	TDV_TRY(dalloc(c, &t.pos, 3 * (size_t)cap));
	TDV_TRY(dalloc(c, &t.mass, cap));
	TDV_TRY(dalloc(c, &t.com, 3 * (size_t)cap));
	TDV_TRY(dalloc(c, &t.dp, 3 * (size_t)cap));
	TDV_TRY(dalloc(c, &t.kicked, cap));
	TDV_TRY(dalloc(c, &t.ntop, 1));
	TDV_TRY(dalloc(c, &t.nnodes, 1));
	TDV_HIP(c, hipMemset(t.ntop, 0, sizeof(int)));
	c->top_cap = cap;
	return 0;
}

static int compute_ewald(tdv_context *c)
{
	double *raw = nullptr;
	TDV_TRY(dalloc(c, &raw, 4 * (size_t)TDV_NE3));
	int r = tdv_ewald_table_device(c, raw);
	if (!r) {
		std::vector<double> h(4 * (size_t)TDV_NE3);
		r = tdv_check(c, hipMemcpyAsync(h.data(), raw, h.size() * 8,
				hipMemcpyDeviceToHost, c->stream), "ewald copy");
		if (!r)
			r = tdv_check(c, hipStreamSynchronize(c->stream), "ewald sync");
		if (!r)
			r = tdv_set_ewald_table(c, h.data());
	}
	dfree(raw);
	return r;
}

extern "C" {

int tdv_create(const tdv_params *p, tdv_context **out)
{
	if (!p || !out)
		return TDV_ERR_ARG;
	*out = nullptr;
	if (p->vector_size != 64)
		return TDV_ERR_ARG;
	if (p->periodic && !(p->boxsize > 0))
		return TDV_ERR_ARG;
	if (p->precision != TDV_WALK_FP64 && p->precision != TDV_WALK_FP32)
		return TDV_ERR_ARG;
	if (p->precision == TDV_WALK_FP32 && p->periodic
			&& !(p->ewald_size > 0 && p->ewald_accuracy > 0))
		return TDV_ERR_ARG;
	if (p->fp32_order < 0 || p->fp32_order > 2)
		return TDV_ERR_ARG;
	if (p->fp32_criterion != 0 && p->fp32_criterion != 2
			&& p->fp32_criterion != 3)
		return TDV_ERR_ARG;
	if (p->fp32_carry < 0 || p->fp32_carry > 4096)
		return TDV_ERR_ARG;
	if (p->fp32_waves != 0 && p->fp32_waves != 6 && p->fp32_waves != 8)
		return TDV_ERR_ARG;
	if (p->fp32_queue != 0 && p->fp32_queue != 1)
		return TDV_ERR_ARG;
	if (p->fp32_split < 0 || p->fp32_split > 256)
		return TDV_ERR_ARG;
	if (p->fp32_bundle_min < -1 || p->fp32_bundle_min > 64)
		return TDV_ERR_ARG;
	tdv_context *c = new (std::nothrow) tdv_context();
	if (!c)
		return TDV_ERR_NOMEM;
	c->par = *p;
	/* Setup_Gravity_Tree, tree_build.c:134-139: K1 length, negative sign */
	c->Epsilon = -41.0 / 32.0 * p->softening;
	c->Epsilon2 = c->Epsilon * c->Epsilon;
	c->Epsilon3 = c->Epsilon * c->Epsilon * c->Epsilon;
	int r = tdv_check(c, hipSetDevice(p->device), "hipSetDevice");
	if (!r)
		r = tdv_check(c, hipStreamCreateWithFlags(&c->stream,
				hipStreamNonBlocking), "hipStreamCreate");
	for (int e = 0; !r && e < EV_N; e++)
		r = tdv_check(c, hipEventCreate(&c->ev[e]), "hipEventCreate");
	if (!r)
		r = dalloc(c, &c->d_err, 1);
	if (!r)
		r = dalloc(c, &c->d_count, 4);
	if (!r)
		r = dalloc(c, &c->d_nactive, 1);
	if (!r)
		r = dalloc(c, &c->d_queue, 1);
	if (!r)
		r = dalloc(c, &c->long_count, 2);
	if (!r)
		r = dalloc(c, &c->q_arrived, 192);
	if (!r)
		r = tdv_check(c, hipMemset(c->d_err, 0, sizeof(unsigned)), "memset");
	if (!r)
		r = alloc_top(c, 64);
	if (!r && p->periodic)
		r = compute_ewald(c);       /* Gravity_Periodic_Init */
	if (r) {
		fprintf(stderr, "tdv_create: %s\n", c->err.c_str());
		tdv_destroy(c);
		return r;
	}
	*out = c;
	return 0;
}

void tdv_destroy(tdv_context *c)
{
	if (!c)
		return;
	if (c->stream)
		(void)hipStreamSynchronize(c->stream);
	free_particles(c);
	free_top(c);
	dfree(c->node_a);
	dfree(c->node_b);
	dfree(c->node_c);
	dfree(c->node_up);
	dfree(c->ewald);
	dfree(c->tl_spline);
	dfree(c->ewald16);
	dfree(c->ewald_s);
	dfree(c->ewald_j);
	dfree(c->p32);
	dfree(c->p32o);
	dfree(c->a32);
	dfree(c->rec32);
	dfree(c->ctr32);
	dfree(c->q32);
	dfree(c->q_s6);
	dfree(c->q_arrived);
	dfree(c->lvl_nodes);
	dfree(c->tq32);
	dfree(c->top32);
	tdv_graph_drop(c);
	dfree(c->d_err);
	dfree(c->d_count);
	dfree(c->d_nactive);
	dfree(c->d_queue);
	dfree(c->partial);
	dfree(c->wmask);
	dfree(c->gkey);
	dfree(c->gorder);
	dfree(c->long_count);
	dfree(c->long_list);
	for (int e = 0; e < EV_N; e++)
		if (c->ev[e])
			(void)hipEventDestroy(c->ev[e]);
	if (c->stream)
		(void)hipStreamDestroy(c->stream);
	delete c;
}

const char *tdv_last_error(const tdv_context *c)
{
	return c ? c->err.c_str() : "no context";
}

size_t tdv_num_particles(const tdv_context *c)
{
	return c ? c->n : 0;
}

int tdv_set_particles(tdv_context *c, size_t n, const double *pos[3],
		const double *vel[3], const double *mass, const double *acc[3],
		const double *last_acc_mag, const int *time_bin, const uint32_t *id)
{
	if (!c || !pos || !mass || n == 0 || n > 0x7FFFFFFFu)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_particles: bad arguments");
	if (n > c->cap)
		TDV_TRY(alloc_particles(c, n));
	c->n = n;
	const size_t b = n * sizeof(double);
	for (int d = 0; d < 3; d++) {
		TDV_HIP(c, hipMemcpyAsync(c->pos[d], pos[d], b, hipMemcpyHostToDevice,
				c->stream));
		if (vel && vel[d])
			TDV_HIP(c, hipMemcpyAsync(c->vel[d], vel[d], b,
					hipMemcpyHostToDevice, c->stream));
		else
			TDV_HIP(c, hipMemsetAsync(c->vel[d], 0, b, c->stream));
		if (acc && acc[d])
			TDV_HIP(c, hipMemcpyAsync(c->acc[d], acc[d], b,
					hipMemcpyHostToDevice, c->stream));
		else
			TDV_HIP(c, hipMemsetAsync(c->acc[d], 0, b, c->stream));
		TDV_HIP(c, hipMemsetAsync(c->gacc[d], 0, b, c->stream));
	}
	TDV_HIP(c, hipMemcpyAsync(c->mass, mass, b, hipMemcpyHostToDevice,
			c->stream));
	if (last_acc_mag)
		TDV_HIP(c, hipMemcpyAsync(c->last_acc, last_acc_mag, b,
				hipMemcpyHostToDevice, c->stream));
	else
		TDV_HIP(c, hipMemsetAsync(c->last_acc, 0, b, c->stream));
	TDV_HIP(c, hipMemsetAsync(c->cost, 0, b, c->stream));
	TDV_HIP(c, hipMemsetAsync(c->tree_parent, 0, n * sizeof(int), c->stream));
	/* It_Drift_Pos = It_Kick_Pos = Int_Time.Beg (timestep.c:137-153) */
	TDV_HIP(c, hipMemsetAsync(c->it_drift, 0, n * sizeof(uint32_t),
			c->stream));
	TDV_HIP(c, hipMemsetAsync(c->it_kick, 0, n * sizeof(uint32_t),
			c->stream));
	if (time_bin)
		TDV_HIP(c, hipMemcpyAsync(c->time_bin, time_bin, n * sizeof(int),
				hipMemcpyHostToDevice, c->stream));
	else
		TDV_HIP(c, hipMemsetAsync(c->time_bin, 0, n * sizeof(int), c->stream));
	if (id) {
		TDV_HIP(c, hipMemcpyAsync(c->id, id, n * sizeof(uint32_t),
				hipMemcpyHostToDevice, c->stream));
	} else {
		std::vector<uint32_t> ids(n);
		for (size_t i = 0; i < n; i++)
			ids[i] = (uint32_t)i;
		TDV_HIP(c, hipMemcpyAsync(c->id, ids.data(), n * sizeof(uint32_t),
				hipMemcpyHostToDevice, c->stream));
		TDV_HIP(c, hipStreamSynchronize(c->stream));
	}
	/* every particle active (Time_Bin <= Max_Active_Bin on a sync point) */
	std::vector<int> act(n);
	for (size_t i = 0; i < n; i++)
		act[i] = (int)i;
	TDV_HIP(c, hipMemcpyAsync(c->active, act.data(), n * sizeof(int),
			hipMemcpyHostToDevice, c->stream));
	TDV_TRY(set_nactive(c, (int)n));
	tdv_graph_drop(c);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	c->have_domain = 0;
	c->have_tree = 0;
	tdv_launch_pack_pm(c);
	return tdv_check(c, hipStreamSynchronize(c->stream), "set_particles");
}

int tdv_get_particles(tdv_context *c, double *pos[3], double *vel[3],
		double *mass, double *acc[3], double *grav_acc[3], double *cost,
		double *last_acc_mag, int *tree_parent, uint64_t *key, uint32_t *id,
		int *time_bin)
{
	if (!c)
		return TDV_ERR_ARG;
	const size_t n = c->n, b = n * sizeof(double);
	auto cp = [&](void *h, const void *d, size_t bytes) {
		return tdv_check(c, hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost,
				c->stream), "get_particles");
	};
	for (int d = 0; d < 3; d++) {
		if (pos && pos[d])
			TDV_TRY(cp(pos[d], c->pos[d], b));
		if (vel && vel[d])
			TDV_TRY(cp(vel[d], c->vel[d], b));
		if (acc && acc[d])
			TDV_TRY(cp(acc[d], c->acc[d], b));
		if (grav_acc && grav_acc[d])
			TDV_TRY(cp(grav_acc[d], c->gacc[d], b));
	}
	if (mass)
		TDV_TRY(cp(mass, c->mass, b));
	if (cost)
		TDV_TRY(cp(cost, c->cost, b));
	if (last_acc_mag)
		TDV_TRY(cp(last_acc_mag, c->last_acc, b));
	if (tree_parent)
		TDV_TRY(cp(tree_parent, c->tree_parent, n * sizeof(int)));
	if (id)
		TDV_TRY(cp(id, c->id, n * sizeof(uint32_t)));
	if (time_bin)
		TDV_TRY(cp(time_bin, c->time_bin, n * sizeof(int)));
	if (key) {
		if (c->have_domain) {       /* P.Key holds reversed keys */
			tdv_launch_reverse_keys(c, n, c->key, c->key_tmp);
			TDV_TRY(cp(key, c->key_tmp, 2 * n * sizeof(uint64_t)));
		} else {
			memset(key, 0, 2 * n * sizeof(uint64_t));
		}
	}
	return tdv_check(c, hipStreamSynchronize(c->stream), "get_particles");
}

/* set_computational_domain (domain.c:732-847) */
static int computational_domain(tdv_context *c)
{
	if (!c->n)
		return tdv_fail(c, TDV_ERR_ARG, "no particles");
	if (c->par.periodic) {                 /* domain.c:732-747 */
		c->dom_size = c->par.boxsize;
		for (int d = 0; d < 3; d++) {
			c->dom_center[d] = 0.5 * c->dom_size;
			c->dom_origin[d] = 0;
		}
		return 0;
	}
	return tdv_domain_nonperiodic(c);
}

__global__ void k_iota(int n, int *__restrict__ a)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		a[i] = i;
}

/* Sort_Particles_By_Peano_Key (peano.c:20-48) with the current domain:
 * keys, sort, reorder of every column.  The active list's indices belong to
 * the old order, so it is reset to every particle (a sync point, as at every
 * domain decomposition of the reference); tdv_set_active narrows it. */
__global__ void k_set_int(int *p, int v)
{
	if (threadIdx.x == 0)
		*p = v;
}

__global__ void k_clear_err_bits(unsigned *err, unsigned bits)
{
	if (threadIdx.x == 0)
		*err &= ~bits;
}

static int keys_sort_reorder(tdv_context *c)
{
	const size_t n = c->n;
	if (!n)
		return tdv_fail(c, TDV_ERR_ARG, "no particles");
	rec(c, EV_KEYS0);
	tdv_launch_keys(c, n, c->pos[0], c->pos[1], c->pos[2], c->dom_origin,
			c->dom_size, c->key_tmp, 0);
	rec(c, EV_KEYS1);
	TDV_TRY(tdv_sort_device(c, n, c->key_tmp, c->perm));
	/* equal keys (particles within Domain.Size / 2^42 on every axis) stay in
	 * their input order: the reference's quicksort leaves them in an order
	 * of its own (sort.c:253-341), the build dumps them all the same */
	hipLaunchKernelGGL(k_clear_err_bits, dim3(1), dim3(64), 0, c->stream,
			c->d_err, TDV_EBIT_DUP);
	rec(c, EV_SORT1);
	/* reorder_collisionless_particles (peano.c:50-82) as gathers */
	tdv_launch_gather_key(c, n, c->perm, c->key_tmp, c->key);
	double **cols[] = { &c->pos[0], &c->pos[1], &c->pos[2], &c->vel[0],
		&c->vel[1], &c->vel[2], &c->acc[0], &c->acc[1], &c->acc[2],
		&c->gacc[0], &c->gacc[1], &c->gacc[2], &c->mass, &c->last_acc,
		&c->cost };
	for (double **col : cols) {
		tdv_launch_gather_f64(c, n, c->perm, *col, c->tmp_f64);
		if (c->graph_mode) {
			/* a captured step keeps every column where it is, so each
			 * replay finds its buffers: copy back instead of swapping */
			TDV_HIP(c, hipMemcpyAsync(*col, c->tmp_f64, n * sizeof(double),
					hipMemcpyDeviceToDevice, c->stream));
			continue;
		}
		double *t = *col;
		*col = c->tmp_f64;
		c->tmp_f64 = t;
	}
	int *icols[] = { c->tree_parent, c->time_bin,
		reinterpret_cast<int *>(c->id), reinterpret_cast<int *>(c->it_drift),
		reinterpret_cast<int *>(c->it_kick) };
	for (int *col : icols) {
		tdv_launch_gather_i32(c, n, c->perm, col, c->lohi);
		TDV_HIP(c, hipMemcpyAsync(col, c->lohi, n * sizeof(int),
				hipMemcpyDeviceToDevice, c->stream));
	}
	tdv_launch_pack_pm(c);
	if (c->graph_mode) {
		/* every particle active after the sort, the count on the device */
		hipLaunchKernelGGL(k_iota, dim3(tdv_blocks(n, 256)), dim3(256), 0,
				c->stream, (int)n, c->active);
		hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream,
				c->d_nactive, (int)n);
		TDV_HIP(c, hipGetLastError());
		c->nactive = (int)n;
		c->have_domain = 1;
		c->have_tree = 0;
		c->have_top = 0;
		return 0;
	}
	if ((size_t)c->nactive != n) {
		hipLaunchKernelGGL(k_iota, dim3(tdv_blocks(n, 256)), dim3(256), 0,
				c->stream, (int)n, c->active);
		TDV_TRY(set_nactive(c, (int)n));
	}
	tdv_graph_drop(c);               /* columns were swapped */
	(void)hipEventRecord(c->ev[EV_REORDER1], c->stream);
	TDV_HIP(c, hipGetLastError());
	TDV_TRY(tdv_check_err_flags(c));
	c->have_domain = 1;
	c->have_tree = 0;
	c->have_top = 0;
	return 0;
}

int tdv_set_domain(tdv_context *c, double size, const double origin[3])
{
	if (!c || !origin || !(size > 0))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_domain: bad arguments");
	c->dom_size = size;
	for (int d = 0; d < 3; d++) {
		c->dom_origin[d] = origin[d];
		c->dom_center[d] = origin[d] + 0.5 * size;
	}
	c->have_domain_set = 1;
	return 0;
}

int tdv_sort_particles(tdv_context *c)
{
	if (!c || !c->have_domain_set)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_sort_particles before a domain");
	return keys_sort_reorder(c);
}

int tdv_setup_domain(tdv_context *c)
{
	if (!c)
		return TDV_ERR_ARG;
	TDV_TRY(computational_domain(c));
	c->have_domain_set = 1;
	return keys_sort_reorder(c);
}

int tdv_set_top_nodes(tdv_context *c, const void *records, int ntop)
{
	if (!c || !records || ntop <= 0 || ntop > TDV_MAX_TOP || !c->have_domain)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_top_nodes: bad arguments");
	const TopRecordIn *r = static_cast<const TopRecordIn *>(records);
	/* contiguous particle ranges in descending key order covering all */
	size_t expect = c->n;
	for (int t = 0; t < ntop; t++) {
		if (r[t].Npart <= 0 || r[t].First_Part < 0 || r[t].Level < 1
				|| r[t].Level > TDV_NT - 1
				|| (size_t)r[t].First_Part + (size_t)r[t].Npart != expect)
			return tdv_fail(c, TDV_ERR_ARG, "tdv_set_top_nodes: top nodes must "
					"tile the particles in descending key order");
		expect = (size_t)r[t].First_Part;
	}
	if (expect != 0)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_top_nodes: particles left over");
	if (ntop > c->top_cap)
		TDV_TRY(alloc_top(c, ntop));
	std::vector<int> first(ntop), npart(ntop), level(ntop), target(ntop, -1);
	std::vector<uint64_t> key(ntop);
	std::vector<float> pos(3 * (size_t)ntop);
	for (int t = 0; t < ntop; t++) {
		first[t] = r[t].First_Part;
		npart[t] = r[t].Npart;
		level[t] = r[t].Level;
		key[t] = r[t].Key;
		for (int d = 0; d < 3; d++)
			pos[3 * t + d] = r[t].Pos[d];
	}
	auto up = [&](void *d, const void *h, size_t b) {
		return tdv_check(c, hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice,
				c->stream), "top copy");
	};
	TDV_TRY(up(c->top.first, first.data(), ntop * 4));
	TDV_TRY(up(c->top.npart, npart.data(), ntop * 4));
	TDV_TRY(up(c->top.level, level.data(), ntop * 4));
	TDV_TRY(up(c->top.target, target.data(), ntop * 4));
	TDV_TRY(up(c->top.key, key.data(), ntop * 8));
	TDV_TRY(up(c->top.pos, pos.data(), ntop * 12));
	TDV_TRY(up(c->top.ntop, &ntop, 4));
	c->ntop = ntop;
	tdv_graph_drop(c);
	TDV_TRY(tdv_assign_top(c));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	c->have_top = 1;
	c->have_tree = 0;
	return 0;
}

int tdv_domain_decomposition(tdv_context *c)
{
	if (!c)
		return TDV_ERR_ARG;
	TDV_TRY(computational_domain(c));
	c->have_domain_set = 1;
	TDV_TRY(keys_sort_reorder(c));
	TDV_TRY(tdv_build_top_nodes(c));
	c->have_top = 1;
	(void)hipEventRecord(c->ev[EV_TOP1], c->stream);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	float ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_KEYS0], c->ev[EV_KEYS1]);
	c->stage_ms[0] = ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_KEYS1], c->ev[EV_SORT1]);
	c->stage_ms[1] = ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_SORT1], c->ev[EV_REORDER1]);
	c->stage_ms[2] = ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_REORDER1], c->ev[EV_TOP1]);
	c->stage_ms[3] = ms;
	return 0;
}

int tdv_domain_decomposition_targets(tdv_context *c, int ntarget,
		double part_alloc_factor, double imbal_ceil, int *tasks, int max_tasks,
		int *ntop, int *iterations)
{
	if (!c)
		return TDV_ERR_ARG;
	if (ntarget < 1 || ntarget > 512 || !(part_alloc_factor >= 1.0)
			|| !(imbal_ceil > 0.0) || (tasks && max_tasks < 0))
		return tdv_fail(c, TDV_ERR_ARG,
				"tdv_domain_decomposition_targets: bad arguments");
	TDV_TRY(computational_domain(c));
	c->have_domain_set = 1;
	TDV_TRY(keys_sort_reorder(c));
	std::vector<TopRecordIn> rec;
	int it = 0;
	TDV_TRY(tdv_domain_refine(c, ntarget, part_alloc_factor, imbal_ceil, rec,
			&it));
	TDV_TRY(tdv_set_top_nodes(c, rec.data(), (int)rec.size()));
	(void)hipEventRecord(c->ev[EV_TOP1], c->stream);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	float ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_KEYS0], c->ev[EV_KEYS1]);
	c->stage_ms[0] = ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_KEYS1], c->ev[EV_SORT1]);
	c->stage_ms[1] = ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_SORT1], c->ev[EV_REORDER1]);
	c->stage_ms[2] = ms;
	(void)hipEventElapsedTime(&ms, c->ev[EV_REORDER1], c->ev[EV_TOP1]);
	c->stage_ms[3] = ms;
	if (ntop)
		*ntop = (int)rec.size();
	if (iterations)
		*iterations = it;
	if (tasks) {
		if ((int)rec.size() > max_tasks)
			return tdv_fail(c, TDV_ERR_ARG, "tdv_domain_decomposition_targets:"
					" tasks array too small");
		for (size_t i = 0; i < rec.size(); i++)
			tasks[i] = -(rec[i].Target + 1);
	}
	return 0;
}

int tdv_tree_build(tdv_context *c)
{
	if (!c || !c->have_domain || !c->have_top)
		return tdv_fail(c, TDV_ERR_ARG, "tree build before top nodes exist");
	(void)hipEventRecord(c->ev[EV_BUILD0], c->stream);
	TDV_TRY(tdv_build_tree(c));
	(void)hipEventRecord(c->ev[EV_BUILD1], c->stream);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	(void)hipEventElapsedTime(&c->stage_ms[4], c->ev[EV_BUILD0], c->ev[EV_BUILD1]);
	return tdv_check_err_flags(c);
}

int tdv_tree_acceleration(tdv_context *c, int use_bh)
{
	if (!c || !c->have_tree)
		return tdv_fail(c, TDV_ERR_ARG, "walk before a tree build");
	TDV_HIP(c, hipMemsetAsync(c->d_count, 0, 4 * sizeof(unsigned long long),
			c->stream));
	(void)hipEventRecord(c->ev[EV_WALK0], c->stream);
	if (c->par.precision == TDV_WALK_FP32) {
		TDV_TRY(tdv_walk32(c, use_bh));     /* Newton + Ewald in one pass */
		(void)hipEventRecord(c->ev[EV_WALK1], c->stream);
	} else {
		TDV_TRY(tdv_walk64(c, use_bh));
		(void)hipEventRecord(c->ev[EV_WALK1], c->stream);
		if (c->par.periodic)
			TDV_TRY(tdv_walk64_ewald(c, use_bh));
	}
	(void)hipEventRecord(c->ev[EV_EWALD1], c->stream);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	(void)hipEventElapsedTime(&c->stage_ms[5], c->ev[EV_WALK0], c->ev[EV_WALK1]);
	(void)hipEventElapsedTime(&c->stage_ms[6], c->ev[EV_WALK1], c->ev[EV_EWALD1]);
	TDV_HIP(c, hipMemcpy(c->count, c->d_count, sizeof(c->count),
			hipMemcpyDeviceToHost));
	return tdv_check_err_flags(c);
}

int tdv_compute_acceleration(tdv_context *c, int prepare_step)
{
	if (!c || !c->have_domain)
		return tdv_fail(c, TDV_ERR_ARG, "acceleration before a decomposition");
	tdv_launch_safe_last_accel(c);
	if (!c->have_tree)
		TDV_TRY(tdv_tree_build(c));
	if (prepare_step) {
		TDV_TRY(tdv_tree_acceleration(c, 1));
		tdv_launch_safe_last_accel(c);
	}
	return tdv_tree_acceleration(c, 0);
}

static int capture_walk(tdv_context *c, int substep)
{
	if (!c || !c->have_tree || c->par.precision != TDV_WALK_FP32)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_capture_walk: FP32 walk over a "
				"built tree only");
	if (!c->p32 || (c->par.periodic && !c->ewald16_ok)
			|| c->n32_cap < c->node_cap || c->top32_cap < c->top_cap
			|| (c->par.fp32_order != 1 && c->nnodes && !c->lvl_ok))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_capture_walk: run one walk on "
				"this tree first");
	tdv_graph_drop(c);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	TDV_HIP(c, hipStreamBeginCapture(c->stream,
			hipStreamCaptureModeThreadLocal));
	c->capturing = 1;
	c->substep_prep = substep;
	tdv_launch_safe_last_accel_dev(c);
	int r = tdv_walk32(c, 0);
	c->capturing = 0;
	c->substep_prep = 0;
	hipGraph_t g = nullptr;
	const hipError_t e = hipStreamEndCapture(c->stream, &g);
	if (r) {
		if (g)
			(void)hipGraphDestroy(g);
		return r;
	}
	TDV_TRY(tdv_check(c, e, "hipStreamEndCapture"));
	c->graph = g;
	TDV_TRY(tdv_check(c, hipGraphInstantiate(&c->graph_exec, g, nullptr,
			nullptr, 0), "hipGraphInstantiate"));
	return 0;
}

int tdv_capture_walk(tdv_context *c)
{
	return capture_walk(c, 0);
}

extern "C" int tdv_capture_substep_walk(tdv_context *c)
{
	return capture_walk(c, 1);
}

int tdv_launch_walk(tdv_context *c)
{
	if (!c || !c->graph_exec)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_launch_walk: nothing captured");
	(void)hipEventRecord(c->ev[EV_WALK0], c->stream);
	TDV_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
	(void)hipEventRecord(c->ev[EV_WALK1], c->stream);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	(void)hipEventElapsedTime(&c->stage_ms[5], c->ev[EV_WALK0], c->ev[EV_WALK1]);
	c->stage_ms[6] = 0.f;
	TDV_HIP(c, hipMemcpy(c->count, c->d_count, sizeof(c->count),
			hipMemcpyDeviceToHost));
	return tdv_check_err_flags(c);
}

/* The whole regular step of a periodic FP32 run as one HIP graph: the
 * decomposition (keys, sort, reorder, level-2 top nodes), safe_last_accel,
 * the build, the FP32 walk's records, moments and walk, and the kick-drift
 * of one global step.  Captured with the capacities of the last eager step
 * (node arrays, walk records), every count read on the device, columns
 * copied back after the reorder instead of swapped, so a replay finds its
 * buffers; nothing runs at capture. */
int tdv_capture_step(tdv_context *c, double kick, double drift)
{
	if (!c || c->par.precision != TDV_WALK_FP32 || !c->par.periodic)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_capture_step: the FP32 walk of a "
				"periodic box only");
	if (!std::isfinite(kick) || !std::isfinite(drift))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_capture_step: factors not "
				"finite");
	const size_t long_need = (c->n / (TDV_LONG_CHAIN + 1) + 1) * (TDV_NT + 2);
	if (!c->have_tree || !c->p32 || !c->ewald16_ok || !c->partial
			|| !c->gorder || c->n32_cap < c->node_cap || c->top32_cap < c->top_cap
			|| c->top_cap < 64 || c->long_cap < long_need
			|| !c->lvl_nodes || c->lvl_cap < c->node_cap)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_capture_step: run one eager "
				"step (decomposition, FP32 walk) on these particles first");
	tdv_graph_drop(c);
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	/* host-side state the captured calls set; nothing has run yet */
	const int nact0 = c->nactive, dom0 = c->have_domain, top0 = c->have_top,
			tree0 = c->have_tree, lvl0 = c->lvl_ok;
	TDV_HIP(c, hipStreamBeginCapture(c->stream,
			hipStreamCaptureModeThreadLocal));
	c->graph_mode = 1;
	c->capturing = 1;
	int r = computational_domain(c);
	if (!r)
		r = keys_sort_reorder(c);
	if (!r)
		r = tdv_build_top_nodes(c);
	if (!r) {
		c->have_top = 1;
		tdv_launch_safe_last_accel_dev(c);
		r = tdv_build_tree(c);
	}
	if (!r)
		r = tdv_walk32(c, 0);
	if (!r)
		tdv_launch_kick_drift(c, kick, drift);
	c->graph_mode = 0;
	c->capturing = 0;
	hipGraph_t g = nullptr;
	const hipError_t e = hipStreamEndCapture(c->stream, &g);
	c->nactive = nact0;
	c->have_domain = dom0;
	c->have_top = top0;
	c->have_tree = tree0;
	c->lvl_ok = lvl0;
	if (r) {
		if (g)
			(void)hipGraphDestroy(g);
		return r;
	}
	TDV_TRY(tdv_check(c, e, "hipStreamEndCapture"));
	c->step_graph = g;
	TDV_TRY(tdv_check(c, hipGraphInstantiate(&c->step_exec, g, nullptr,
			nullptr, 0), "hipGraphInstantiate"));
	c->step_kick = kick;
	c->step_drift = drift;
	return 0;
}

/* Replay the captured step; the counts and error flags are read after it. */
int tdv_launch_step(tdv_context *c)
{
	if (!c || !c->step_exec)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_launch_step: nothing captured");
	(void)hipEventRecord(c->ev[EV_KEYS0], c->stream);
	TDV_HIP(c, hipGraphLaunch(c->step_exec, c->stream));
	(void)hipEventRecord(c->ev[EV_KICK1], c->stream);
	int h[2];
	TDV_HIP(c, hipMemcpyAsync(&h[0], c->top.ntop, sizeof(int),
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipMemcpyAsync(&h[1], c->top.nnodes, sizeof(int),
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipMemcpyAsync(c->count, c->d_count, sizeof(c->count),
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	float ms = 0.f;
	(void)hipEventElapsedTime(&ms, c->ev[EV_KEYS0], c->ev[EV_KICK1]);
	for (int k = 0; k < 8; k++)
		c->stage_ms[k] = 0.f;
	c->stage_ms[5] = ms;                 /* the whole step */
	c->ntop = h[0];
	c->nactive = (int)c->n;
	c->have_domain = 1;
	c->have_top = 1;
	c->have_tree = h[1] >= 0;
	c->lvl_ok = h[1] >= 0;
	c->nnodes = h[1] >= 0 ? (size_t)h[1] : 0;
	unsigned err = 0;
	TDV_HIP(c, hipMemcpy(&err, c->d_err, sizeof(err), hipMemcpyDeviceToHost));
	if (err & TDV_EBIT_TIES) {
		TDV_HIP(c, hipMemset(c->d_err, 0, sizeof(unsigned)));
		c->have_tree = 0;
		return tdv_fail(c, TDV_ERR_TIES, "tdv_launch_step: a run of more "
				"than 4096 equal 64-bit key prefixes needs the sort's "
				"fallback, which a captured step does not take: run the step "
				"eagerly");
	}
	if (h[1] < 0) {
		TDV_HIP(c, hipMemset(c->d_err, 0, sizeof(unsigned)));
		return tdv_fail(c, TDV_ERR_RANGE, "tdv_launch_step: the tree needs "
				"more nodes than the captured step holds: run the step "
				"eagerly and capture again");
	}
	return tdv_check_err_flags(c);
}

enum { FT_TABLE, FT_EXACT_SPLIT, FT_EXACT_PAIRS };

static int forcetest(tdv_context *c, int nt, const int *idx, double *acc_out,
		int mode);

int tdv_forcetest(tdv_context *c, int nt, const int *idx, double *acc_out)
{
	return forcetest(c, nt, idx, acc_out, FT_TABLE);
}

int tdv_forcetest_exact(tdv_context *c, int nt, const int *idx,
		double *acc_out)
{
	return forcetest(c, nt, idx, acc_out, FT_EXACT_SPLIT);
}

int tdv_forcetest_exact_pairs(tdv_context *c, int nt, const int *idx,
		double *acc_out)
{
	return forcetest(c, nt, idx, acc_out, FT_EXACT_PAIRS);
}

static int forcetest(tdv_context *c, int nt, const int *idx, double *acc_out,
		int mode)
{
	if (!c || nt < 0 || (nt && (!idx || !acc_out)) || !c->n)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_forcetest: bad arguments");
	if (!nt)
		return 0;
	for (int t = 0; t < nt; t++)
		if (idx[t] < 0 || (size_t)idx[t] >= c->n)
			return tdv_fail(c, TDV_ERR_ARG, "tdv_forcetest: index out of range");
	int *di = nullptr;
	double *dout = nullptr;
	int r = dalloc(c, &di, (size_t)nt);
	if (!r)
		r = dalloc(c, &dout, 3 * (size_t)nt);
	if (!r)
		r = tdv_check(c, hipMemcpyAsync(di, idx, nt * sizeof(int),
				hipMemcpyHostToDevice, c->stream), "copy");
	if (!r)
		r = mode == FT_EXACT_SPLIT ? tdv_direct_sum_split(c, nt, di, dout)
				: mode == FT_EXACT_PAIRS ? tdv_direct_sum_exact(c, nt, di, dout)
				: tdv_direct_sum(c, nt, di, dout);
	if (!r)
		r = tdv_check(c, hipMemcpyAsync(acc_out, dout, 3 * nt * sizeof(double),
				hipMemcpyDeviceToHost, c->stream), "copy");
	if (!r)
		r = tdv_check(c, hipStreamSynchronize(c->stream), "forcetest");
	dfree(di);
	dfree(dout);
	return r;
}

int tdv_peano_keys_host(size_t n, const double *x, const double *y,
		const double *z, const double origin[3], double size, int reversed,
		uint64_t *keys_out)
{
	if ((n && (!x || !y || !z || !origin || !keys_out)) || !(size > 0))
		return TDV_ERR_ARG;
	const double fac = (double)(1ull << 63) / size;
	for (size_t i = 0; i < n; i++)
		peano_key_words(x[i], y[i], z[i], origin[0], origin[1], origin[2],
				fac, reversed, &keys_out[2 * i], &keys_out[2 * i + 1]);
	return 0;
}

int tdv_peano_keys(tdv_context *c, size_t n, const double *x, const double *y,
		const double *z, const double origin[3], double size, int reversed,
		uint64_t *keys_out)
{
	if (!c || !x || !y || !z || !keys_out)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_peano_keys: bad arguments");
	if (!n)
		return 0;
	double *d[3] = { nullptr, nullptr, nullptr };
	uint64_t *k = nullptr;
	int r = 0;
	const double *h[3] = { x, y, z };
	for (int q = 0; q < 3 && !r; q++) {
		r = dalloc(c, &d[q], n);
		if (!r)
			r = tdv_check(c, hipMemcpyAsync(d[q], h[q], n * 8,
					hipMemcpyHostToDevice, c->stream), "copy");
	}
	if (!r)
		r = dalloc(c, &k, 2 * n);
	if (!r) {
		tdv_launch_keys(c, n, d[0], d[1], d[2], origin, size, k, reversed);
		r = tdv_check(c, hipMemcpyAsync(keys_out, k, 16 * n,
				hipMemcpyDeviceToHost, c->stream), "copy");
	}
	if (!r)
		r = tdv_check(c, hipStreamSynchronize(c->stream), "keys");
	for (int q = 0; q < 3; q++)
		dfree(d[q]);
	dfree(k);
	return r;
}

int tdv_sort_keys(tdv_context *c, size_t n, const uint64_t *keys,
		uint32_t *perm_out)
{
	if (!c || !keys || !perm_out)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_sort_keys: bad arguments");
	if (!n)
		return 0;
	if (n > c->cap) {        /* grows the scratch; drops particle state */
		TDV_TRY(alloc_particles(c, n));
		c->n = 0;
		c->have_domain = c->have_tree = 0;
	}
	TDV_HIP(c, hipMemcpyAsync(c->key_tmp, keys, 16 * n, hipMemcpyHostToDevice,
			c->stream));
	TDV_TRY(tdv_sort_device(c, n, c->key_tmp, c->perm));
	TDV_HIP(c, hipMemcpyAsync(perm_out, c->perm, 4 * n, hipMemcpyDeviceToHost,
			c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	return tdv_check_err_flags(c);
}

int tdv_ewald_correction(tdv_context *c, size_t n, const double *dr,
		double *f)
{
	if (!c || !dr || !f || !c->have_ewald)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_ewald_correction: no table");
	if (!n)
		return 0;
	double *ddr = nullptr, *df = nullptr;
	int r = dalloc(c, &ddr, 3 * n);
	if (!r)
		r = dalloc(c, &df, 3 * n);
	if (!r)
		r = tdv_check(c, hipMemcpyAsync(ddr, dr, 24 * n, hipMemcpyHostToDevice,
				c->stream), "copy");
	if (!r) {
		tdv_launch_ewald_correction(c, n, ddr, df);
		r = tdv_check(c, hipMemcpyAsync(f, df, 24 * n, hipMemcpyDeviceToHost,
				c->stream), "copy");
	}
	if (!r)
		r = tdv_check(c, hipStreamSynchronize(c->stream), "ewald");
	dfree(ddr);
	dfree(df);
	return r;
}

int tdv_ewald_correction_host(const tdv_context *c, size_t n,
		const double *dr, double *f)
{
	if (!c || !dr || !f || !c->have_ewald
			|| c->ewald_host.size() != 4 * (size_t)TDV_NE3)
		return TDV_ERR_ARG;
	const double b2e = 2 * TDV_NE / c->par.boxsize;
	for (size_t i = 0; i < n; i++)
		ewald_corr(c->ewald_host.data(), dr[3 * i], dr[3 * i + 1],
				dr[3 * i + 2], b2e, f + 3 * i);
	return 0;
}

int tdv_get_ewald_table(tdv_context *c, double *t4)
{
	if (!c || !t4 || !c->have_ewald)
		return tdv_fail(c, TDV_ERR_ARG, "no Ewald table");
	std::vector<double> h(4 * (size_t)TDV_NE3);
	TDV_HIP(c, hipMemcpyAsync(h.data(), c->ewald, h.size() * 8,
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	const double box = c->par.boxsize;
	for (size_t i = 0; i < (size_t)TDV_NE3; i++) {
		/* undo the 1/Box^2, 1/Box normalisation */
		for (int q = 0; q < 3; q++)
			t4[q * TDV_NE3 + i] = h[q * TDV_NE3 + i] * (box * box);
		t4[3 * TDV_NE3 + i] = h[3 * TDV_NE3 + i] * box;
	}
	return 0;
}

int tdv_set_ewald_table(tdv_context *c, const double *t4)
{
	if (!c || !t4 || !c->par.periodic)
		return tdv_fail(c, TDV_ERR_ARG, "Ewald table on a non-periodic run");
	/* Gravity_Periodic_Init, Gravity/periodic.c:215-228 */
	std::vector<double> h(4 * (size_t)TDV_NE3);
	const double box = c->par.boxsize;
	for (size_t i = 0; i < (size_t)TDV_NE3; i++) {
		for (int q = 0; q < 3; q++)
			h[q * TDV_NE3 + i] = t4[q * TDV_NE3 + i] / (box * box);
		h[3 * TDV_NE3 + i] = t4[3 * TDV_NE3 + i] / box;
	}
	if (!c->ewald)
		TDV_TRY(dalloc(c, &c->ewald, h.size()));
	TDV_HIP(c, hipMemcpyAsync(c->ewald, h.data(), h.size() * 8,
			hipMemcpyHostToDevice, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	c->ewald_host.swap(h);
	c->have_ewald = 1;
	c->ewald16_ok = 0;               /* the FP32 records follow the table */
	return 0;
}

int tdv_get_ewald_derivs(tdv_context *c, double *s10)
{
	if (!c || !s10 || !c->par.periodic)
		return tdv_fail(c, TDV_ERR_ARG, "no Ewald derivatives on this run");
	TDV_TRY(tdv_ewald_derivs_ready(c));
	TDV_HIP(c, hipMemcpyAsync(s10, c->ewald_s, 10 * (size_t)TDV_NE3 * 8,
			hipMemcpyDeviceToHost, c->stream));
	return tdv_check(c, hipStreamSynchronize(c->stream), "ewald derivs");
}

int tdv_get_ewald_jacobian(tdv_context *c, double *j6)
{
	if (!c || !j6 || !c->par.periodic)
		return tdv_fail(c, TDV_ERR_ARG, "no Ewald derivatives on this run");
	TDV_TRY(tdv_ewald_derivs_ready(c));
	TDV_HIP(c, hipMemcpyAsync(j6, c->ewald_j, 6 * (size_t)TDV_NE3 * 8,
			hipMemcpyDeviceToHost, c->stream));
	return tdv_check(c, hipStreamSynchronize(c->stream), "ewald jacobian");
}

int tdv_get_walk32_quadrupoles(tdv_context *c, float *nodes, float *top,
		int *node_first)
{
	if (!c || !c->have_tree || !c->q32 || !c->tq32
			|| c->n32_cap < c->nnodes || c->top32_cap < c->ntop)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_get_walk32_quadrupoles: no "
				"FP32 walk with quadrupoles on this tree");
	if (nodes && c->nnodes)
		TDV_HIP(c, hipMemcpyAsync(nodes, c->q32, c->nnodes * 32,
				hipMemcpyDeviceToHost, c->stream));
	if (top && c->ntop)
		TDV_HIP(c, hipMemcpyAsync(top, c->tq32, (size_t)c->ntop * 32,
				hipMemcpyDeviceToHost, c->stream));
	if (node_first && c->nnodes) {
		std::vector<int4> nc(c->nnodes);
		TDV_HIP(c, hipMemcpyAsync(nc.data(), c->node_c, c->nnodes * 16,
				hipMemcpyDeviceToHost, c->stream));
		TDV_HIP(c, hipStreamSynchronize(c->stream));
		for (size_t j = 0; j < c->nnodes; j++)
			node_first[j] = nc[j].w < 0 ? -1 : nc[j].w;
	}
	return tdv_check(c, hipStreamSynchronize(c->stream), "quadrupoles");
}

int tdv_get_domain(tdv_context *c, double *size, double origin[3],
		double center[3])
{
	if (!c || !c->have_domain)
		return tdv_fail(c, TDV_ERR_ARG, "no domain yet");
	if (size)
		*size = c->dom_size;
	for (int d = 0; d < 3; d++) {
		if (origin)
			origin[d] = c->dom_origin[d];
		if (center)
			center[d] = c->dom_center[d];
	}
	return 0;
}

/* Top_Tree_Node layout, domain.h:40-52 */
struct TopRecord {
	uint64_t Key;
	int Target, Level, First_Part, Npart;
	float Pos[3], Mass, CoM[3], Dp[3];
};

int tdv_get_top_nodes(tdv_context *c, void *out, int max, int *ntop)
{
	if (!c || !ntop || !c->have_top)
		return tdv_fail(c, TDV_ERR_ARG, "no top nodes yet");
	const int n = c->ntop;
	*ntop = n;
	if (!out)
		return 0;
	if (n > max)
		return tdv_fail(c, TDV_ERR_ARG, "top node buffer too small");
	std::vector<int> first(n), npart(n), level(n), target(n);
	std::vector<uint64_t> key(n);
	std::vector<float> pos(3 * n), mass(n), com(3 * n), tdp(3 * n);
	std::vector<int> kicked(n);
	auto cp = [&](void *h, const void *d, size_t b) {
		return b ? tdv_check(c, hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost,
				c->stream), "top copy") : 0;
	};
	TDV_TRY(cp(first.data(), c->top.first, n * 4));
	TDV_TRY(cp(npart.data(), c->top.npart, n * 4));
	TDV_TRY(cp(level.data(), c->top.level, n * 4));
	TDV_TRY(cp(target.data(), c->top.target, n * 4));
	TDV_TRY(cp(key.data(), c->top.key, n * 8));
	TDV_TRY(cp(pos.data(), c->top.pos, n * 12));
	TDV_TRY(cp(mass.data(), c->top.mass, n * 4));
	TDV_TRY(cp(com.data(), c->top.com, n * 12));
	TDV_TRY(cp(tdp.data(), c->top.dp, n * 12));
	TDV_TRY(cp(kicked.data(), c->top.kicked, n * 4));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	TopRecord *rec = static_cast<TopRecord *>(out);
	for (int t = 0; t < n; t++) {
		TopRecord r;
		memset(&r, 0, sizeof(r));
		r.Key = key[t];
		r.Target = c->have_tree ? target[t] : -1;
		/* kicked since the last drift: Level < 0 (tree_update.c:46-47) */
		r.Level = c->have_tree && kicked[t] ? -level[t] : level[t];
		r.First_Part = first[t];
		r.Npart = npart[t];
		for (int d = 0; d < 3; d++) {
			r.Pos[d] = pos[3 * t + d];
			r.CoM[d] = c->have_tree ? com[3 * t + d] : 0.f;
			r.Dp[d] = c->have_tree ? tdp[3 * t + d] : 0.f;
		}
		r.Mass = c->have_tree ? mass[t] : 0.f;
		rec[t] = r;
	}
	return 0;
}

/* Tree_Node layout, Gravity/tree.h:23-32 (Float = double) */
int tdv_get_tree(tdv_context *c, void *out, size_t max, size_t *nnodes)
{
	if (!c || !nnodes || !c->have_tree)
		return tdv_fail(c, TDV_ERR_ARG, "no tree yet");
	const size_t n = c->nnodes;
	*nnodes = n;
	if (!out)
		return 0;
	if (n > max)
		return tdv_fail(c, TDV_ERR_ARG, "tree buffer too small");
	if (!n)
		return 0;
	std::vector<double4> a(n), b(n);
	std::vector<int4> cc(n);
	std::vector<int> up(n);
	std::vector<double4> dp(n);
	TDV_HIP(c, hipMemcpyAsync(dp.data(), c->node_dp, n * 32,
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipMemcpyAsync(a.data(), c->node_a, n * 32,
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipMemcpyAsync(b.data(), c->node_b, n * 32,
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipMemcpyAsync(cc.data(), c->node_c, n * 16,
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipMemcpyAsync(up.data(), c->node_up, n * 4,
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	NodeRecord *rec = static_cast<NodeRecord *>(out);
	for (size_t j = 0; j < n; j++) {
		NodeRecord r;
		memset(&r, 0, sizeof(r));
		r.DNext = cc[j].x;
		r.Bitfield = (uint32_t)cc[j].z;
		r.DUp = up[j];
		r.Npart = cc[j].y;
		r.Pos[0] = b[j].x;
		r.Pos[1] = b[j].y;
		r.Pos[2] = b[j].z;
		r.Mass = a[j].w;
		r.CoM[0] = a[j].x;
		r.CoM[1] = a[j].y;
		r.CoM[2] = a[j].z;
		r.Dp[0] = dp[j].x;
		r.Dp[1] = dp[j].y;
		r.Dp[2] = dp[j].z;
		rec[j] = r;
	}
	return 0;
}

int tdv_get_keys(tdv_context *c, uint64_t *keys, int reversed)
{
	if (!c || !keys || !c->have_domain)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_get_keys: no sorted keys");
	const size_t n = c->n;
	const uint64_t *src = c->key;
	if (reversed) {
		tdv_launch_reverse_keys(c, n, c->key, c->key_tmp);
		src = c->key_tmp;
	}
	TDV_HIP(c, hipMemcpyAsync(keys, src, 16 * n, hipMemcpyDeviceToHost,
			c->stream));
	return tdv_check(c, hipStreamSynchronize(c->stream), "tdv_get_keys");
}

int tdv_permute_host_columns(tdv_context *c, int ncols, void **cols,
		const int *elem_bytes)
{
	if (!c || ncols < 0 || (ncols && (!cols || !elem_bytes)) || !c->have_domain)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_permute_host_columns: no sort");
	const size_t n = c->n;
	for (int k = 0; k < ncols; k++)
		if (!cols[k] || (elem_bytes[k] != 4 && elem_bytes[k] != 8))
			return tdv_fail(c, TDV_ERR_ARG, "tdv_permute_host_columns: columns "
					"of 4 or 8 bytes (P.Key: tdv_get_keys)");
	for (int k = 0; k < ncols; k++) {
		const size_t b = elem_bytes[k];
		/* key_tmp holds 16 B per particle, tmp_f64 8 B */
		TDV_HIP(c, hipMemcpyAsync(c->key_tmp, cols[k], n * b,
				hipMemcpyHostToDevice, c->stream));
		if (b == 4)
			tdv_launch_gather_i32(c, n, c->perm, (const int *)c->key_tmp,
					(int *)c->tmp_f64);
		else
			tdv_launch_gather_f64(c, n, c->perm, (const double *)c->key_tmp,
					c->tmp_f64);
		TDV_HIP(c, hipMemcpyAsync(cols[k], c->tmp_f64, n * b,
				hipMemcpyDeviceToHost, c->stream));
	}
	return tdv_check(c, hipStreamSynchronize(c->stream), "permute columns");
}

int tdv_update_particles(tdv_context *c, const double *pos[3],
		const double *mass, const double *acc[3], const double *last_acc_mag)
{
	if (!c || !pos || !c->n)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_update_particles: no particles");
	const size_t b = c->n * sizeof(double);
	for (int d = 0; d < 3; d++) {
		if (!pos[d])
			return tdv_fail(c, TDV_ERR_ARG, "tdv_update_particles: pos");
		TDV_HIP(c, hipMemcpyAsync(c->pos[d], pos[d], b, hipMemcpyHostToDevice,
				c->stream));
		if (acc && acc[d])
			TDV_HIP(c, hipMemcpyAsync(c->acc[d], acc[d], b,
					hipMemcpyHostToDevice, c->stream));
	}
	if (mass)
		TDV_HIP(c, hipMemcpyAsync(c->mass, mass, b, hipMemcpyHostToDevice,
				c->stream));
	if (last_acc_mag)
		TDV_HIP(c, hipMemcpyAsync(c->last_acc, last_acc_mag, b,
				hipMemcpyHostToDevice, c->stream));
	tdv_launch_pack_pm(c);
	return tdv_check(c, hipStreamSynchronize(c->stream), "update_particles");
}

int tdv_set_active(tdv_context *c, int nactive, const int *list)
{
	if (!c || nactive < 0 || (size_t)nactive > c->n || (nactive && !list))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_active: bad arguments");
	for (int i = 0; i < nactive; i++)
		if (list[i] < 0 || (size_t)list[i] >= c->n)
			return tdv_fail(c, TDV_ERR_ARG, "tdv_set_active: index out of range");
	if (nactive)
		TDV_HIP(c, hipMemcpyAsync(c->active, list, nactive * sizeof(int),
				hipMemcpyHostToDevice, c->stream));
	TDV_TRY(set_nactive(c, nactive));
	return tdv_check(c, hipStreamSynchronize(c->stream), "tdv_set_active");
}

int tdv_set_tree_moments(tdv_context *c, const void *tree, size_t nnodes,
		const void *top, int ntop)
{
	if (!c || !c->have_tree || nnodes != c->nnodes || ntop != c->ntop
			|| (nnodes && !tree) || !top)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_tree_moments: shape mismatch");
	const NodeRecord *nr = static_cast<const NodeRecord *>(tree);
	std::vector<double4> a(nnodes);
	for (size_t j = 0; j < nnodes; j++)
		a[j] = make_double4(nr[j].CoM[0], nr[j].CoM[1], nr[j].CoM[2],
				nr[j].Mass);
	const TopRecordIn *tr = static_cast<const TopRecordIn *>(top);
	std::vector<float> tm(ntop), tc(3 * (size_t)ntop);
	for (int t = 0; t < ntop; t++) {
		tm[t] = tr[t].Mass;
		for (int d = 0; d < 3; d++)
			tc[3 * t + d] = tr[t].CoM[d];
	}
	if (nnodes)
		TDV_HIP(c, hipMemcpyAsync(c->node_a, a.data(), nnodes * 32,
				hipMemcpyHostToDevice, c->stream));
	TDV_HIP(c, hipMemcpyAsync(c->top.mass, tm.data(), ntop * 4,
			hipMemcpyHostToDevice, c->stream));
	TDV_HIP(c, hipMemcpyAsync(c->top.com, tc.data(), ntop * 12,
			hipMemcpyHostToDevice, c->stream));
	return tdv_check(c, hipStreamSynchronize(c->stream), "set_tree_moments");
}

int tdv_interaction_counts(tdv_context *c, uint64_t counts[2])
{
	if (!c || !counts)
		return TDV_ERR_ARG;
	counts[0] = c->count[0];
	counts[1] = c->count[1];
	return 0;
}

extern "C" int tdv_interaction_kinds(tdv_context *c, uint64_t counts[4])
{
	if (!c || !counts)
		return TDV_ERR_ARG;
	for (int k = 0; k < 4; k++)
		counts[k] = c->count[k];
	return 0;
}

__global__ void k_active_flags(size_t n, const int *__restrict__ bin,
		int max_bin, uint32_t *__restrict__ flag)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		flag[i] = bin[i] <= max_bin ? 1u : 0u;
}

__global__ void k_active_scatter(size_t n, const uint32_t *__restrict__ flag,
		const uint32_t *__restrict__ off, int *__restrict__ active,
		int *__restrict__ nact)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flag[i])
		active[off[i]] = (int)i;
	if (i == 0)
		*nact = (int)off[n];
}

int tdv_make_active_list(tdv_context *c, int max_active_bin,
		const int *time_bin)
{
	if (!c || !c->n)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_make_active_list: no particles");
	const size_t n = c->n;
	if (time_bin)
		TDV_HIP(c, hipMemcpyAsync(c->time_bin, time_bin, n * sizeof(int),
				hipMemcpyHostToDevice, c->stream));
	uint32_t *flag = reinterpret_cast<uint32_t *>(c->cnt);
	hipLaunchKernelGGL(k_active_flags, dim3(tdv_blocks(n, 256)), dim3(256), 0,
			c->stream, n, c->time_bin, max_active_bin, flag);
	TDV_TRY(tdv_scan_u32(c, flag, c->off, n));
	hipLaunchKernelGGL(k_active_scatter, dim3(tdv_blocks(n, 256)), dim3(256),
			0, c->stream, n, flag, c->off, c->active, c->d_nactive);
	TDV_HIP(c, hipGetLastError());
	int nact = 0;
	TDV_HIP(c, hipMemcpyAsync(&nact, c->d_nactive, sizeof(int),
			hipMemcpyDeviceToHost, c->stream));
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	c->nactive = nact;
	if (nact <= 0)                   /* timestep.c:296-297 */
		return tdv_fail(c, TDV_ERR_ARG, "tdv_make_active_list: no active "
				"particles, bin max " + std::to_string(max_active_bin));
	return 0;
}

int tdv_set_fp32_options(tdv_context *c, int waves, int queue, int split)
{
	if (!c || (waves != 0 && waves != 6 && waves != 8)
			|| (queue != 0 && queue != 1) || split < 0 || split > 256)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_set_fp32_options: waves 0, 6 "
				"or 8, queue 0 or 1, split 0..256");
	c->par.fp32_waves = waves;
	c->par.fp32_queue = queue;
	c->par.fp32_split = split;
	tdv_graph_drop(c);
	return 0;
}

int tdv_stage_times(tdv_context *c, float ms[8])
{
	if (!c || !ms)
		return TDV_ERR_ARG;
	memcpy(ms, c->stage_ms, sizeof(c->stage_ms));
	return 0;
}

} /* extern "C" */
