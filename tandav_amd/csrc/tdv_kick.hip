/*
 * The KDK step of a run on one global time step, on the device, in FP64:
 * every particle is active and in the same time bin, so the half-step kicks
 * (kick.c:9-70) and the drift (drift.c:12-64) take one factor each, the
 * caller's Particle_Kick_Step / Particle_Drift_Step (comoving.c:30-50 on
 * comoving runs, the plain time difference otherwise): v += dt_kick a,
 * x += dt_drift v, then Periodic_Constrain_Particles_To_Box
 * (periodic.c:30-53).  The walk's packed positions follow the drift.
 * Hierarchical steps, where each particle kicks and drifts from its own
 * place on the integer timeline, are tdv_kick_halfstep and
 * tdv_drift_to_sync_point (tdv_integrate.hip).
 */
#include "tdv_internal.h"

__global__ void k_kick(int nact, const int *__restrict__ active,
		const double *__restrict__ ax, const double *__restrict__ ay,
		const double *__restrict__ az, double *__restrict__ vx,
		double *__restrict__ vy, double *__restrict__ vz, double dt)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nact)
		return;
	const int i = active[g];
	vx[i] = vx[i] + dt * ax[i];
	vy[i] = vy[i] + dt * ay[i];
	vz[i] = vz[i] + dt * az[i];
}

/* while (x < 0) x += box; while (x >= box) x -= box, with the loops
 * bounded: a position many boxes out is wrapped in one step */
__device__ __forceinline__ double wrap_box(double x, double box)
{
	if (!(fabs(x) < 64.0 * box))
		return x - floor(x / box) * box;
	while (x < 0)
		x += box;
	while (x >= box)
		x -= box;
	return x;
}

__global__ void k_drift(size_t n, double *__restrict__ x,
		double *__restrict__ y, double *__restrict__ z,
		const double *__restrict__ vx, const double *__restrict__ vy,
		const double *__restrict__ vz, double dt, int periodic, double box,
		double4 *__restrict__ pm)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double px = x[i] + dt * vx[i], py = y[i] + dt * vy[i],
			pz = z[i] + dt * vz[i];
	if (periodic) {
		px = wrap_box(px, box);
		py = wrap_box(py, box);
		pz = wrap_box(pz, box);
	}
	x[i] = px;
	y[i] = py;
	z[i] = pz;
	const double4 o = pm[i];
	pm[i] = make_double4(px, py, pz, o.w);
}

/* the kernels of one global step: every particle active */
void tdv_launch_kick_drift(tdv_context *c, double kick, double drift)
{
	if (kick != 0 && c->n)
		hipLaunchKernelGGL(k_kick, dim3(tdv_blocks(c->n, 256)), dim3(256), 0,
				c->stream, (int)c->n, c->active, c->acc[0], c->acc[1],
				c->acc[2], c->vel[0], c->vel[1], c->vel[2], kick);
	if (drift != 0 && c->n)
		hipLaunchKernelGGL(k_drift, dim3(tdv_blocks(c->n, 256)), dim3(256), 0,
				c->stream, c->n, c->pos[0], c->pos[1], c->pos[2], c->vel[0],
				c->vel[1], c->vel[2], drift, c->par.periodic,
				c->par.boxsize, c->pm);
}

extern "C" int tdv_kick_drift(tdv_context *c, double kick, double drift)
{
	if (!c || !c->n || !std::isfinite(kick) || !std::isfinite(drift))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_kick_drift: no particles or "
				"factors not finite");
	if ((size_t)c->nactive != c->n)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_kick_drift: one global step "
				"needs every particle active; hierarchical steps use "
				"tdv_kick_halfstep and tdv_drift_to_sync_point");
	(void)hipEventRecord(c->ev[EV_KICK0], c->stream);
	tdv_launch_kick_drift(c, kick, drift);
	(void)hipEventRecord(c->ev[EV_KICK1], c->stream);
	TDV_HIP(c, hipGetLastError());
	TDV_HIP(c, hipStreamSynchronize(c->stream));
	(void)hipEventElapsedTime(&c->stage_ms[7], c->ev[EV_KICK0],
			c->ev[EV_KICK1]);
	return 0;
}
