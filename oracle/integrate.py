"""CPU restatement of Tandav's KDK integrator on the integer timeline, in
numpy.  TEST INFRASTRUCTURE ONLY: the checker of the engine's
tdv_kick_halfstep, tdv_drift_to_sync_point and tdv_set_new_timebins.

  it2time        Integer_Time2Integration_Time (timestep.c:340-372)
  kick_step      Particle_Kick_Step (kick.c:81-87; comoving.c:30-39)
  drift_step     Particle_Drift_Step (drift.c:115-121; comoving.c:41-50)
  kick_halfstep  Kick_First/Second_Halfstep (kick.c:9-70)
  drift_to_sync  Drift_To_Sync_Point (drift.c:12-64) with
                 Periodic_Constrain_Particles_To_Box (periodic.c:30-53)
  timebins       set_new_particle_timebins (timestep.c:171-201),
                 get_physical_timestep, cosmological_timestep (:495-526),
                 timestep2timebin (:380-388), Acceleration_Physical
                 (unit.c:55-71)
  NaturalSpline  gsl_interp_cspline as comoving.c:146-150 uses it (natural
                 end conditions, c_i = y''_i / 2), restated from the
                 textbook construction; pinned against scipy's natural
                 CubicSpline in tests/test_oracle_integrate.py.

How each is pinned: the spline against scipy; kick/drift factors of
non-comoving runs against the closed form t(it) = Time.Begin + it
Step_Min; timebins by brute force over the powers of two (the bin is the
largest b with 2^b Step_Min < dt) and by hand-worked float cases; the
comoving tables (computed by the engine's host code) against scipy's
adaptive quadrature and the Einstein-de Sitter closed forms.  GSL itself
is absent here, so agreement with gsl_integration_qag and gsl_spline at
the last bits is parity unpinned.
"""
import numpy as np

NBITS = 32                  # N_INT_BINS: intime_t is uint32_t (globals.h:33)
KNOTS = 512                 # TABLESIZE (comoving.c:5)
FLT_MAX = float(np.finfo(np.float32).max)


class NaturalSpline:
    """Cubic spline through (x, y) with zero second derivative at both
    ends; eval() picks the interval by bisection (x_i <= a < x_{i+1}) and
    evaluates y_i + u (b_i + u (c_i + u d_i))."""

    def __init__(self, x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        n = len(x)
        h = np.diff(x)
        m = n - 2
        # tridiagonal system for c_1..c_{n-2} (c = y'' / 2)
        a = np.zeros(m)
        b = 2.0 * (h[:-1] + h[1:])
        r = 3.0 * ((y[2:] - y[1:-1]) / h[1:] - (y[1:-1] - y[:-2]) / h[:-1])
        a[1:] = h[1:-1]                      # sub/super diagonal
        cp = np.zeros(m)
        dp = np.zeros(m)
        for i in range(m):                   # Thomas algorithm
            den = b[i] - (a[i] * cp[i - 1] if i else 0.0)
            cp[i] = (h[i + 1] if i + 1 < m else 0.0) / den
            dp[i] = (r[i] - (a[i] * dp[i - 1] if i else 0.0)) / den
        c = np.zeros(n)
        for i in range(m - 1, -1, -1):
            c[i + 1] = dp[i] - (cp[i] * c[i + 2] if i + 1 < m else 0.0)
        self.x, self.y, self.c = x, y, c

    def __call__(self, a):
        a = np.asarray(a, np.float64)
        x, y, c = self.x, self.y, self.c
        i = np.clip(np.searchsorted(x, a, side="right") - 1, 0, len(x) - 2)
        h = x[i + 1] - x[i]
        u = a - x[i]
        b = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0
        d = (c[i + 1] - c[i]) / (3.0 * h)
        return y[i] + u * (b + u * (c[i] + u * d))


class Timeline:
    """Time.Begin, Step_Min and Step_Max (timestep.c:130-147) and, when
    comoving, the splines of the kick and drift tables."""

    def __init__(self, comoving, time_begin, time_end, tables=None):
        self.comoving = bool(comoving)
        self.begin = float(time_begin)
        if comoving:
            self.step_max = np.log(time_end) - np.log(time_begin)
            a, kick, drift = tables
            self.kick_spline = NaturalSpline(a, kick)
            self.drift_spline = NaturalSpline(a, drift)
        else:
            self.step_max = float(time_end) - float(time_begin)
        self.step_min = self.step_max / float(1 << (NBITS - 1))

    def it2time(self, it):
        it = np.asarray(it, np.uint32).astype(np.float64)
        if self.comoving:
            return self.begin * np.exp(it * self.step_min)
        return self.begin + it * self.step_min

    def kick_step(self, i0, i1):
        a0, a1 = self.it2time(i0), self.it2time(i1)
        if not self.comoving:
            return a1 - a0
        return self.kick_spline(a1) - self.kick_spline(a0)

    def drift_step(self, i0, i1):
        a0, a1 = self.it2time(i0), self.it2time(i1)
        if not self.comoving:
            return a1 - a0
        return self.drift_spline(a1) - self.drift_spline(a0)


def kick_halfstep(tl, active, time_bin, it_kick, acc, vel):
    """Kick_*_Halfstep: returns (vel, it_kick) after the kick."""
    vel = vel.copy()
    it_kick = it_kick.copy()
    i = np.asarray(active)
    step = (np.left_shift(np.uint32(1), time_bin[i].astype(np.uint32))
            >> np.uint32(1)).astype(np.uint32)
    i0 = it_kick[i]
    i1 = (i0 + step).astype(np.uint32)
    dt = tl.kick_step(i0, i1)
    vel[i] = vel[i] + dt[:, None] * acc[i]
    it_kick[i] = i1
    return vel, it_kick


def _wrap(x, box):
    x = x.copy()
    while (x < 0).any():
        x = np.where(x < 0, x + box, x)
    while (x >= box).any():
        x = np.where(x >= box, x - box, x)
    return x


def drift_to_sync(tl, active, it_drift, it_next, pos, vel, periodic, box):
    """Drift_To_Sync_Point: returns (pos, it_drift)."""
    pos = pos.copy()
    it_drift = it_drift.copy()
    i = np.asarray(active)
    dt = tl.drift_step(it_drift[i], np.full(len(i), it_next, np.uint32))
    p = pos[i] + dt[:, None] * vel[i]
    if periodic:
        p = np.stack([_wrap(p[:, d], box) for d in range(3)], 1)
    pos[i] = p
    it_drift[i] = it_next
    return pos, it_drift


def timebins(tl, grav_acc, time_bin, accuracy, softening, dt_max_global,
             max_active_bin, expansion=1.0, hubble=1.0, accel_factor=1.0):
    """set_new_particle_timebins: returns (bins, bin_min, bin_max, ok)."""
    acc = np.sqrt((grav_acc * grav_acc).sum(1))
    if tl.comoving:
        acc = acc * accel_factor
    eps = 105.0 / 32.0 * softening
    with np.errstate(divide="ignore"):
        if tl.comoving:
            f = accuracy * np.sqrt(2 * expansion * eps / acc)
        else:
            f = accuracy * np.sqrt(2 * eps / acc)
    dts = np.minimum(np.float64(np.float32(FLT_MAX)),
                     f.astype(np.float32).astype(np.float64))
    dt = dts.astype(np.float32).astype(np.float64)
    if tl.comoving:
        dt = dt * hubble
    dt = np.minimum(dt, float(np.float32(dt_max_global)))
    ok = bool(np.all(dt >= tl.step_min))
    _, e = np.frexp(tl.step_max / dt)
    want = NBITS - 1 - e
    allowed = np.maximum(max_active_bin, time_bin)
    b = np.minimum(want, allowed).astype(np.int32)
    return b, int(b.min()), int(b.max()), ok
