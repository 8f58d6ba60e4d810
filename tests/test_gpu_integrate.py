"""The device KDK integrator against oracle/integrate.py on the same
inputs: Kick_*_Halfstep, Drift_To_Sync_Point and set_new_particle_timebins
with particles spread over many time bins and places on the integer
timeline, an active subset, non-comoving (bit for bit) and comoving (the
device's exp and the host's may differ in the last bit, so to 1e-13)."""
import numpy as np
import pytest

from oracle import integrate as I
from tandav_amd import Engine, TDVError, engine, ics

pytestmark = pytest.mark.gpu


def _state(n, seed, comoving):
    g = np.random.Generator(np.random.Philox(seed))
    pos = g.uniform(0.0, 1.0, (n, 3))
    vel = g.normal(0.0, 1.0, (n, 3))
    acc = g.normal(0.0, 5.0, (n, 3))
    bins = g.integers(0, 12, n).astype(np.int32)
    it_k = (g.integers(0, 1 << 18, n) << 12).astype(np.uint32)
    it_d = (g.integers(0, 1 << 18, n) << 12).astype(np.uint32)
    act = np.sort(g.choice(n, n // 3, replace=False)).astype(np.int32)
    if comoving:
        tl = engine.timeline(True, 0.1, 1.0, hubble=np.sqrt(8 * np.pi / 0.9),
                             omega_matter=0.3, omega_lambda=0.7)
    else:
        tl = engine.timeline(False, 0.25, 4.0)
    return pos, vel, acc, bins, it_k, it_d, act, tl


def _oracle_timeline(tl):
    tables = engine.comoving_tables(tl) if tl.comoving else None
    return I.Timeline(tl.comoving, tl.time_begin, tl.time_end, tables)


@pytest.mark.parametrize("comoving", [False, True])
def test_kick_and_drift_match_oracle(comoving):
    n = 30000
    pos, vel, acc, bins, it_k, it_d, act, tl = _state(n, 7, comoving)
    m = np.full(n, 1.0 / n)
    e = Engine(periodic=True, softening=0.01)
    e.set_particles(pos, m, vel=vel, acc=acc, time_bin=bins)
    e.set_timeline(tl)
    e.set_integer_times(it_d, it_k)
    e.set_active(act)
    otl = _oracle_timeline(tl)
    # two half kicks (first and second), then the drift to the next sync
    v, ik = vel, it_k
    for _ in range(2):
        e.kick_halfstep()
        v, ik = I.kick_halfstep(otl, act, bins, ik, acc, v)
    got = e.get("vel", "pos")
    gd, gk = e.integer_times()
    np.testing.assert_array_equal(gk, ik)
    it_next = int(it_d.max()) + (1 << 12)
    e.drift_to_sync_point(it_next)
    x, idr = I.drift_to_sync(otl, act, it_d, it_next, pos, v, True, 1.0)
    got2 = e.get("pos")
    gd, _ = e.integer_times()
    np.testing.assert_array_equal(gd, idr)
    if comoving:
        np.testing.assert_allclose(got["vel"], v, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(got2["pos"], x, rtol=1e-13, atol=1e-15)
    else:
        np.testing.assert_array_equal(got["vel"], v)
        np.testing.assert_array_equal(got2["pos"], x)
    # inactive particles did not move
    off = np.setdiff1d(np.arange(n), act)
    np.testing.assert_array_equal(got2["pos"][off], pos[off])
    # the walk's packed positions follow the drift
    idx = np.array([int(act[0]), int(act[-1]), int(off[0])])
    f = Engine(periodic=True, softening=0.01)
    f.set_particles(got2["pos"], m)
    np.testing.assert_array_equal(e.forcetest(idx), f.forcetest(idx))
    # beyond the timeline's end: refused
    with pytest.raises(TDVError):
        e.drift_to_sync_point((1 << 31) + 1)


@pytest.mark.parametrize("comoving", [False, True])
def test_timebins_match_oracle(comoving):
    pos, _, m = ics.clustered_box(1 << 14, nhalo=16, seed=5)
    n = len(m)
    e = Engine(periodic=True, softening=1e-3, theta_bh=0.1)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    ga = e.get("grav_acc")["grav_acc"]
    amag = np.linalg.norm(ga, axis=1)
    assert amag.max() / np.median(amag) > 10      # many bins
    tl = (engine.timeline(True, 0.02, 1.0, hubble=2.0) if comoving
          else engine.timeline(False, 0.0, 1.0))
    e.set_timeline(tl)
    otl = _oracle_timeline(tl)
    g = np.random.Generator(np.random.Philox(8))
    start = g.integers(0, 20, n).astype(np.int32)
    e.make_active_list(31, time_bin=start)
    kw = dict(time_int_accuracy=0.01, softening=1e-3, dt_max_global=0.01,
              max_active_bin=9)
    if comoving:
        kw.update(expansion_factor=0.05, hubble_parameter=3.1,
                  grav_accel_factor=1 / 0.05 ** 2)
    lo, hi = e.set_new_timebins(**kw)
    got = e.get("time_bin")["time_bin"]
    want, wlo, whi, ok = I.timebins(
        otl, ga, start, kw["time_int_accuracy"], kw["softening"],
        kw["dt_max_global"], kw["max_active_bin"],
        kw.get("expansion_factor", 1.0), kw.get("hubble_parameter", 1.0),
        kw.get("grav_accel_factor", 1.0))
    assert ok
    np.testing.assert_array_equal(got, want)
    assert (lo, hi) == (wlo, whi)
    assert len(np.unique(got)) > 4
    # a step below the timeline's resolution: the reference's Assert
    with pytest.raises(TDVError, match="resolution"):
        e.set_new_timebins(1e-9, 1e-3, 0.01, 9)
