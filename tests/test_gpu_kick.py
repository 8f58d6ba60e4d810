"""tdv_kick_drift (one global step, every particle active) against the same
FP64 operations on the CPU: v += kick a (kick.c:9-70), x += drift v
(drift.c:12-64), wrapped into the box (periodic.c:30-53); bit for bit.  The
drifted positions reach the walk's packed copy: a direct sum after the
drift equals one of a context given the drifted positions.  With only some
particles active it refuses (hierarchical steps are tdv_kick_halfstep and
tdv_drift_to_sync_point, test_gpu_integrate.py)."""
import numpy as np
import pytest

from tandav_amd import Engine

pytestmark = pytest.mark.gpu


def _wrap(x, box):
    x = x.copy()
    while (x < 0).any():
        x = np.where(x < 0, x + box, x)
    while (x >= box).any():
        x = np.where(x >= box, x - box, x)
    return x


@pytest.mark.parametrize("periodic", [True, False])
def test_kick_drift_matches_cpu(periodic):
    g = np.random.Generator(np.random.Philox(21))
    n = 20000
    pos = g.uniform(0.0, 1.0, (n, 3))
    vel = g.normal(0.0, 1.0, (n, 3))
    acc = g.normal(0.0, 5.0, (n, 3))
    m = np.full(n, 1.0 / n)
    e = Engine(periodic=periodic, softening=0.01, alpha_rel=0.01)
    e.set_particles(pos, m, vel=vel, acc=acc)
    kick, drift = 0.0123, 0.0456
    e.kick_drift(kick, drift)
    got = e.get("pos", "vel")
    v = vel + kick * acc
    x = pos + drift * v
    if periodic:
        x = np.stack([_wrap(x[:, d], 1.0) for d in range(3)], 1)
    np.testing.assert_array_equal(got["vel"], v)
    np.testing.assert_array_equal(got["pos"], x)
    assert e.stage_times()["kick"] > 0
    # the walk's packed positions moved too
    idx = np.array([0, 17, 4242, n - 1])
    f = Engine(periodic=periodic, softening=0.01, alpha_rel=0.01)
    f.set_particles(x, m)
    np.testing.assert_array_equal(e.forcetest(idx), f.forcetest(idx))
    # a subset active: refused
    from tandav_amd import TDVError
    e.set_active(np.arange(0, n, 3, dtype=np.int32))
    with pytest.raises(TDVError, match="every particle active"):
        e.kick_drift(kick, drift)
