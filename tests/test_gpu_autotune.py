"""The autotuner on the GPU: it times every candidate on one tree, leaves
its pick set and the particles as they were, and the walk with the picked
options takes the same interactions as with the defaults."""
import numpy as np
import pytest

from tandav_amd import Engine, ics
from tandav_amd.autotune import QUEUE, SPLITS, WAVES, autotune

pytestmark = pytest.mark.gpu


def _engine(pos, m):
    e = Engine(periodic=True, softening=1 / 1280., alpha_rel=0.01,
               precision="fp32", compat=False, ewald_size=0.125,
               ewald_accuracy=1.0, fp32_order=2, fp32_criterion=3)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    return e


def test_autotune_picks_and_keeps_results():
    pos, _, m = ics.zeldovich_box(32, seed=7)
    n = len(m)
    a, b = _engine(pos, m), _engine(pos, m)
    before = b.get("acc", "last_acc_mag")
    sub = np.sort(np.random.Generator(np.random.Philox(4)).choice(
        n, 200, replace=False)).astype(np.int32)
    r = autotune(b, active_sets=[sub], repeats=1)
    assert len(r["trials"]["full"]) == len(WAVES) * len(QUEUE)
    assert len(r["trials"]["sparse"]) == len(SPLITS)
    assert all(t["walk_ms"] > 0 for t in r["trials"]["full"])
    p = b.params
    assert (p.fp32_waves, p.fp32_queue, p.fp32_split) == (
        r["waves"], r["queue"], r["split"])
    after = b.get("acc", "last_acc_mag")
    np.testing.assert_array_equal(after["acc"], before["acc"])
    np.testing.assert_array_equal(after["last_acc_mag"],
                                  before["last_acc_mag"])
    for act in (np.arange(n, dtype=np.int32), sub):
        for e in (a, b):
            e.set_active(act)
            e.update_particles(before_pos(a), acc=before["acc"],
                               last_acc_mag=before["last_acc_mag"])
            e.tree_acceleration(False)
        ga, gb = a.get("acc", "cost"), b.get("acc", "cost")
        np.testing.assert_array_equal(ga["cost"][act], gb["cost"][act])
        assert a.interaction_counts() == b.interaction_counts()
        err = (np.linalg.norm(ga["acc"][act] - gb["acc"][act], axis=1)
               / np.linalg.norm(ga["acc"][act], axis=1))
        assert err.max() < 1e-4


def before_pos(e):
    return e.get("pos")["pos"]
