"""The persistent work-queue walk and the captured walk graph against the
plain FP32 walk: the same groups run the same code, so the results must be
bit-identical, on every replay, with eager work and synchronisations between
the replays, for the full active set and for a subset."""
import numpy as np
import pytest

from tandav_amd import Engine, ics

pytestmark = pytest.mark.gpu


def _engine(pos, m, queue):
    e = Engine(periodic=True, softening=1 / 1280., alpha_rel=0.01,
               precision="fp32", compat=False, ewald_size=0.125,
               ewald_accuracy=1.0, fp32_order=2, fp32_criterion=3,
               fp32_queue=queue)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    e.compute_acceleration(False)
    return e


def test_queue_walk_matches_grid_walk():
    pos, _, m = ics.zeldovich_box(32, seed=7)
    a = _engine(pos, m, 0)
    b = _engine(pos, m, 1)
    ga, gb = a.get("acc", "cost"), b.get("acc", "cost")
    np.testing.assert_array_equal(ga["acc"], gb["acc"])
    np.testing.assert_array_equal(ga["cost"], gb["cost"])
    assert a.interaction_counts() == b.interaction_counts()


def test_graph_replays_match_eager_walk():
    pos, _, m = ics.zeldovich_box(32, seed=7)
    e = _engine(pos, m, 1)
    e.capture_walk()
    st = e.get("pos", "mass", "acc", "last_acc_mag")
    n = len(m)
    g = np.random.Generator(np.random.Philox(3))
    subsets = [np.arange(n), np.sort(g.choice(n, n // 7, replace=False)),
               np.sort(g.choice(n, 100, replace=False)), np.arange(n)]
    for k, sub in enumerate(subsets):
        e.set_active(sub)
        e.update_particles(st["pos"], acc=st["acc"],
                           last_acc_mag=st["last_acc_mag"])
        e.compute_acceleration(False)          # eager: safe_last_accel + walk
        want = e.get("acc", "cost", "last_acc_mag")
        want_counts = e.interaction_counts()
        for rep in range(3):
            e.update_particles(st["pos"], acc=st["acc"],
                               last_acc_mag=st["last_acc_mag"])
            e.launch_walk()
            got = e.get("acc", "cost", "last_acc_mag")
            for f in ("acc", "cost", "last_acc_mag"):
                np.testing.assert_array_equal(got[f], want[f],
                                              err_msg="%s subset %d rep %d"
                                              % (f, k, rep))
            assert e.interaction_counts() == want_counts
            # eager work between replays
            e.get("pos")
            e.forcetest([0, 1])
