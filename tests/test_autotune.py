"""The autotuner's selection and bookkeeping on a stand-in engine (CPU):
it times every candidate on the same restored state, keeps the fastest,
samples whole 64-particle groups, and leaves the particles as they were."""
import numpy as np

from tandav_amd import autotune as at


class FakeEngine:
    def __init__(self, n, cost):
        self.n = n
        self.cost = cost                      # (waves, queue, split) -> ms
        self.opts = (0, 0, 0)
        self.active = np.arange(n)
        self.acc = np.ones((n, 3))
        self.restored = 0
        self.log = []

    def get(self, *names):
        return {"pos": np.zeros((self.n, 3)), "acc": self.acc.copy(),
                "last_acc_mag": np.ones(self.n)}

    def set_fp32_options(self, waves=None, queue=None, split=None):
        self.opts = (waves, queue, split)

    def set_active(self, a):
        self.active = np.asarray(a)

    def update_particles(self, pos, acc=None, last_acc_mag=None):
        self.acc = acc.copy()
        self.restored += 1

    def tree_acceleration(self, use_bh):
        self.log.append((self.opts, len(self.active)))
        self.acc = self.acc + 1.0             # a walk changes acc

    def stage_times(self):
        return {"walk": self.cost(*self.opts)}


def test_picks_fastest_and_restores():
    def cost(w, q, s):             # unsplit sparse walks are slow
        return 10.0 + (w == 6) * -3.0 + q * 1.0 + (5.0 if s == 1 else
                                                   abs(s - 32) / 8.0)
    e = FakeEngine(1000, cost)
    sparse = np.arange(0, 100)
    r = at.autotune(e, active_sets=[np.arange(500), sparse], repeats=2)
    assert (r["waves"], r["queue"], r["split"]) == (6, 0, 32)
    assert e.opts == (6, 0, 32)
    # every trial started from the original accelerations
    assert e.restored == len(e.log) + 1
    np.testing.assert_array_equal(e.acc, np.ones((1000, 3)))
    assert len(e.active) == 1000
    assert len(r["trials"]["full"]) == len(at.WAVES) * len(at.QUEUE)
    assert all(t["n_active"] == 100 for t in r["trials"]["sparse"])


def test_group_sample_takes_whole_groups():
    idx = at._group_sample(64 * 100 + 10, 7, seed=1)
    assert len(idx) % 64 in (0, 10)
    starts = idx[::64]
    assert np.all(starts % 64 == 0)
    assert np.all(np.diff(idx) > 0)
    np.testing.assert_array_equal(at._group_sample(300, 7, seed=1),
                                  np.arange(300))


def test_keeps_default_within_margin():
    def cost(w, q, s):             # 1% faster than the default: noise
        return 10.0 - (0.1 if (w, q) == (8, 1) else 0.0)
    e = FakeEngine(500, cost)
    r = at.autotune(e, repeats=1)
    assert (r["waves"], r["queue"], r["split"]) == (0, 0, 0)
    # the first of the candidates within the margin, not the fastest
    def cost2(w, q, s):
        return {(0, 0): 10.0, (0, 1): 10.5, (6, 0): 9.0}.get((w, q), 8.9)
    r = at.autotune(FakeEngine(500, cost2), repeats=1)
    assert (r["waves"], r["queue"]) == (6, 0)
