"""Autotuner of the FP32 group walk's launch options.

Group size is the wavefront (64 particles, one per lane) and the walk keeps
its interaction list in registers and the scalar cache, so what is left to
pick per workload is the schedule:

* ``waves``: the register budget, 0 (the compiler's), 6 or 8 wavefronts per
  SIMD (``fp32_waves``);
* ``queue``: one wavefront launched per group, or per (group, part) work
  item when the walk splits (0), or the persistent work-queue kernel (1);
* ``split``: for substeps with few active groups, how many parts each
  group's subtree walks may be cut into (1 = never; n >= 2 = up to n).

Of the candidates within MARGIN of the fastest, the first in the order
above is kept (the default schedule first), so noise does not flip it.

None of them changes the interactions a particle takes (waves and queue not
even the rounding), so they are timed on one tree: the walk of the full
active set for ``waves`` x ``queue`` (on a sample of whole groups when the
set is larger than the chip needs), then the sparsest of the caller's active
sets for ``split``.  Each trial restores the particles' accelerations and
|a_prev| first, so every trial walks the same work; the best options are
left set on the engine and the particles as they were.
"""
import numpy as np

WAVES = (0, 6, 8)
QUEUE = (0, 1)
SPLITS = (1, 8, 16, 32, 64, 128, 256)
SAMPLE_GROUPS = 65536
MARGIN = 0.03          # candidates this close to the fastest count as equal


def _trial(e, state, active, repeats):
    best = float("inf")
    for _ in range(repeats):
        e.set_active(active)
        e.update_particles(state["pos"], acc=state["acc"],
                           last_acc_mag=state["last_acc_mag"])
        e.tree_acceleration(False)
        best = min(best, float(e.stage_times()["walk"]))
    return best


def _group_sample(n, groups, seed):
    """Whole 64-particle PH groups, sorted, at most `groups` of them."""
    ng = (n + 63) // 64
    if ng <= groups:
        return np.arange(n, dtype=np.int32)
    g = np.random.Generator(np.random.Philox(seed))
    pick = np.sort(g.choice(ng, groups, replace=False))
    idx = (pick[:, None] * 64 + np.arange(64)[None, :]).reshape(-1)
    return idx[idx < n].astype(np.int32)


def autotune(e, active_sets=(), repeats=2, sample_groups=SAMPLE_GROUPS,
             seed=3):
    """Time the FP32 walk's launch options on engine `e` (a built tree, one
    FP32 walk done) and leave the fastest set.  `active_sets`: particle index
    arrays of the substeps to tune `split` for (the sparsest is used).
    Returns {"waves", "queue", "split", "trials"}."""
    state = e.get("pos", "acc", "last_acc_mag")
    n = len(state["pos"])
    full = _group_sample(n, sample_groups, seed)
    trials = {"full": [], "sparse": []}
    for w in WAVES:
        for q in QUEUE:
            e.set_fp32_options(waves=w, queue=q, split=1)
            trials["full"].append({"waves": w, "queue": q,
                                   "walk_ms": _trial(e, state, full,
                                                     repeats)})
    # the first candidate (the default first) within MARGIN of the fastest:
    # choices inside the timing noise do not flip from run to run
    fastest = min(r["walk_ms"] for r in trials["full"])
    best = next(r for r in trials["full"]
                if r["walk_ms"] <= (1.0 + MARGIN) * fastest)
    split = 0
    sets = [np.asarray(a, dtype=np.int32) for a in active_sets if len(a)]
    if sets:
        sparse = min(sets, key=len)
        for sp in SPLITS:
            e.set_fp32_options(waves=best["waves"], queue=best["queue"],
                               split=sp)
            trials["sparse"].append({"split": sp, "n_active": len(sparse),
                                     "walk_ms": _trial(e, state, sparse,
                                                       repeats)})
        split = min(trials["sparse"], key=lambda r: r["walk_ms"])["split"]
    e.set_fp32_options(waves=best["waves"], queue=best["queue"], split=split)
    e.set_active(np.arange(n, dtype=np.int32))
    e.update_particles(state["pos"], acc=state["acc"],
                       last_acc_mag=state["last_acc_mag"])
    return {"waves": best["waves"], "queue": best["queue"], "split": split,
            "sample_particles": len(full), "trials": trials}
