"""The oracle's Domain_Decomposition with NTarget > 1 (domain.c:67-145),
pinned by cases whose outcome follows by hand from the reference's rules,
by the single-target case against an independent cell histogram, and by
invariants on clustered keys (the bunches tile the particles in descending
key order, each bunch is the PH cell its key names, the loop stops only
balanced or after 20 passes)."""
import numpy as np
import pytest

import oracle

L2 = np.uint64(58)


def _cell_keys(prefixes):
    """short keys of particles in the given (level, prefix) cells, one
    particle per entry, low bits a fixed pattern"""
    out = []
    for lvl, pref in prefixes:
        sh = 64 - 3 * lvl
        out.append((pref << sh) | ((1 << (sh - 1)) - 12345 % (1 << (sh - 1))))
    return np.array(out, dtype=np.uint64)


def _check_tiling(rec, keys):
    n = len(keys)
    expect = n
    for r in rec:                           # descending key order
        assert r["Npart"] > 0
        assert r["First_Part"] + r["Npart"] == expect
        expect = r["First_Part"]
        sh = np.uint64(64 - 3 * int(r["Level"]))
        low = (np.uint64(1) << sh) - np.uint64(1)
        assert (np.uint64(r["Key"]) & low) == low   # largest key of its cell
        seg = keys[r["First_Part"]:r["First_Part"] + r["Npart"]]
        assert np.all((seg >> sh) == (np.uint64(r["Key"]) >> sh))
    assert expect == 0


def test_one_target_is_the_level2_cells():
    g = np.random.Generator(np.random.Philox(3))
    k = np.sort(g.integers(0, 2 ** 63, 50000, dtype=np.uint64)
                * np.uint64(2) + g.integers(0, 2, 50000, dtype=np.uint64))
    k[:5000] = k[0] | np.uint64(0x3F)       # a crowded cell
    k = np.sort(k)
    rec, it = oracle.domain_targets(k, np.zeros(len(k)), 1)
    cells, first, count = np.unique(k >> L2, return_index=True,
                                    return_counts=True)
    assert it == 0
    np.testing.assert_array_equal(rec["Level"], 2)
    np.testing.assert_array_equal(rec["Key"] >> L2, cells[::-1])
    np.testing.assert_array_equal(rec["First_Part"], first[::-1])
    np.testing.assert_array_equal(rec["Npart"], count[::-1])
    np.testing.assert_array_equal(rec["Target"], -1)


def test_equal_cells_go_round_robin():
    """64 level-2 cells of equal cost, 8 targets: every pass of distribute
    gives the next bunch to the first target furthest below the mean
    (strictly larger deficit), and equal costs keep the key order (stable
    sort), so bunch i in descending key order goes to target i mod 8; the
    split is exact and the loop stops at once."""
    k = np.sort(np.concatenate([_cell_keys([(2, c)] * 5) for c in range(64)]))
    rec, it = oracle.domain_targets(k, np.zeros(len(k)), 8)
    assert it == 0 and len(rec) == 64
    np.testing.assert_array_equal(-(rec["Target"] + 1), np.arange(64) % 8)
    _check_tiling(rec, k)


def test_two_passes_worked_by_hand():
    """Cell A (level-2 prefix 0) holds 3 particles in each of its 8
    level-3 children, child 7's three in level-4 cells 0, 3 and 5; cell B
    (prefix 1) holds 8 particles in its level-3 child 2.  Costs are 0, so
    cost_metric = 1 per particle; 2 targets, mean 16.
    Pass 0: A (24) -> target 0, B (8) -> 1; imbalance 0.5.  Split_Idx starts
    at [0, 1] (bunch indices in key order, B first): A beats B for target
    0 and bunch 1 (A) is already target 1's, so A splits.
    Pass 1: B 8, A7..A0 3 each; greedy gives 17 : 15 (imbalance 1/16 >
    0.05).  No bunch of either target beats its initial Split_Idx (B for
    target 0, A7 for target 1), so B and A7 split: B into its one
    non-empty child, A7 into three level-4 cells of 1.
    Pass 2: 8, seven 3s, three 1s; greedy gives 16 : 16 and the loop
    stops after two splitting passes."""
    ka = []
    for c in range(7):
        ka += [(3, c)] * 3
    ka += [(4, (7 << 3) | j) for j in (0, 3, 5)]
    kb = [(3, (1 << 3) | 2)] * 8
    k = np.sort(np.concatenate([_cell_keys(ka), _cell_keys(kb)]))
    rec, it = oracle.domain_targets(k, np.zeros(len(k)), 2)
    assert it == 2
    lv = rec["Level"].tolist()
    assert lv == [3, 4, 4, 4, 3, 3, 3, 3, 3, 3, 3]
    np.testing.assert_array_equal(rec["Npart"], [8, 1, 1, 1, 3, 3, 3, 3, 3,
                                                 3, 3])
    tasks = -(rec["Target"] + 1)
    np.testing.assert_array_equal(tasks, [0, 0, 0, 1, 1, 1, 1, 0, 1, 0, 1])
    _check_tiling(rec, k)


def test_fewer_bunches_than_targets_is_an_error():
    k = np.sort(_cell_keys([(3, 9)] * 100))
    with pytest.raises(ValueError):
        oracle.domain_targets(k, np.zeros(len(k)), 2)


@pytest.mark.parametrize("ntarget", [2, 3, 8, 16, 64])
def test_clustered_invariants(ntarget):
    g = np.random.Generator(np.random.Philox(40 + ntarget))
    n = 120000
    parts = [g.integers(0, 2 ** 63, n // 3, dtype=np.uint64) * np.uint64(2)]
    for c in range(4):                      # four dense clumps
        base = np.uint64(g.integers(0, 2 ** 20)) << np.uint64(44)
        parts.append(base + g.integers(0, 2 ** (40 - 4 * c), n // 6,
                                       dtype=np.uint64))
    k = np.unique(np.concatenate(parts))
    cost = g.integers(0, 4000, len(k)).astype(np.float64)
    rec, it = oracle.domain_targets(k, cost, ntarget)
    _check_tiling(rec, k)
    assert rec["Level"].min() >= 2 and rec["Level"].max() <= 21
    tasks = -(rec["Target"] + 1)
    assert tasks.min() >= 0 and tasks.max() < ntarget
    # the loop stops balanced (in exact integer sums, within float error)
    # unless it ran its 20 passes
    w = np.maximum(1, cost).astype(np.int64)
    bc = np.add.reduceat(w, rec["First_Part"][::-1])[::-1]
    tc = np.bincount(tasks, weights=bc, minlength=ntarget)
    tn = np.bincount(tasks, weights=rec["Npart"], minlength=ntarget)
    if it <= 19:
        assert np.max(np.abs(tc - tc.mean()) / tc.mean()) < 0.05 + 1e-3
        assert np.max(np.abs(tn - len(k) / ntarget) / (len(k) / ntarget)) \
            < 0.1 + 1e-9
