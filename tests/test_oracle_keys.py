"""Pins the oracle's Peano-Hilbert keys (peano.c:122-276), index sort
(sort.c:90-341), reorder (aux.c:187-300) and domain (domain.c:732-847,
select.c:19-80) by what the mathematics fixes, not by retyping them.

Keys: on the 2^L grid (L = 1..5) the level-L prefixes of the cell-centre keys
form a bijection onto 0..8^L-1, consecutive prefixes are face-adjacent cells
(the Hilbert property), every level refines the previous one, the curve
starts in the origin cell, and the reversed layout is the triplet
permutation of the standard layout.  A dropped Gray step, a swapped exchange
or a wrong shift breaks adjacency or the bijection.  Which of the
adjacency-preserving variants (axis relabelling) the reference uses is fixed
only by its source text: the restatement follows it line by line, and the
variant is "parity unpinned" against a compiled reference.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.cpu


def grid_keys(L, reversed_=False):
    n = 1 << L
    i = np.arange(n)
    gx, gy, gz = np.meshgrid(i, i, i, indexing="ij")
    c = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    x = (c + 0.5) / n
    k = oracle.peano_keys(x[:, 0], x[:, 1], x[:, 2], [0, 0, 0], 1.0,
                          reversed_=reversed_)
    return c, oracle.keys_to_u128(k)


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_hilbert_bijection_and_adjacency(L):
    c, keys = grid_keys(L)
    # level-L prefix = key bits [128-3L, 127]
    prefix = np.array([k >> (128 - 3 * L) for k in keys], dtype=np.int64)
    assert sorted(prefix.tolist()) == list(range(8 ** L))
    order = np.argsort(prefix)
    cells = c[order]
    step = np.abs(np.diff(cells, axis=0)).sum(1)
    assert (step == 1).all()
    assert (cells[0] == 0).all()
    # bits below the used levels are the finer levels of the centre point,
    # the two lowest bits are always zero
    assert all(k & 3 == 0 for k in keys)


def test_levels_refine():
    c3, k3 = grid_keys(3)
    c2, k2 = grid_keys(2)
    p3 = {tuple(cc): k >> (128 - 9) for cc, k in zip(c3, k3)}
    p2 = {tuple(cc): k >> (128 - 6) for cc, k in zip(c2, k2)}
    for cc, v in p3.items():
        assert v >> 3 == p2[tuple(np.array(cc) // 2)]


def test_reversed_is_triplet_permutation():
    g = np.random.Generator(np.random.Philox(3))
    x = g.uniform(0, 1, (2000, 3))
    o = [0.0, 0.0, 0.0]
    std = oracle.keys_to_u128(oracle.peano_keys(x[:, 0], x[:, 1], x[:, 2],
                                                o, 1.0))
    rev = oracle.keys_to_u128(oracle.peano_keys(x[:, 0], x[:, 1], x[:, 2],
                                                o, 1.0, reversed_=True))
    for s, r in zip(std, rev):
        assert r & 7 == 0                       # explicit level 0
        for lvl in range(1, 42):
            ts = (s >> (128 - 3 * lvl)) & 7
            tr = (r >> (3 * lvl)) & 7
            assert ts == tr
        assert (r >> 126) == ((s >> 2) & 3)     # level 42 loses its top bit


def test_key_depends_only_on_quantised_cell():
    g = np.random.Generator(np.random.Philox(4))
    n = 1 << 10
    cells = g.integers(0, n, (500, 3))
    a = (cells + 0.25) / n
    b = (cells + 0.75) / n
    ka = oracle.keys_to_u128(oracle.peano_keys(a[:, 0], a[:, 1], a[:, 2],
                                               [0, 0, 0], 1.0))
    kb = oracle.keys_to_u128(oracle.peano_keys(b[:, 0], b[:, 1], b[:, 2],
                                               [0, 0, 0], 1.0))
    for u, v in zip(ka, kb):
        assert u >> (128 - 30) == v >> (128 - 30)


def test_qsort_index_sorts_and_permutes():
    g = np.random.Generator(np.random.Philox(5))
    for n in (1, 2, 31, 33, 65, 1000, 50000):
        k = g.integers(0, 2 ** 63, (n, 2), dtype=np.uint64)
        k[: n // 3, 1] = 7                      # many equal high words
        perm = oracle.qsort_index(k, nthreads=4)
        assert sorted(perm.tolist()) == list(range(n))
        ks = oracle.keys_to_u128(k[perm])
        assert all(a < b for a, b in zip(ks, ks[1:]))


def test_select_is_order_statistic_without_element_zero():
    g = np.random.Generator(np.random.Philox(6))
    for n in (3, 4, 10, 101, 1000, 4097):
        d = g.normal(size=n)
        k = n >> 1
        got = oracle.select(k, d)
        want = np.sort(d[1:])[k - 1]
        assert got == want


def test_domain_nonperiodic_closed_form():
    g = np.random.Generator(np.random.Philox(8))
    x = g.standard_cauchy((3001, 3))
    c, size, o = oracle.domain_nonperiodic(x[:, 0], x[:, 1], x[:, 2])
    n = x.shape[0]
    for d in range(3):
        assert c[d] == np.sort(x[1:, d])[(n >> 1) - 1]
    assert size == 2.001 * np.abs(x - c).max()
    np.testing.assert_array_equal(o, c - 0.5 * size)
