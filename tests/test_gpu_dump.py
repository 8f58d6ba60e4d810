"""Particles below the key resolution, GPU against the oracle.

build_subtree (Gravity/tree_build.c:337-342) gives a particle whose key
matches the previous one down to level 41 no node of its own: it is
"dumped" into the level-41 node it reached (P.Tree_Parent), it starts no
collapse event and leaves last_key / last_parent as they were; if the next
collapse does not take its cell, that level-41 node stays an inner node
over the dumped particle and a level-42 node holding the one before.  Two
constructions on the 32^3 periodic box: partners one level-42 cell apart
(keys equal but for the last triplet) in the bulk, where the next event
collapses them into a bundle, and after the last particle of three top
nodes, where no event follows and the level-41 node survives; then exact
copies of particles (equal keys), which the build dumps too and which the
sort keeps in input order (the reference's quicksort in an order of its
own: only the ids of the copies may differ).  Everything else -- keys,
top nodes, every node, Tree_Parent, Cost, accelerations -- bit for bit as
in test_gpu_parity.py."""
import numpy as np
import pytest

import oracle
from tandav_amd import ics
from test_gpu_parity import _compare, _pipeline_pair

pytestmark = pytest.mark.gpu

L42 = 2.0 ** -42          # a level-42 cell of the unit box


def _keys(p):
    return oracle.peano_keys(p[:, 0], p[:, 1], p[:, 2], [0, 0, 0], 1.0)


def _snap(p):
    """p moved to the centre of its level-42 cell, the cell index made even
    on every axis (so the next cell up each axis shares the level-41 cell)."""
    c = np.floor(p / L42)
    c -= c % 2
    return (c + 0.5) * L42, c


def _partner(p, axis):
    a, c = _snap(p)
    b = a.copy()
    b[axis] = (c[axis] + 1.5) * L42
    return a, b


def _dump_box():
    pos, _, _ = ics.zeldovich_box(32, seed=7)
    pos = pos.copy()
    k = _keys(pos)
    cell = k[:, 1] >> np.uint64(58)
    extra = []
    # after the last particle of three top nodes: a partner whose key is
    # larger and equal down to level 41
    for cl in (3, 17, 40):
        idx = np.flatnonzero(cell == cl)
        j = idx[np.lexsort((k[idx, 0], k[idx, 1]))[-1]]
        for axis in range(3):
            a, b = _partner(pos[j], axis)
            ka, kb = _keys(a[None])[0], _keys(b[None])[0]
            if (ka[1] == kb[1] and ka[0] >> np.uint64(5) == kb[0] >> np.uint64(5)
                    and kb[0] > ka[0]):
                pos[j] = a
                extra.append(b)
                break
    assert len(extra) == 3
    g = np.random.Generator(np.random.Philox(3))
    for j in g.choice(len(pos), 6, replace=False):
        a, b = _partner(pos[j], 0)
        pos[j] = a
        extra.append(b)
    x = np.vstack([pos, extra])
    return x, np.full(len(x), 1.0 / len(x))


@pytest.mark.parametrize("prepare", [True, False])
def test_particles_below_key_resolution(prepare):
    x, m = _dump_box()
    s, e = _pipeline_pair(x, m, True, oracle.cached_ewald_table(),
                          prepare=prepare, softening=1 / 1280.,
                          alpha_rel=0.005, nthreads=16)
    assert s.ndumped == 9
    tr = s.tree
    tp = s.tree_parent
    inner = tp[(tp >= 0)]
    # the three dumped after a top node's last particle keep a level-41
    # inner node as their parent
    assert np.sum((tr["Bitfield"][inner] & 63 == 41)
                  & (tr["DNext"][inner] > 0)) == 3
    _compare(s, e)


def test_particles_with_equal_keys():
    pos, _, _ = ics.zeldovich_box(32, seed=7)
    g = np.random.Generator(np.random.Philox(4))
    copies = pos[g.choice(len(pos), 7, replace=False)]
    x = np.vstack([pos, copies, copies[:2]])       # pairs and triples
    m = np.full(len(x), 1.0 / len(x))
    s, e = _pipeline_pair(x, m, True, oracle.cached_ewald_table(),
                          softening=1 / 1280., alpha_rel=0.005, nthreads=16)
    got = e.get("id", "pos")
    # copies may come in another order; the positions do not tell them apart
    np.testing.assert_array_equal(got["pos"], s.positions())
    ids_g, ids_s = got["id"].copy(), s.id.copy()
    tied = np.flatnonzero(ids_g != ids_s)
    for i in tied:
        same = np.flatnonzero((x == x[ids_g[i]]).all(1))
        assert ids_s[i] in same
    s.id[:] = ids_g                  # compare everything else bit for bit
    _compare(s, e)
