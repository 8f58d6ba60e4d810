"""Hand-worked cases that pin the oracle's decision rules where the closed
forms of test_oracle_walk.py (theta = 0, alpha = 0: everything opened)
cannot: the walks' opening tests at theta > 0 and alpha > 0, the Ewald
walk's straddle bound and size cap, its any-zero-component skip, the
collapse threshold of the tree build and the orientation of the PH curve.

Each case is a periodic unit box (Domain.Size = 1, so node sizes are
2^-level and node cells are known) with particles on small lattices
("clumps") in chosen cells.  Costs are read per particle (P.Cost counts
interact and interact_with_ewald_cube calls, tree_accel.c:391,
tree_periodic.c:437).  Where a case adds a far clump A to a box that holds
the target's own clump B, the target's walk of B is the same with and
without A (same particles, keys and subtree), so Cost(B + A) - Cost(B) is
exactly what the walks take from A; that difference is worked out by hand
in each docstring from the reference's rules:

  main BH walk, tree_accel.c:297-353: open when s^2 > theta r^2;
  main relative walk, tree_accel.c:211-290: open when M s^2 > r^4 |a| alpha
      / G, else open when the target lies within 0.6 s of the centre on
      every axis, else take the node;
  Ewald walks, tree_periodic.c:216-422: want to open by the same criteria
      (theta^2 for BH) or the 0.6 s test; open only when the node straddles
      the half box (|ds_k| > (Box - s) / 2 on an axis) or s > 0.2 Box,
      else take the node; bundle particles are skipped when any component
      of their separation is 0;
  top nodes, tree_accel.c:137-176 and tree_periodic.c:112-177: the same
      with theta (not squared) and the straddle test on the CoM separation.

tests/test_oracle_mutations.py plants a plausible mistake in each rule of
oracle/orc_sim.c and checks that these cases catch it.
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.cpu


def lattice(level, cell, shape, frac=0.9):
    """shape[0] x shape[1] x shape[2] points evenly inside the level-`level`
    cell `cell`, filling the middle `frac` of it on each axis."""
    s = 1.0 / (1 << level)
    lo = np.asarray(cell, dtype=np.float64) * s
    axes = [lo[d] + s * (0.5 * (1 - frac) + frac * (np.arange(shape[d]) + 0.5)
                         / shape[d]) for d in range(3)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([a.ravel() for a in g], 1)


PMASS = 1e-3          # every particle's mass, the same in every run


def run(x, theta=0.1, alpha=0.0, last_acc=None, use_bh=True, active=None):
    """One walk of the oracle over particles x (masses PMASS); returns
    (Cost per input particle, the Sim)."""
    n = len(x)
    m = np.full(n, PMASS)
    s = oracle.Sim(x, m, periodic=True, softening=1e-5, theta_bh=theta,
                   alpha_rel=alpha, nthreads=1)
    s.setup_domain()
    s.domain_decomposition()
    s.tree_build()
    ids = s.id.astype(np.int64)
    if last_acc is not None:
        la = np.broadcast_to(np.asarray(last_acc, dtype=np.float64), (n,))
        s.last_acc_mag[:] = la[ids]
    if active is not None:
        on = np.zeros(n, bool)
        on[active] = True
        s.time_bin[:] = np.where(on[ids], 0, 1)
        s.set_active_bin(0)
    s.tree_acceleration(use_bh=use_bh)
    cost = np.zeros(n)
    cost[ids] = s.cost
    return cost, s


def find_node(tree, level, lo):
    """Index of the node of `level` whose cell has lower corner `lo`."""
    s = 1.0 / (1 << level)
    lvl = tree["Bitfield"] & 63
    for k in np.flatnonzero((lvl == level) & (tree["Npart"] > 0)):
        if np.allclose(np.floor(tree["CoM"][k] / s) * s, lo, atol=0):
            return int(k)
    raise AssertionError("no node at level %d, corner %s" % (level, lo))


def children(tree, k):
    """Direct children of node k, following the skip links."""
    out, c, end = [], k + 1, k + tree["DNext"][k]
    while c < end:
        out.append(c)
        c = c + tree["DNext"][c] if tree["DNext"][c] > 0 else c + 1
    return out


def opened_cost(tree, k):
    """Cost of a far target's walk of node k when k is opened and every
    child node is taken whole: bundles give one interaction per particle,
    other children one each."""
    ch = children(tree, k)
    return sum(int(tree["Npart"][c]) if tree["DNext"][c] < 0 else 1
               for c in ch)


# clump A: 100 particles filling the level-4 cell (8,5,5), so A's node is
# that cell (x in [0.5, 0.5625), centre 0.53125) under a chain of the
# level-3 cell (4,2,2) (centre x 0.5625) and the level-2 top node (2,1,1);
# clump B: 200 targets filling the level-3 cell (0,2,2), x from 0.017 to
# 0.108, so |dx| to A's centre is 0.42 ... 0.51.
A_CELL = (8, 5, 5)
QA = 0.53125


def clump_a():
    return lattice(4, A_CELL, (5, 5, 4))


def clump_b():
    return lattice(3, (0, 2, 2), (8, 5, 5), frac=0.83)


def test_bh_walks_theta_and_straddle():
    """theta_BH = 0.1, prepare-step walks.  Main BH walk of A: the top node
    and its root (s^2 = 1/16) open since r < 0.79; the level-3 node
    (s^2 = 1/64) is taken since r > 0.395: one interaction.  Ewald BH walk
    of A: the top node and root want to open and are larger than 0.2 Box
    (walked); the level-3 node wants to open (s^2 > theta^2 r^2) and
    straddles (|0.5625 - x| > 0.4375 for every target); A's level-4 node
    wants to open (1/256 > 0.01 r^2) and straddles iff
    |0.53125 - x| > (1 - 1/16) / 2: then it is opened and its children
    taken (level-5 nodes: 1/1024 < 0.01 r^2), else it is taken whole.  So
    Cost(B+A) - Cost(B) = 1 + opened_cost(A) for straddling targets and
    1 + 1 for the others."""
    a, b = clump_a(), clump_b()
    cab, s = run(np.vstack([b, a]))
    cb, _ = run(b)
    d = cab[:len(b)] - cb
    qa = find_node(s.tree, 4, np.array(A_CELL) / 16.0)
    assert s.tree["Npart"][qa] == 100
    full = opened_cost(s.tree, qa)
    assert full > 60                    # mostly bundles of A's particles
    straddle = np.abs(QA - b[:, 0]) > 0.5 * (1 - 1 / 16.0)
    assert straddle.any() and (~straddle).any()
    np.testing.assert_array_equal(d[straddle], 1 + full)
    np.testing.assert_array_equal(d[~straddle], 2)


def test_ewald_skips_any_zero_component():
    """As test_bh_walks_theta_and_straddle, with one particle of A moved
    (within its level-6 cell, so the tree keeps its shape) to the y of a
    straddling target: the Ewald walks of the straddling targets with that
    y (the lattice repeats it) skip it (dy = 0, tree_periodic.c:241-244),
    the main walk does not (r^2 != 0); nothing else changes."""
    a, b = clump_a(), clump_b()
    t = int(np.flatnonzero((np.abs(QA - b[:, 0]) > 0.47)
                           & (np.abs(b[:, 1] - 0.3333) < 1e-3))[0])
    k = int(np.argmin(np.abs(a[:, 1] - b[t, 1]) + 10 * (a[:, 1] > b[t, 1])))
    assert abs(a[k, 1] - b[t, 1]) < 1e-3
    assert np.floor(a[k, 1] * 64) == np.floor(b[t, 1] * 64)
    c0, _ = run(np.vstack([b, a]))
    a[k, 1] = b[t, 1]
    c1, _ = run(np.vstack([b, a]))
    hit = (b[:, 1] == b[t, 1]) & (np.abs(QA - b[:, 0]) > 0.5 * (1 - 1 / 16.))
    assert hit.sum() > 1
    np.testing.assert_array_equal(c1[:len(b)], c0[:len(b)] - hit)


def test_top_node_theta_not_squared():
    """theta_BH = 0.5 with A's top node at 0.42 < r < 0.53 from every
    target: both walks take the top node whole (main: 1/16 <= 0.5 r^2,
    i.e. r >= 0.354; Ewald: the same unsquared test, and not inside), so
    Cost(B+A) - Cost(B) = 2.  With theta^2 = 0.25 the top node would be
    opened for the targets at r < 0.5."""
    a, b = clump_a(), clump_b()
    cab, _ = run(np.vstack([b, a]), theta=0.5)
    cb, _ = run(b, theta=0.5)
    r = np.linalg.norm(b - a.mean(0), axis=1)
    assert (r > 0.354).all() and (r < 0.5).sum() > 50
    np.testing.assert_array_equal(cab[:len(b)] - cb, 2)


def test_ewald_top_node_theta_not_squared():
    """theta_BH = 0.3, clump A2: two 100-particle clumps in the level-4
    cells (8,5,5) and (8,5,6), i.e. two level-3 children of the top node
    (2,1,1).  Targets of B at r > 0.46 from A2's CoM: the main walk takes
    the top node (1/16 <= 0.3 r^2 for r >= 0.456); the Ewald walk's top-node
    test is the same unsquared test, so it does not want to open it either
    (and the target is not inside): one interaction each, Cost(B+A2) -
    Cost(B) = 2.  With theta^2 there the Ewald walk would want to open it,
    find it straddling (|dx| > 0.375), and take the two level-3 nodes."""
    a2 = np.vstack([lattice(4, (8, 5, 5), (5, 5, 4)),
                    lattice(4, (8, 5, 6), (5, 5, 4))])
    b = clump_b()
    cab, _ = run(np.vstack([b, a2]), theta=0.3)
    cb, _ = run(b, theta=0.3)
    d = a2.mean(0) - b
    d -= np.round(d)
    r = np.linalg.norm(d, axis=1)
    far = r > 0.46
    assert far.sum() > 20 and (np.abs(d[far, 0]) > 0.375).all()
    np.testing.assert_array_equal(cab[:len(b)][far] - cb[far], 2)


def test_relative_criterion_exponents():
    """Relative walks (alpha 0.02) with |a_prev| chosen per target so that
    k = |a| alpha / G * r^4 / M (r to A's CoM, M = A's mass; every node of
    A's chain has that CoM and mass) is 0.01 or 0.003.  Main walk: open
    when s^2 > k: k = 0.01 opens the top node, root (1/16) and level 3
    (1/64) and takes A's level-4 node (1/256): one interaction; k = 0.003
    opens it too (its children, s^2 = 1/1024 < k, are taken):
    opened_cost.  Ewald walk: level 3 straddles for every target, A's
    level-4 node straddles for the targets of the BH case; with k = 0.01
    it is not wanted open (taken: 1), with k = 0.003 it is wanted open and
    opened where it straddles.  None of the nodes is within 0.6 s of a
    target."""
    a, b = clump_a(), clump_b()
    x = np.vstack([b, a])
    n, nb = len(x), len(b)
    mass_a = len(a) * PMASS
    alpha = 0.02
    com = a.mean(0)
    dr = com - b
    dr -= np.round(dr)
    r4 = (dr * dr).sum(1) ** 2
    straddle = np.abs(QA - b[:, 0]) > 0.5 * (1 - 1 / 16.0)
    for k in (0.01, 0.003):
        acc = np.ones(n)
        acc[:nb] = k * mass_a / (alpha * r4)
        cab, s = run(x, alpha=alpha, last_acc=acc, use_bh=False)
        accb = np.ones(nb)
        accb[:] = acc[:nb]
        cb, _ = run(b, alpha=alpha, last_acc=accb, use_bh=False)
        d = cab[:nb] - cb
        qa = find_node(s.tree, 4, np.array(A_CELL) / 16.0)
        full = opened_cost(s.tree, qa)
        if k == 0.01:
            np.testing.assert_array_equal(d, 2)
        else:
            np.testing.assert_array_equal(d[straddle], 2 * full)
            np.testing.assert_array_equal(d[~straddle], full + 1)


@pytest.mark.parametrize("axis", [0, 1])
def test_main_walk_inside_test_at_0_6_s(axis):
    """A target in A's top node, alone in the level-4 cell next to A's
    (+x or +y), offset by f s from the centre of A's cell (s = 1/16) on
    `axis` and by 0.1 s on the others; relative walk with |a_prev| so large
    that the criterion never opens.  The main walk opens the nodes that
    hold the target (0.6 s test) and takes every other node; A's node is
    within 0.6 s of the target iff f < 0.6: then it is opened and its
    children taken (opened_cost), else it is taken (1).  The Ewald walk
    takes the level-3 node (the target is inside it, it does not straddle
    and 1/8 <= 0.2): the same for both offsets.  So Cost(f = 0.59) -
    Cost(f = 0.61) = opened_cost - 1."""
    a = clump_a()
    s4 = 1.0 / 16
    centre = (np.array(A_CELL) + 0.5) * s4
    costs, tree = [], None
    for f in (0.59, 0.61):
        t = centre + 0.1 * s4
        t[axis] = centre[axis] + f * s4
        x = np.vstack([a, t[None, :]])
        c, s = run(x, alpha=0.02, last_acc=1e30, use_bh=False,
                   active=[len(a)])
        costs.append(c[-1])
        tree = s.tree
    qa = find_node(tree, 4, np.array(A_CELL) / 16.0)
    assert costs[0] - costs[1] == opened_cost(tree, qa) - 1


def test_ewald_size_cap():
    """theta_BH = 0.08.  Clump C: four 70-particle clumps filling the
    level-4 cells (6,4,4), (6,5,4) [level-3 cell (3,2,2)] and (6,4,6),
    (6,5,6) [(3,2,3)], top node (1,1,1); targets: B's particles with
    x >= 0.031.  Main BH walk of C: top node, root and both level-3 nodes
    open (r < 0.44), the level-4 nodes are taken (r > 0.22): 4.  Ewald BH
    walk: the top node wants to open (1/16 > 0.08 r^2), does not straddle
    (CoM separation < 0.375 per axis) and is larger than 0.2 Box: walked;
    the root likewise; each level-3 node wants to open (theta^2) and is
    taken (no straddle, 1/8 <= 0.2 Box): 2.  So Cost(B+C) - Cost(B) = 6."""
    b = clump_b()
    c = np.vstack([lattice(4, cell, (7, 5, 2)) for cell in
                   ((6, 4, 4), (6, 5, 4), (6, 4, 6), (6, 5, 6))])
    tg = np.flatnonzero(b[:, 0] >= 0.031)
    cbc, s = run(np.vstack([b, c]), theta=0.08)
    cb, _ = run(b, theta=0.08)
    # the premises: distances to the level-3 and level-4 CoMs, straddling
    tr = s.tree
    lv = tr["Bitfield"] & 63
    for level, lo, hi in ((3, 0.0, np.sqrt(1 / 64. / 0.08)),
                          (4, np.sqrt(1 / 256. / 0.08), 1.0)):
        ks = np.flatnonzero((lv == level) & (tr["CoM"][:, 0] > 0.25)
                            & (tr["CoM"][:, 0] < 0.5) & (tr["Npart"] > 64))
        assert len(ks) == (2 if level == 3 else 4)
        for k in ks:
            d = tr["CoM"][k] - b[tg]
            d -= np.round(d)
            r = np.linalg.norm(d, axis=1)
            assert (r > lo).all() and (r < hi).all()
    dc = c.mean(0) - b[tg]
    assert (np.abs(dc - np.round(dc)) < 0.375).all()
    np.testing.assert_array_equal(cbc[tg] - cb[tg], 6)


# ---------------------------------------------------------------------------
# the tree build's collapse threshold (collapse_last_branch,
# tree_build.c:404-433; VECTOR_SIZE = 64)
# ---------------------------------------------------------------------------

def _ph_first_last_cells(level, parent_level, parent_cell):
    """The children of a cell in PH order (keys of their centres)."""
    s = 1.0 / (1 << level)
    k = 1 << (level - parent_level)
    base = np.asarray(parent_cell) * k
    cells = np.array([base + np.array([i, j, l]) for i in range(k)
                      for j in range(k) for l in range(k)])
    c = (cells + 0.5) * s
    keys = oracle.peano_keys(c[:, 0], c[:, 1], c[:, 2], [0, 0, 0], 1.0)
    order = np.lexsort((keys[:, 0], keys[:, 1]))
    return cells[order]


def _build(x):
    n = len(x)
    s = oracle.Sim(x, np.full(n, 1.0 / n), periodic=True, softening=1e-5,
                   nthreads=1)
    s.setup_domain()
    s.domain_decomposition()
    s.tree_build()
    return s


@pytest.mark.parametrize("k", [63, 64, 65])
def test_collapse_threshold_at_64(k):
    """A top node (level-2 cell (1,1,1)) holding clump C1 of k particles in
    its first level-3 child (PH order) and 40 particles in its last.  When
    the first particle of the second clump arrives it starts a new branch
    at level 3 and the finished sibling cell, C1's, is collapsed into one
    bundle iff it holds <= 64 particles (collapse_last_branch's first
    test): for k = 63, 64 C1's level-3 node is a bundle of k particles
    starting at the top node's first particle; for k = 65 it stays a node
    (DNext > 0) with 65 particles."""
    kids = _ph_first_last_cells(3, 2, (1, 1, 1))
    first, last = kids[0], kids[-1]
    shapes = {63: (3, 3, 7), 64: (4, 4, 4), 65: (5, 13, 1)}
    c1 = lattice(3, first, shapes[k])
    c2 = lattice(3, last, (2, 4, 5))
    s = _build(np.vstack([c1, c2]))
    tr = s.tree
    top = s.top
    assert len(top) == 1 and top["Npart"][0] == k + 40
    node = find_node(tr, 3, first / 8.0)
    assert tr["Npart"][node] == k
    if k <= 64:
        assert tr["DNext"][node] == -(int(top["First_Part"][0]) + 1)
    else:
        assert tr["DNext"][node] > 0


def test_collapse_falls_back_to_last_parent():
    """C1 = 65 particles in the first level-4 cell of the top node's first
    level-3 child, then 63 particles in the first seven level-5 children
    (PH order) of that child's last level-4 cell and one particle alone in
    its last level-5 child, then 40 particles in the top node's last
    level-3 child.  When that 40-clump's first particle arrives, the
    finished sibling (the first level-3 node, 129 particles) is too big, so
    collapse_last_branch collapses last_parent instead: the parent of the
    lone particle, which is the level-4 cell holding 64 particles -- a
    bundle of 64.  The 65-particle level-4 cell stays a node."""
    l3 = _ph_first_last_cells(3, 2, (1, 1, 1))
    l4 = _ph_first_last_cells(4, 3, l3[0])
    l5 = _ph_first_last_cells(5, 4, l4[-1])
    c1a = lattice(4, l4[0], (5, 13, 1))
    c1b = np.vstack([lattice(5, l5[i], (3, 3, 1)) for i in range(7)])
    c1c = lattice(5, l5[7], (1, 1, 1))
    c2 = lattice(3, l3[-1], (2, 4, 5))
    s = _build(np.vstack([c1a, c1b, c1c, c2]))
    tr = s.tree
    n4a = find_node(tr, 4, l4[0] / 16.0)
    n4b = find_node(tr, 4, l4[-1] / 16.0)
    assert tr["Npart"][n4a] == 65 and tr["DNext"][n4a] > 0
    assert tr["Npart"][n4b] == 64 and tr["DNext"][n4b] < 0


def test_small_top_node_is_discarded_at_64():
    """finalise_subtree discards a subtree whose root holds <= 64
    particles (tree_build.c:459-466): a top node of 64 particles gets no
    nodes (its Target is the node count at that point), one of 65 does."""
    for k, built in ((64, False), (65, True)):
        c = lattice(3, (2, 2, 2), {64: (4, 4, 4), 65: (5, 13, 1)}[k])
        far = lattice(3, (6, 6, 6), (5, 5, 5))
        s = _build(np.vstack([c, far]))
        top = s.top
        j = int(np.flatnonzero(top["Npart"] == k)[0])
        has_root = any((s.tree["Bitfield"] & 63)[i] == 2
                       and s.tree["Npart"][i] == k for i in range(len(s.tree)))
        assert has_root == built


# ---------------------------------------------------------------------------
# orientation of the Hilbert curve (Peano_Key, peano.c:122-195)
# ---------------------------------------------------------------------------

def test_level1_octants_follow_gray_code_z_major():
    """peano.c:132-153 starts the inverse undo at q = 2^63, one bit above
    the coordinates (X < 2^63): there P = 2^63 - 1 and no coordinate has
    bit 63 set, so X0 exchanges all its 63 bits with X1 and then with X2,
    leaving (X0, X1, X2) = (z, x, y) before the levels proper (q = 2^62 ...
    2, which never touch a coordinate's bit 62).  The Gray encoding then
    gives X1 ^= X0, X2 ^= X1 at bit 62 (its t correction only reaches
    bits below the top), and the interleave puts X0 X1 X2 from the most
    significant bit down.  So the octant of level-1 triplet t is gray(t) =
    t ^ (t >> 1) read as (z, x, y), z the most significant bit: in x y z
    order 000 001 011 010 110 111 101 100 become (0,0,0) (0,0,1) (0,1,1)
    (0,1,0) (1,1,0) (1,1,1) (1,0,1) (1,0,0) as (z, x, y).  A curve without
    the extra bit (q from 2^62, Skilling's own loop), or with the axes
    relabelled, orders the octants differently."""
    c = (np.array([[i, j, k] for i in range(2) for j in range(2)
                   for k in range(2)]) + 0.5) / 2.0
    keys = oracle.peano_keys(c[:, 0], c[:, 1], c[:, 2], [0, 0, 0], 1.0)
    t = (keys[:, 1] >> np.uint64(61)).astype(int)       # level-1 triplet
    xyz = np.floor(c * 2).astype(int)
    zxy = xyz[:, 2] * 4 + xyz[:, 0] * 2 + xyz[:, 1]
    np.testing.assert_array_equal(zxy, t ^ (t >> 1))
    assert sorted(t) == list(range(8))


def test_curve_enters_and_leaves_along_z():
    """The curve starts in the cell at the origin and, its level-1 octant
    sequence ending in gray(7) = 100 read as (z, x, y), leaves through the
    corner cell on the z axis: at level 2 the first cell is (0,0,0) and
    the last (0,0,3)."""
    c = (np.array([[i, j, k] for i in range(4) for j in range(4)
                   for k in range(4)]) + 0.5) / 4.0
    keys = oracle.peano_keys(c[:, 0], c[:, 1], c[:, 2], [0, 0, 0], 1.0)
    order = np.lexsort((keys[:, 0], keys[:, 1]))
    cells = np.floor(c[order] * 4).astype(int)
    np.testing.assert_array_equal(cells[0], [0, 0, 0])
    np.testing.assert_array_equal(cells[-1], [0, 0, 3])
    # every step of the level-2 curve is a face step
    assert (np.abs(np.diff(cells, axis=0)).sum(1) == 1).all()
