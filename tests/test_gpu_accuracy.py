"""GPU accuracy: the FP64 direct-sum oracle on the GPU (tdv_forcetest) against
the CPU oracle's direct sum, then the FP32 group walk's error against it.

The FP32 walk does not reproduce the reference's interaction sets (it opens
a node when any particle of its 64-particle group asks for it), so it is
held to an error bound, written in each test, against FP64 direct
summation with the same pairwise kernel and the same Ewald table.
"""
import numpy as np
import pytest

import oracle
from tandav_amd import Engine, ics

pytestmark = pytest.mark.gpu


def rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


@pytest.fixture(scope="module")
def table():
    return oracle.cached_ewald_table()


@pytest.mark.parametrize("periodic", [False, True])
def test_forcetest_matches_cpu_direct_sum(periodic, table):
    if periodic:
        pos, _, m = ics.zeldovich_box(24, seed=7)
        soft = 1 / 960.
    else:
        pos, _, m = ics.hernquist_sphere(20000, seed=1)
        soft = 0.01
    e = Engine(periodic=periodic, softening=soft, ewald_table=table
               if periodic else None)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    s = oracle.Sim(pos, m, periodic=periodic, softening=soft,
                   ewald_table=table if periodic else None, nthreads=16)
    s.setup_domain()
    s.domain_decomposition()
    np.testing.assert_array_equal(e.get("pos")["pos"], s.positions())
    idx = np.arange(0, len(m), 97)
    got = e.forcetest(idx)
    want = s.direct_sum(idx)
    # same terms, different summation order (block reduction vs sequential)
    assert rel_err(got, want).max() < 1e-11


def _fp32_error(pos, m, periodic, soft, alpha, esize, table, order=2):
    e64 = Engine(periodic=periodic, softening=soft, alpha_rel=alpha,
                 ewald_table=table if periodic else None)
    e32 = Engine(periodic=periodic, softening=soft, alpha_rel=alpha,
                 precision="fp32", ewald_size=esize, compat=False,
                 ewald_table=table if periodic else None, fp32_order=order)
    out = []
    for e in (e64, e32):
        e.set_particles(pos, m)
        e.setup_domain()
        e.domain_decomposition()
        e.compute_acceleration(True)
        e.domain_decomposition()
        e.compute_acceleration(False)
        out.append(e.get("acc")["acc"])
    g = np.random.Generator(np.random.Philox(5))
    idx = np.sort(g.choice(len(m), 2048, replace=False))
    ref = e64.forcetest(idx)
    return rel_err(out[1][idx], ref), rel_err(out[0][idx], ref)


def test_fp32_group_walk_hernquist(table):
    pos, _, m = ics.hernquist_sphere(65536, seed=1)
    r32, r64 = _fp32_error(pos, m, False, 0.01, 0.005, 0.0625, table)
    # p99 |da|/|a| <= 2e-3 (the accuracy bar of the FP32 path)
    assert np.percentile(r32, 99) <= 2e-3
    assert np.median(r32) <= np.median(r64)


def test_fp32_group_walk_periodic_box(table):
    pos, _, m = ics.zeldovich_box(48, seed=7)
    r32, r64 = _fp32_error(pos, m, True, 1 / 1920., 0.005, 1 / 64., table)
    # p99 |da|/|a| <= 2e-3 with per-particle corrections in bundles; the
    # reference's own FP64 walk is far less accurate on this box
    assert np.percentile(r32, 99) <= 2e-3
    assert np.percentile(r32, 99) < np.percentile(r64, 99)


@pytest.mark.parametrize("periodic", [False, True])
def test_fp32_quadrupoles_reduce_the_error(periodic, table):
    """The quadrupole term of accepted nodes (fp32_order=2) against the
    monopole-only walk (fp32_order=1) at the same opening criterion."""
    if periodic:
        pos, _, m = ics.zeldovich_box(48, seed=7)
        args = (True, 1 / 1920., 0.005, 1 / 64., table)
    else:
        pos, _, m = ics.hernquist_sphere(65536, seed=1)
        args = (False, 0.01, 0.005, 0.0625, table)
    r2, _ = _fp32_error(pos, m, *args, order=2)
    r1, _ = _fp32_error(pos, m, *args, order=1)
    assert np.percentile(r2, 99) < 0.7 * np.percentile(r1, 99)
    assert np.median(r2) < 0.7 * np.median(r1)


def test_ewald_second_derivatives(table):
    """The walk's second-order Ewald table against second differences of the
    oracle's FP64 correction table (grid step 1/128 box), and the symmetry
    and tracelessness the correction's curvature must have."""
    e = Engine(periodic=True, softening=1e-3)
    S = e.ewald_derivs()
    E = -table[:3]                       # the correction on the octant
    h = 0.5 / 64
    sl = (slice(2, 62),) * 3             # interior nodes, away from the faces
    comp = {"xxx": (0, 0), "yyy": (1, 1), "zzz": (2, 2), "xxy": (1, 0),
            "xxz": (2, 0), "xyy": (0, 1), "yyz": (2, 1), "xzz": (0, 2),
            "yzz": (1, 2)}                # name: (E component, axis)
    names = ["xxx", "yyy", "zzz", "xxy", "xxz", "xyy", "yyz", "xzz", "yzz",
             "xyz"]
    scale = np.abs(S[:, 2:62, 2:62, 2:62]).max()
    for q, name in enumerate(names[:9]):
        c, ax = comp[name]
        fd = (np.roll(E[c], -1, ax) - 2 * E[c] + np.roll(E[c], 1, ax)) / h**2
        # O(h^2) differences of the table against the fine-step derivative
        assert np.abs(fd[sl] - S[q][sl]).max() < 0.02 * scale, name
    fd = (np.roll(np.roll(E[0], -1, 1), -1, 2) - np.roll(np.roll(E[0], -1, 1),
          1, 2) - np.roll(np.roll(E[0], 1, 1), -1, 2)
          + np.roll(np.roll(E[0], 1, 1), 1, 2)) / (4 * h * h)
    assert np.abs(fd[sl] - S[9][sl]).max() < 0.02 * scale
    # traceless in the last two indices: S_axx + S_ayy + S_azz = 0
    for a, (i1, i2, i3) in enumerate([(0, 5, 7), (3, 1, 8), (4, 6, 2)]):
        tr = S[i1] + S[i2] + S[i3]
        assert np.abs(tr).max() < 1e-4 * scale, a
    # odd in the axes that appear an odd number of times: zero on those faces
    assert np.abs(S[0][0]).max() < 1e-9 * scale       # xxx at x = 0
    assert np.abs(S[9][:, 0]).max() < 1e-9 * scale    # xyz at y = 0
    # E''(d) grows from zero at the origin (cubic lattice)
    assert np.abs(S[:, 0, 0, 0]).max() < 1e-6 * scale


def test_ewald_first_derivatives(table):
    """The first derivatives J_ab = d_b E_a the FP32 walk expands the Ewald
    correction with: against central differences of the oracle's FP64
    table, symmetric by construction, and with the divergence the closed
    form fixes: away from the particle the periodic force of a unit mass in
    a unit box has div = 4 pi (the uniform background), the nearest-image
    Newton force none, so div E = -4 pi for E = -(that difference) at
    every node, origin and faces included (to the truncation of the
    reference's sums, alpha 2 with |n|, |h| <= 4: ~7e-5 of 4 pi)."""
    e = Engine(periodic=True, softening=1e-3)
    J = e.ewald_jacobian()
    E = -table[:3]
    h = 0.5 / 64
    sl = (slice(2, 62),) * 3
    comp = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]   # (E comp, axis)
    scale = np.abs(J[:, 2:62, 2:62, 2:62]).max()
    for q, (c, ax) in enumerate(comp):
        fd = (np.roll(E[c], -1, ax) - np.roll(E[c], 1, ax)) / (2 * h)
        assert np.abs(fd[sl] - J[q][sl]).max() < 2e-3 * scale, q
    tr = J[0] + J[1] + J[2]
    np.testing.assert_allclose(tr, -4 * np.pi, rtol=0, atol=2e-4 * 4 * np.pi)
    # E is odd in x: J_xy vanishes on the x = 0 face, J_xx does not
    assert np.abs(J[3][0]).max() < 1e-9 * scale


def test_forcetest_exact_on_table_nodes(table):
    """The exact-Ewald direct sum against the oracle's table: separations
    placed on table nodes, where trilinear interpolation returns the node
    value, so both sums must agree to rounding."""
    g = np.random.Generator(np.random.Philox(12))
    h = 1.0 / 128
    base = np.array([0.3, 0.4, 0.45])
    steps = g.integers(-60, 61, (40, 3))
    steps = steps[np.abs(steps).sum(1) > 0]
    pos = np.vstack([base, np.mod(base + steps * h, 1.0)])
    m = np.full(len(pos), 1.0 / len(pos))
    e = Engine(periodic=True, softening=1e-4, ewald_table=table)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    ids = e.get("id")["id"]
    i0 = int(np.flatnonzero(ids == 0)[0])          # the base particle
    want = e.forcetest([i0])
    for pairs in (True, False):
        got = e.forcetest_exact([i0], pairs=pairs)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("box", ["zeldovich", "clustered"])
def test_exact_split_matches_pair_sums(box):
    """The exact force split at screening 4 (real-space images per pair,
    Fourier part through structure factors) against the reference's own
    sums (screening 2, |n|, |h| <= 4) evaluated per pair: the same lattice
    sum, so they agree to rounding on every sampled particle, near
    neighbours and softened pairs included."""
    if box == "zeldovich":
        pos, _, m = ics.zeldovich_box(24, seed=7)
        soft = 1 / 960.
    else:
        pos, _, m = ics.clustered_box(16384, nhalo=8, seed=3)
        soft = 1 / 1000.
    e = Engine(periodic=True, softening=soft)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    g = np.random.Generator(np.random.Philox(8))
    idx = np.sort(g.choice(len(m), 48, replace=False))
    pairs = e.forcetest_exact(idx, pairs=True)
    split = e.forcetest_exact(idx)
    assert rel_err(split, pairs).max() < 1e-11


def test_fp32_walk_against_exact_ewald():
    """The FP32 group walk at the benchmark's settings (alpha 0.01 with the
    next-order criterion, quadrupoles, second-order Ewald) against FP64
    direct summation with exact Ewald sums: p99 |da|/|a| <= 2e-3, and far
    below the FP64 walk's error on the same box."""
    pos, _, m = ics.zeldovich_box(48, seed=7)
    soft = 1 / 1920.
    e64 = Engine(periodic=True, softening=soft, alpha_rel=0.005)
    e32 = Engine(periodic=True, softening=soft, alpha_rel=0.01,
                 precision="fp32", compat=False, ewald_size=0.125,
                 ewald_accuracy=1.0, fp32_order=2, fp32_criterion=3)
    out = []
    for e in (e64, e32):
        e.set_particles(pos, m)
        e.setup_domain()
        e.domain_decomposition()
        e.compute_acceleration(True)
        e.domain_decomposition()
        e.compute_acceleration(False)
        out.append(e.get("acc")["acc"])
    g = np.random.Generator(np.random.Philox(5))
    idx = np.sort(g.choice(len(m), 256, replace=False))
    ref = e64.forcetest_exact(idx)
    r32, r64 = rel_err(out[1][idx], ref), rel_err(out[0][idx], ref)
    assert np.percentile(r32, 99) <= 2e-3
    assert np.percentile(r32, 99) < 0.5 * np.percentile(r64, 99)


def test_fp32_bundles_taken_whole():
    """Bundles (collapsed leaves) that no lane's criterion opens are taken
    whole at their centre of mass (fp32_bundle_min, default 4): far fewer
    interactions than taking every bundle particle by particle (-1, the
    reference's walk), with the error against the exact periodic force
    still within the bar."""
    pos, _, m = ics.zeldovich_box(48, seed=7)
    soft = 1 / 1920.
    out = {}
    for bmin in (-1, 0, 64):
        e = Engine(periodic=True, softening=soft, alpha_rel=0.01,
                   precision="fp32", compat=False, ewald_size=0.125,
                   ewald_accuracy=1.0, fp32_order=2, fp32_criterion=3,
                   fp32_bundle_min=bmin)
        e.set_particles(pos, m)
        e.setup_domain()
        e.domain_decomposition()
        e.compute_acceleration(True)
        e.domain_decomposition()
        e.compute_acceleration(False)
        out[bmin] = (e.get("acc")["acc"], e.interaction_counts()[0])
        if bmin == -1:
            g = np.random.Generator(np.random.Philox(5))
            idx = np.sort(g.choice(len(m), 256, replace=False))
            ref = e.forcetest_exact(idx)
        e.close()
    assert out[0][1] < 0.5 * out[-1][1]
    assert out[0][1] <= out[64][1] <= out[-1][1]
    for bmin in (-1, 0, 64):
        err = rel_err(out[bmin][0][idx], ref)
        assert np.percentile(err, 99) <= 2e-3, bmin
    from tandav_amd import TDVError
    with pytest.raises(TDVError):
        Engine(precision="fp32", compat=False, fp32_bundle_min=65)
