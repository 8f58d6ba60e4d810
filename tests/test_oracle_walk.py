"""Pins the oracle's walks, force kernel and Ewald machinery
(Gravity/tree_accel.c, Gravity/tree_periodic.c, Gravity/periodic.c) with
closed forms and brute force:

  * theta_BH = 0 opens every node, so the BH walk must equal the brute-force
    direct sum over all other particles -- and, where a top node holds
    <= 64 particles, over the particle range [Target, Target+Npart) that the
    reference's interact_with_topnode_particles reads (tree_accel.c:183-203);
  * alpha_rel = 0 does the same for the relative walk;
  * two particles closer than |h| feel the K1 kernel with the negative
    softening length of Setup_Gravity_Tree (tree_build.c:136-139);
  * Ewald table: an independent numpy/scipy evaluation of the same finite
    sums at sampled nodes, odd symmetry, F(0) = 0, Fp(0) = 2.8372975;
  * Ewald lookup: exact table values at grid nodes, with the sign rule of
    Ewald_Correction (Gravity/periodic.c:35-113), and trilinear between.
"""
import numpy as np
import pytest
from scipy.special import erfc

import oracle

pytestmark = pytest.mark.cpu


def uniform_cube(n, seed):
    g = np.random.Generator(np.random.Philox(seed))
    return g.uniform(-1, 1, (n, 3)), g.uniform(0.5, 1.5, n) / n


def direct_numpy(x, m, G, eps, periodic=False, box=1.0, targets=None,
                 sources=None):
    """K1 softened forces (tree_accel.c:362-394), float64 numpy."""
    targets = np.arange(len(x)) if targets is None else targets
    out = np.zeros((len(targets), 3))
    eps2, eps3 = eps * eps, eps ** 3
    for t, i in enumerate(targets):
        src = np.arange(len(x)) if sources is None else sources[t]
        dr = x[src] - x[i]
        if periodic:
            dr = np.where(dr > box / 2, dr - box, dr)
            dr = np.where(dr < -box / 2, dr + box, dr)
        r2 = (dr * dr).sum(1)
        keep = r2 != 0
        dr, r2, mm = dr[keep], r2[keep], m[src][keep]
        fac = G * mm
        inner = r2 < eps2
        u2 = r2 / eps2
        fac = np.where(inner, fac * (175 - u2 * (294 - u2 * 135)) / (16 * eps3),
                       fac / np.sqrt(r2) ** 3)
        out[t] = (fac[:, None] * dr).sum(0)
    return out


def test_bh_theta0_is_direct_sum():
    x, m = uniform_cube(64 * 150, 1)
    s = oracle.Sim(x, m, softening=0.02, theta_bh=0.0, nthreads=8)
    s.setup_domain()
    s.domain_decomposition()
    assert (s.top["Npart"] > 64).all()
    s.tree_build()
    s.tree_acceleration(use_bh=True)
    xs = s.positions()
    want = direct_numpy(xs, s.mass, 1.0, s.epsilon)
    np.testing.assert_allclose(s.accelerations(), want, rtol=1e-10,
                               atol=1e-12 * np.abs(want).max())
    assert (s.cost == len(x) - 1).all()


def test_relative_alpha0_is_direct_sum():
    x, m = uniform_cube(64 * 120, 2)
    s = oracle.Sim(x, m, softening=0.02, alpha_rel=0.0, nthreads=8)
    s.setup_domain()
    s.domain_decomposition()
    s.tree_build()
    s.last_acc_mag[:] = 1.0
    s.tree_acceleration(use_bh=False)
    want = direct_numpy(s.positions(), s.mass, 1.0, s.epsilon)
    np.testing.assert_allclose(s.accelerations(), want, rtol=1e-10,
                               atol=1e-12 * np.abs(want).max())


def test_small_topnodes_read_target_range():
    # a dense core plus a sparse halo: outer top nodes hold <= 64 particles
    g = np.random.Generator(np.random.Philox(3))
    core = g.normal(0, 0.05, (6000, 3))
    halo = g.uniform(-1, 1, (500, 3))
    x = np.concatenate([core, halo])
    m = np.full(len(x), 1.0 / len(x))
    s = oracle.Sim(x, m, softening=0.01, theta_bh=0.0, nthreads=8)
    s.setup_domain()
    s.domain_decomposition()
    s.tree_build()
    top = s.top
    small = top[top["Npart"] <= 64]
    assert len(small) > 0 and (small["Target"] > 0).any()
    s.tree_acceleration(use_bh=True)
    xs = s.positions()
    # sources of every particle: real particles of big top nodes, the
    # [Target, Target+Npart) window for small ones
    src = []
    for t in top:
        f, n = int(t["First_Part"]), int(t["Npart"])
        if n > 64:
            src.append(np.arange(f, f + n))
        else:
            src.append(np.arange(int(t["Target"]), int(t["Target"]) + n))
    src = np.concatenate(src)
    tg = np.arange(0, len(x), 37)
    want = direct_numpy(xs, s.mass, 1.0, s.epsilon, targets=tg,
                        sources=[src] * len(tg))
    np.testing.assert_allclose(s.accelerations()[tg], want, rtol=1e-9,
                               atol=1e-12 * np.abs(want).max())


def test_k1_kernel_negative_length_inside_h():
    # a close pair inside |h| plus a featherweight uniform cloud that keeps
    # every top node above 64 particles
    soft = 0.1
    g = np.random.Generator(np.random.Philox(12))
    cloud = g.uniform(-1, 1, (64 * 160, 3))
    x = np.concatenate([[[0.01, 0.02, 0.0], [0.04, 0.03, -0.02]], cloud])
    m = np.concatenate([[0.3, 0.7], np.full(len(cloud), 1e-14)])
    s = oracle.Sim(x, m, softening=soft, theta_bh=0.0, nthreads=8)
    s.setup_domain()
    s.domain_decomposition()
    assert (s.top["Npart"] > 64).all()
    s.tree_build()
    s.tree_acceleration(use_bh=True)
    eps = -41.0 / 32.0 * soft
    assert s.epsilon == eps
    xs = s.positions()
    i0 = int(np.where((xs == x[0]).all(1))[0][0])
    i1 = int(np.where((xs == x[1]).all(1))[0][0])
    dr = xs[i1] - xs[i0]
    u2 = (dr @ dr) / eps ** 2
    assert u2 < 1
    a = 0.7 * (175 - 294 * u2 + 135 * u2 * u2) / (16 * eps ** 3) * dr
    assert np.dot(a, dr) < 0          # repulsive inside h
    np.testing.assert_allclose(s.accelerations()[i0], a, rtol=1e-9)


def _ewald_force_numpy(r):
    alpha = 2.0
    f = r / np.linalg.norm(r) ** 3
    n = np.array([(a, b, c) for a in range(-4, 5) for b in range(-4, 5)
                  for c in range(-4, 5)], dtype=float)
    dx = r[None, :] - n
    rr = np.linalg.norm(dx, axis=1)
    val = erfc(alpha * rr) + 2 * alpha * rr / np.sqrt(np.pi) * np.exp(
        -(alpha * rr) ** 2)
    f = f - (dx / rr[:, None] ** 3 * val[:, None]).sum(0)
    h = n[(n * n).sum(1) > 0]
    h2 = (h * h).sum(1)
    val = 2.0 / h2 * np.exp(-np.pi ** 2 * h2 / alpha ** 2) * np.sin(
        2 * np.pi * (h @ r))
    return f - (h * val[:, None]).sum(0)


def _ewald_pot_numpy(r):
    alpha = 2.0
    n = np.array([(a, b, c) for a in range(-4, 5) for b in range(-4, 5)
                  for c in range(-4, 5)], dtype=float)
    rr = np.linalg.norm(r[None, :] - n, axis=1)
    s1 = (erfc(alpha * rr) / rr).sum()
    h = n[(n * n).sum(1) > 0]
    h2 = (h * h).sum(1)
    s2 = (1.0 / (np.pi * h2) * np.exp(-np.pi ** 2 * h2 / alpha ** 2)
          * np.cos(2 * np.pi * (h @ r))).sum()
    return np.pi / alpha ** 2 - s1 - s2 + 1 / np.linalg.norm(r)


@pytest.fixture(scope="module")
def table():
    return oracle.cached_ewald_table()


def test_ewald_table_independent_sums(table):
    g = np.random.Generator(np.random.Philox(9))
    idx = g.integers(0, 65, (60, 3))
    for i, j, k in idx:
        if i == j == k == 0:
            continue
        r = np.array([i, j, k]) * 0.5 / 64
        np.testing.assert_allclose(table[:3, i, j, k], _ewald_force_numpy(r),
                                   rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(table[3, i, j, k], _ewald_pot_numpy(r),
                                   rtol=1e-11, atol=1e-12)
    assert (table[:3, 0, 0, 0] == 0).all()
    assert table[3, 0, 0, 0] == 2.8372975
    # the correction force vanishes on the faces it is odd across
    assert np.abs(table[0, 0, :, :]).max() < 1e-12
    assert np.abs(table[1, :, 0, :]).max() < 1e-12
    # the table is (nearest-image force) - (periodic force); at x = 1/2 the
    # periodic x force vanishes by symmetry, leaving x / r^3
    jj, kk = np.meshgrid(np.arange(65), np.arange(65), indexing="ij")
    r = np.sqrt(0.25 + (jj / 128.0) ** 2 + (kk / 128.0) ** 2)
    np.testing.assert_allclose(table[0, 64], 0.5 / r ** 3, rtol=1e-9)
    # axis symmetry: Fx(i,j,k) = Fy(j,i,k)
    np.testing.assert_allclose(table[0], np.transpose(table[1], (1, 0, 2)),
                               rtol=1e-10, atol=1e-13)


def test_ewald_lookup_nodes_and_sign(table):
    box = 2.0
    x = np.array([[0.1, 0.2, 0.3]] * 2)
    s = oracle.Sim(x, np.ones(2), periodic=True, boxsize=box,
                   ewald_table=table, nthreads=1)
    for (i, j, k) in [(1, 2, 3), (10, 40, 63), (64, 0, 5), (33, 33, 33)]:
        dr = np.array([i, j, k]) * box / 128.0
        f = s.ewald_correction(dr)
        want = -table[:3, i, j, k] / box ** 2
        if i == 64:      # the index is clamped to 63 with weight u = 1
            want[0] = -table[0, 64, j, k] / box ** 2
        np.testing.assert_allclose(f, want, rtol=1e-13, atol=1e-18)
        fm = s.ewald_correction(-dr)
        # odd in each component; a zero component keeps the sign of +0
        want_m = np.where(dr != 0, -f, f)
        np.testing.assert_array_equal(fm, want_m)
    # trilinear between nodes
    dr = np.array([10.25, 20.5, 30.75]) * box / 128.0
    u = np.array([0.25, 0.5, 0.75])
    acc = np.zeros(3)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = ((u[0] if a else 1 - u[0]) * (u[1] if b else 1 - u[1])
                     * (u[2] if c else 1 - u[2]))
                acc += w * table[:3, 10 + a, 20 + b, 30 + c]
    np.testing.assert_allclose(s.ewald_correction(dr), -acc / box ** 2,
                               rtol=1e-12)
