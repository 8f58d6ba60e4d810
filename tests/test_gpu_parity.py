"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on
the same seeded inputs.  Integer and index results (keys, permutations,
top nodes, the node array's links and counts, Tree_Parent, Cost) must be
bit-exact; node masses, centres of mass and centres must be bit-exact too
(same FP64 operations in the same order); accelerations are held to
|delta a| / |a| <= 1e-12 per particle, far inside the 1e-6 drop-in bar.
"""
import numpy as np
import pytest

import oracle
from tandav_amd import Engine, TDVError, ics

pytestmark = pytest.mark.gpu

ACC_RTOL = 1e-12


@pytest.fixture(scope="module")
def ewald():
    return oracle.cached_ewald_table()


def _engine(**kw):
    return Engine(**kw)


def test_keys_bit_exact():
    g = np.random.Generator(np.random.Philox(21))
    x = g.uniform(-3.0, 5.0, (200000, 3))
    origin = np.array([-3.0, -3.0, -3.0])
    size = 8.0001
    e = _engine()
    for rev in (False, True):
        got = e.peano_keys(x[:, 0], x[:, 1], x[:, 2], origin, size, rev)
        want = oracle.peano_keys(x[:, 0], x[:, 1], x[:, 2], origin, size, rev)
        np.testing.assert_array_equal(got, want)
    # grid corners and faces, the extremes of the quantiser
    i = np.arange(17) / 16.0
    gx, gy, gz = np.meshgrid(i, i, i, indexing="ij")
    p = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) * 0.9999999999
    for rev in (False, True):
        np.testing.assert_array_equal(
            e.peano_keys(p[:, 0], p[:, 1], p[:, 2], [0, 0, 0], 1.0, rev),
            oracle.peano_keys(p[:, 0], p[:, 1], p[:, 2], [0, 0, 0], 1.0, rev))


@pytest.mark.parametrize("n", [1, 2, 3, 4095, 4096, 4097, 300001])
def test_sort_matches_index_quicksort(n):
    g = np.random.Generator(np.random.Philox(22 + n))
    k = g.integers(0, 2 ** 63, (n, 2), dtype=np.uint64)
    k[: n // 2, 1] = k[0, 1]              # long runs of equal high words
    k[: n // 2, 1] |= np.uint64(1 << 63)
    e = _engine()
    perm = e.sort_keys(k)
    want = oracle.qsort_index(k, nthreads=4)
    np.testing.assert_array_equal(perm.astype(np.int64), want)


def test_sort_rejects_duplicate_keys():
    k = np.array([[5, 1], [3, 1], [5, 1]], dtype=np.uint64)
    e = _engine()
    with pytest.raises(TDVError, match="share"):
        e.sort_keys(k)


def test_ewald_table_and_lookup(ewald):
    e = _engine(periodic=True, boxsize=2.0)
    t = e.ewald_table()
    # device erfc/exp/sincos vs glibc differ by an ulp here and there; the
    # entries next to the origin (r = 1/128) are x/r^3 ~ 1.6e4 minus an
    # n = 0 term of the same size, so one ulp of 1.6e4 (3.6e-12) survives
    # the cancellation as an absolute difference
    np.testing.assert_allclose(t, ewald, rtol=1e-12, atol=1e-11)
    e.set_ewald_table(ewald)
    g = np.random.Generator(np.random.Philox(23))
    dr = g.uniform(-1.0, 1.0, (50000, 3))
    dr[:10, 1] = 0.0
    got = e.ewald_correction(dr)
    s = oracle.Sim(np.zeros((2, 3)) + 0.5, np.ones(2), periodic=True,
                   boxsize=2.0, ewald_table=ewald, nthreads=1)
    want = np.array([s.ewald_correction(d) for d in dr[:2000]])
    np.testing.assert_array_equal(got[:2000], want)
    # the per-element form on the host: the same code, the same table
    np.testing.assert_array_equal(e.ewald_correction_host(dr), got)


def test_host_keys_match_device_keys():
    """tdv_peano_keys_host (the binding's per-element Peano_Key) computes on
    the host what the device computes, bit for bit."""
    g = np.random.Generator(np.random.Philox(24))
    x = g.uniform(-3.0, 5.0, (100000, 3))
    e = _engine()
    for rev in (False, True):
        np.testing.assert_array_equal(
            e.peano_keys_host(x[:, 0], x[:, 1], x[:, 2], [-3.0] * 3, 8.0001,
                              rev),
            e.peano_keys(x[:, 0], x[:, 1], x[:, 2], [-3.0] * 3, 8.0001, rev))


def _pipeline_pair(pos, mass, periodic, ewald, prepare=True, **kw):
    """Run the reference calling sequence on the oracle and on the GPU."""
    box = kw.pop("boxsize", 1.0)
    s = oracle.Sim(pos, mass, periodic=periodic, boxsize=box,
                   ewald_table=ewald if periodic else None, **kw)
    s.setup_domain()
    s.domain_decomposition()
    s.compute_acceleration(prepare)
    e = _engine(periodic=periodic, boxsize=box, compat=True,
                **{k: v for k, v in kw.items() if k != "nthreads"})
    if periodic:
        e.set_ewald_table(ewald)
    e.set_particles(pos, mass)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(prepare)
    return s, e


def _compare(s, e):
    size, origin, center = e.domain
    os_, oo, oc = s.domain
    assert size == os_
    np.testing.assert_array_equal(origin, oo)
    got = e.get()
    np.testing.assert_array_equal(got["pos"], s.positions())
    np.testing.assert_array_equal(got["id"], s.id)
    np.testing.assert_array_equal(got["key"], s.key)
    gt = e.top_nodes()
    st = s.top
    for f in ("Key", "Target", "Level", "First_Part", "Npart", "Pos", "Mass",
              "CoM"):
        np.testing.assert_array_equal(gt[f], st[f], err_msg=f)
    tree_g = e.tree()
    tree_s = s.tree
    assert len(tree_g) == len(tree_s)
    for f in ("DNext", "Bitfield", "DUp", "Npart", "Pos", "Mass", "CoM"):
        np.testing.assert_array_equal(tree_g[f], tree_s[f], err_msg=f)
    np.testing.assert_array_equal(got["tree_parent"], s.tree_parent)
    np.testing.assert_array_equal(got["cost"], s.cost)
    a_g, a_s = got["acc"], s.accelerations()
    rel = np.linalg.norm(a_g - a_s, axis=1) / np.linalg.norm(a_s, axis=1)
    assert rel.max() <= ACC_RTOL, rel.max()
    np.testing.assert_allclose(got["last_acc_mag"], s.last_acc_mag,
                               rtol=ACC_RTOL)
    return rel


def test_hernquist_small_prepare_step():
    pos, vel, m = ics.hernquist_sphere(16384, seed=1)
    s, e = _pipeline_pair(pos, m, False, None, softening=0.01,
                          alpha_rel=0.005, nthreads=8)
    _compare(s, e)


def test_clustered_blobs():
    from test_oracle_tree import clustered
    pos, m = clustered(30000, 5)
    s, e = _pipeline_pair(pos, m, False, None, softening=1e-3,
                          alpha_rel=0.005, nthreads=8)
    _compare(s, e)


def test_periodic_zeldovich_32(ewald):
    pos, vel, m = ics.zeldovich_box(32, seed=7)
    s, e = _pipeline_pair(pos, m, True, ewald, softening=1 / 1280.,
                          alpha_rel=0.005, nthreads=8)
    _compare(s, e)


def test_config1_hernquist_65536():
    """C1: 65,536 Hernquist particles, eps 0.01, alpha 0.005, one full
    force evaluation (keys, sort, build, BH + relative walk)."""
    pos, vel, m = ics.hernquist_sphere(65536, seed=1)
    s, e = _pipeline_pair(pos, m, False, None, softening=0.01,
                          alpha_rel=0.005, nthreads=16)
    _compare(s, e)


@pytest.mark.parametrize("case", ["clustered", "periodic"])
def test_refined_top_nodes(case, ewald):
    """Top nodes at mixed levels, as the reference's decomposition makes
    with NTarget > 1, handed to both sides (tdv_set_top_nodes)."""
    from test_oracle_tree import clustered
    from topnodes import refined_top_nodes
    if case == "periodic":
        pos, _, m = ics.zeldovich_box(32, seed=7)
        kw = dict(periodic=True, softening=1 / 1280.)
    else:
        pos, m = clustered(30000, 5)
        kw = dict(periodic=False, softening=1e-3)
    table = ewald if kw["periodic"] else None
    s = oracle.Sim(pos, m, alpha_rel=0.005, nthreads=16, ewald_table=table,
                   **kw)
    s.setup_domain()
    s.domain_decomposition()
    size, origin, _ = s.domain
    x = s.positions()
    rec = refined_top_nodes(oracle.peano_keys(x[:, 0], x[:, 1], x[:, 2],
                                              origin, size), nsplit=10)
    s.set_top_nodes(rec)
    s.compute_acceleration(True)
    e = _engine(alpha_rel=0.005, compat=True, **kw)
    if kw["periodic"]:
        e.set_ewald_table(ewald)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.set_top_nodes(s.top)
    e.compute_acceleration(True)
    _compare(s, e)


def test_set_domain_and_sort_particles():
    """Sort_Particles_By_Peano_Key with the caller's Domain."""
    pos, _, m = ics.hernquist_sphere(50000, seed=3)
    c, size, o = oracle.domain_nonperiodic(pos[:, 0], pos[:, 1], pos[:, 2])
    e = _engine()
    e.set_particles(pos, m)
    e.set_domain(size * 1.5, o - 0.25 * size)
    e.sort_particles()
    got = e.get("pos", "id")
    keys = oracle.peano_keys(pos[:, 0], pos[:, 1], pos[:, 2],
                             o - 0.25 * size, size * 1.5)
    perm = oracle.qsort_index(keys, nthreads=4)
    np.testing.assert_array_equal(got["id"], perm.astype(np.uint32))
    np.testing.assert_array_equal(got["pos"], pos[perm])
