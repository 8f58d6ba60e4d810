"""The reference-side binding on the GPU: a host program calling Tandav's
own functions (Sort_Particles_By_Peano_Key, Reverse_Peano_Keys,
Compute_Acceleration, Gravity_Tree_Update_*, Peano_Key, Ewald_Correction,
Gravity_Forcetest) from a 4-thread OpenMP team leaves exactly the state the
engine's own calling sequence produces, and the engine's own sequence is
pinned to the oracle by test_gpu_parity.py."""
import os
import sys

import numpy as np
import pytest

import oracle
from tandav_amd import Engine, ics

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "glue"))
import host as glue_host  # noqa: E402

pytestmark = pytest.mark.gpu


def _engine_step(e, prepare):
    e.domain_decomposition()
    top = e.top_nodes()
    e.compute_acceleration(prepare)
    return top


def _compare(h, e):
    got = e.get("pos", "vel", "mass", "acc", "grav_acc", "cost",
                "last_acc_mag", "tree_parent", "key", "id", "time_bin")
    np.testing.assert_array_equal(h.vec("Pos"), got["pos"])
    np.testing.assert_array_equal(h.vec("Vel"), got["vel"])
    np.testing.assert_array_equal(h.column("Mass"), got["mass"])
    np.testing.assert_array_equal(h.column("ID"), got["id"])
    np.testing.assert_array_equal(h.column("Time_Bin"), got["time_bin"])
    np.testing.assert_array_equal(h.column("Key"), got["key"])
    # host-only columns followed the same permutation
    ids = h.column("ID").astype(np.uint64)
    np.testing.assert_array_equal(h.column("It_Drift_Pos"),
                                  (3 * ids).astype(np.uint32))
    np.testing.assert_array_equal(h.column("It_Kick_Pos"),
                                  (5 * ids).astype(np.uint32))
    for f in ("acc", "grav_acc"):
        np.testing.assert_array_equal(
            h.vec({"acc": "Acc", "grav_acc": "Grav_Acc"}[f]), got[f])
    np.testing.assert_array_equal(h.column("Cost"), got["cost"])
    np.testing.assert_array_equal(h.column("Last_Acc_Mag"),
                                  got["last_acc_mag"])
    np.testing.assert_array_equal(h.column("Tree_Parent"),
                                  got["tree_parent"])
    tree, etree = h.tree(), e.tree()
    assert len(tree) == len(etree)
    for f in etree.dtype.names:
        np.testing.assert_array_equal(tree[f], etree[f], err_msg=f)
    top, etop = h.top_nodes(), e.top_nodes()
    for f in ("Target", "Mass", "CoM", "First_Part", "Npart", "Level"):
        np.testing.assert_array_equal(top[f], etop[f], err_msg=f)


@pytest.mark.parametrize("periodic", [False, True])
def test_binding_step_matches_engine(periodic):
    # no table handed in: the engine computes it on the GPU, as the
    # binding's context does (Gravity_Periodic_Init)
    if periodic:
        pos, vel, m = ics.zeldovich_box(24, seed=7)
        soft, box = 1 / 960., 1.0
    else:
        pos, vel, m = ics.hernquist_sphere(20000, seed=1)
        soft, box = 0.01, 1.0
    alpha = 0.005
    e = Engine(periodic=periodic, softening=soft, alpha_rel=alpha,
               theta_bh=0.1)
    e.set_particles(pos, m, vel=vel)
    e.setup_domain()
    h = glue_host.Host(periodic)
    h.init(pos, vel, m, boxsize=box, softening=soft, alpha_rel=alpha,
           theta_bh=0.1, threads=4)

    # prepare step: decomposition + BH walk + relative walk
    top = _engine_step(e, True)
    size, origin, _ = e.domain
    h.decomposition(size, origin, top)
    h.L.host_compute_acceleration(1)
    _compare(h, e)

    # a regular step after a new decomposition
    top = _engine_step(e, False)
    h.decomposition(size, origin, top)
    h.L.host_compute_acceleration(0)
    _compare(h, e)

    # a step without a rebuild: tree kicks and drift on the host mirror
    dt = 1e-3
    tree0 = h.tree()
    acc = h.vec("Acc")
    mass = h.column("Mass")
    parent = h.column("Tree_Parent")
    h.L.host_tree_kicks(dt)
    h.L.host_drift(dt)
    dp = np.zeros((len(tree0), 3))
    for i in range(len(mass)):          # Gravity/tree_update.c:12-57
        node = parent[i]
        while node >= 0 and not (tree0["Bitfield"][node] >> 10) & 1:
            dp[node] += mass[i] * dt * acc[i] / tree0["Mass"][node]
            node -= tree0["DUp"][node]
    tree1 = h.tree()
    np.testing.assert_allclose(tree1["CoM"], tree0["CoM"] + dt * dp,
                               rtol=1e-12, atol=1e-15)
    assert not ((tree1["Bitfield"] >> 11) & 1).any()
    h.L.host_compute_acceleration(0)
    pos1 = h.vec("Pos")
    e.update_particles(pos1)
    e.set_tree_moments(tree1, h.top_nodes())
    e.compute_acceleration(False)
    got = e.get("acc", "cost")
    np.testing.assert_array_equal(h.vec("Acc"), got["acc"])
    np.testing.assert_array_equal(h.column("Cost"), got["cost"])
    assert (h.column("Cost") > 0).all()


def test_binding_active_subset_and_tree_walk():
    """Active_Particle_List from Time_Bin (timestep.c:280-299) reaches the
    walk; Gravity_Tree_Acceleration alone with the BH criterion."""
    pos, vel, m = ics.hernquist_sphere(12000, seed=3)
    g = np.random.Generator(np.random.Philox(8))
    tb = g.integers(0, 3, len(m)).astype(np.int32)
    e = Engine(periodic=False, softening=0.01, alpha_rel=0.005)
    e.set_particles(pos, m, vel=vel, time_bin=tb)
    e.setup_domain()
    top = _engine_step(e, True)          # all active on the prepare step
    h = glue_host.Host(False)
    h.init(pos, vel, m, softening=0.01, alpha_rel=0.005, time_bin=tb)
    h.L.host_set_max_active_bin(10)       # a sync point: every bin active
    size, origin, _ = e.domain
    h.decomposition(size, origin, top)
    h.L.host_compute_acceleration(1)
    _compare(h, e)

    h.L.host_set_max_active_bin(0)
    h.L.host_make_active()                # Make_Active_Particle_List
    nact = h.L.host_nactive()
    tbh = h.column("Time_Bin")
    assert nact == int((tbh == 0).sum()) and 0 < nact < len(m)
    h.L.host_tree_acceleration(1)
    act = np.flatnonzero(tbh == 0)
    e.set_active(act)
    acc0 = e.get("acc")["acc"]
    acc0[act] = 0
    e.update_particles(h.vec("Pos"), acc=acc0)
    e.tree_acceleration(True)
    got = e.get("acc", "cost")
    np.testing.assert_array_equal(h.vec("Acc"), got["acc"])
    np.testing.assert_array_equal(h.column("Cost")[act], got["cost"][act])


def test_binding_element_functions():
    """Peano_Key / Reversed / Short keys and Ewald_Correction called from
    many host threads at once."""
    table = oracle.cached_ewald_table()
    h = glue_host.Host(True)
    pos, vel, m = ics.zeldovich_box(8, seed=2)
    h.init(pos, vel, m, softening=1 / 320.)
    h.set_domain(1.0, np.zeros(3))
    g = np.random.Generator(np.random.Philox(6))
    x, y, z = g.uniform(0, 1, (3, 300))
    std, rev, sh, rsh = h.peano_keys(x, y, z)
    np.testing.assert_array_equal(std, oracle.peano_keys(x, y, z,
                                                         np.zeros(3), 1.0))
    np.testing.assert_array_equal(rev, oracle.peano_keys(
        x, y, z, np.zeros(3), 1.0, reversed_=True))
    np.testing.assert_array_equal(sh, std[:, 1])
    np.testing.assert_array_equal(rsh, rev[:, 0] & np.uint64(0xFFFFFFFF))
    dr = g.uniform(-0.5, 0.5, (300, 3))
    s = oracle.Sim(pos, m, periodic=True, ewald_table=table)
    want = np.array([s.ewald_correction(d) for d in dr])
    np.testing.assert_allclose(h.ewald_correction(dr), want, rtol=0,
                               atol=1e-11)


def test_binding_momentum_check_output(tmp_path, capfd):
    """check_total_momentum (Gravity/tree_accel.c:442-469) brackets each
    walk: its running sums are never reset, so a fresh program prints the
    relative changes 1 and 1/3 on the prepare step (BH walk, then relative
    walk) and 1/5 on the next step, whatever the velocities."""
    import shutil
    fresh = tmp_path / "libtandav_host_fresh.so"
    shutil.copy(glue_host.build_host.lib_path(False), fresh)
    pos, vel, m = ics.hernquist_sphere(4096, seed=5)
    e = Engine(periodic=False, softening=0.01, alpha_rel=0.005)
    e.set_particles(pos, m, vel=vel)
    e.setup_domain()
    top = _engine_step(e, True)
    h = glue_host.Host(False, path=str(fresh))
    h.init(pos, vel, m, softening=0.01, alpha_rel=0.005)
    h.L.host_set_max_active_bin(10)
    h.L.host_set_verbose(1)
    size, origin, _ = e.domain
    h.decomposition(size, origin, top)
    capfd.readouterr()
    h.L.host_compute_acceleration(1)
    h.L.host_compute_acceleration(0)
    import ctypes
    ctypes.CDLL(None).fflush(None)
    out = capfd.readouterr().out
    h.L.host_set_verbose(0)
    vals = [float(l.split(":")[1]) for l in out.splitlines()
            if l.startswith("Total change in momentum due to gravity:")]
    assert vals == pytest.approx([1.0, 1 / 3, 1 / 5], rel=1e-5), out
    assert out.count("Tree acceleration  done") == 3, out


@pytest.mark.parametrize("periodic", [False, True])
def test_binding_fp32_walk_accuracy(periodic):
    """TANDAV_WALK=fp32 in a fresh program (the binding reads it when it
    creates its context): Tandav's calling sequence through the binding
    with the FP32 group walk at the binding's settings holds p99
    |da|/|a| <= 2e-3 against FP64 direct summation (exact periodic force)."""
    import json
    import subprocess
    env = dict(os.environ, TANDAV_WALK="fp32")
    for k in ("TANDAV_FP32_ALPHA", "TANDAV_FP32_CRITERION",
              "TANDAV_FP32_EWALD_ACCURACY"):
        env.pop(k, None)
    script = os.path.join(os.path.dirname(__file__), "glue", "fp32_step.py")
    r = subprocess.run([sys.executable, script, "1" if periodic else "0"],
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res)
    assert res["p99"] <= 2e-3
    assert res["cost_min"] > 0


def test_binding_ewald_table_file(tmp_path, monkeypatch):
    """Gravity_Periodic_Init (Gravity/periodic.c:180-233, 303-342): without
    ./ewald_tables.dat it writes the table there (Fx, Fy, Fz, Fp, 65^3
    doubles each, box units, raw); with the file it takes the table from
    it.  A file holding twice the table doubles every Ewald_Correction."""
    import shutil
    monkeypatch.chdir(tmp_path)
    pos, vel, m = ics.zeldovich_box(8, seed=2)
    libs = []
    for k in range(2):
        lib = tmp_path / ("libtandav_host_periodic_%d.so" % k)
        shutil.copy(glue_host.build_host.lib_path(True), lib)
        libs.append(str(lib))
    h1 = glue_host.Host(True, path=libs[0])
    h1.init(pos, vel, m, softening=1 / 320.)
    f = tmp_path / "ewald_tables.dat"
    assert f.stat().st_size == 4 * 65 ** 3 * 8
    written = np.fromfile(f, dtype=np.float64).reshape(4, -1)
    e = Engine(periodic=True, softening=1e-3)
    np.testing.assert_array_equal(written.reshape(-1),
                                  e.ewald_table().reshape(-1))
    e.close()
    (2.0 * written).tofile(f)
    h2 = glue_host.Host(True, path=libs[1])
    h2.init(pos, vel, m, softening=1 / 320.)
    g = np.random.Generator(np.random.Philox(9))
    dr = g.uniform(-0.5, 0.5, (200, 3))
    np.testing.assert_array_equal(h2.ewald_correction(dr),
                                  2.0 * h1.ewald_correction(dr))
