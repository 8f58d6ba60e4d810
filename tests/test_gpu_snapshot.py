"""Gadget format-2 snapshots through the C ABI (tdv_write_snapshot,
tdv_read_snapshot) against the numpy statement of the format in
snapshot_format.py: the engine's file holds the engine's columns block by
block, and files written by numpy load into the engine as written."""
import numpy as np
import pytest

import snapshot_format as sf
from tandav_amd import Engine, TDVError

pytestmark = pytest.mark.gpu


def _particles(n, seed):
    g = np.random.Generator(np.random.Philox(seed))
    pos = g.random((n, 3))
    vel = g.normal(size=(n, 3))
    m = g.uniform(0.5, 1.5, n) / n
    return pos, vel, m


def test_write_matches_engine_columns(tmp_path):
    pos, vel, m = _particles(6000, 1)
    e = Engine(periodic=False, softening=1e-3)
    e.set_particles(pos, m, vel=vel, ids=np.arange(6000) + 7)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    path = str(tmp_path / "snap_000")
    e.write_snapshot(path, npart=[0, 6000, 0, 0, 0, 0], time=0.5,
                     redshift=1.0, boxsize=1.0, omega0=0.3,
                     omega_lambda=0.7, hubble_param=0.7)
    blocks, head = sf.read(path)
    assert list(blocks) == ["HEAD", "POS ", "VEL ", "ID  ", "MASS", "ACC ",
                            "GACC"]
    assert list(head["Npart"]) == [0, 6000, 0, 0, 0, 0]
    assert list(head["Nall"]) == [0, 6000, 0, 0, 0, 0]
    assert head["Num_Files"] == 1 and head["Time"] == 0.5
    assert head["Boxsize"] == 1.0 and head["Omega0"] == 0.3
    got = e.get("pos", "vel", "mass", "id", "acc", "grav_acc")
    f8 = lambda b, k: np.frombuffer(blocks[b], "<f8").reshape(-1, k)
    np.testing.assert_array_equal(f8("POS ", 3), got["pos"])
    np.testing.assert_array_equal(f8("VEL ", 3), got["vel"])
    np.testing.assert_array_equal(f8("MASS", 1)[:, 0], got["mass"])
    np.testing.assert_array_equal(f8("ACC ", 1)[:, 0], got["acc"][:, 0])
    np.testing.assert_array_equal(f8("GACC", 3), got["grav_acc"])
    np.testing.assert_array_equal(np.frombuffer(blocks["ID  "], "<u4"),
                                  got["id"])
    # and back: the same columns, the same header
    e2 = Engine(periodic=False, softening=1e-3)
    h2 = e2.read_snapshot(path)
    back = e2.get("pos", "vel", "mass", "id")
    for k in back:
        np.testing.assert_array_equal(back[k], got[k], err_msg=k)
    assert h2["Time"] == 0.5 and h2["Npart"] == [0, 6000, 0, 0, 0, 0]


@pytest.mark.parametrize("header_mass", [False, True])
def test_read_numpy_file(tmp_path, header_mass):
    pos, vel, m = _particles(3000, 2)
    ids = np.arange(3000, dtype=np.uint32)
    npart = [1000, 2000, 0, 0, 0, 0]
    massarr = (0.002, 0.0005, 0, 0, 0, 0) if header_mass else (0.0,) * 6
    blocks = [("POS ", pos), ("VEL ", vel), ("ID  ", ids)]
    if not header_mass:
        blocks.append(("MASS", m))
    sf.write(tmp_path / "ics", sf.header(npart, massarr, Time=0.1), blocks)
    e = Engine(periodic=False, softening=1e-3)
    head = e.read_snapshot(str(tmp_path / "ics"))
    got = e.get("pos", "vel", "mass", "id")
    np.testing.assert_array_equal(got["pos"], pos)
    np.testing.assert_array_equal(got["vel"], vel)
    np.testing.assert_array_equal(got["id"], ids)
    want = (np.r_[np.full(1000, 0.002), np.full(2000, 0.0005)]
            if header_mass else m)
    np.testing.assert_array_equal(got["mass"], want)
    assert head["Time"] == 0.1


def test_read_multiple_files(tmp_path):
    pos, vel, m = _particles(5000, 3)
    ids = np.arange(5000, dtype=np.uint32)
    cut = [0, 2100, 5000]
    for k in range(2):
        s = slice(cut[k], cut[k + 1])
        npk = [0, cut[k + 1] - cut[k], 0, 0, 0, 0]
        sf.write(tmp_path / ("ics.%d" % k), sf.header(npk), [
            ("POS ", pos[s]), ("VEL ", vel[s]), ("ID  ", ids[s]),
            ("MASS", m[s])])
    e = Engine(periodic=False, softening=1e-3)
    e.read_snapshot(str(tmp_path / "ics"))
    got = e.get("pos", "mass", "id")
    np.testing.assert_array_equal(got["pos"], pos)
    np.testing.assert_array_equal(got["mass"], m)
    np.testing.assert_array_equal(got["id"], ids)


def test_read_refuses_broken_files(tmp_path):
    pos, vel, m = _particles(100, 4)
    ids = np.arange(100, dtype=np.uint32)
    h = sf.header([0, 100, 0, 0, 0, 0])
    sf.write(tmp_path / "swapped", h.astype(sf.HEADER.newbyteorder(">")),
             [("POS ", pos.astype(">f8")), ("VEL ", vel.astype(">f8")),
              ("ID  ", ids.astype(">u4")), ("MASS", m.astype(">f8"))],
             endian=">")
    sf.write(tmp_path / "noid", h, [("POS ", pos), ("VEL ", vel),
                                     ("MASS", m)])
    e = Engine(periodic=False, softening=1e-3)
    for name in ("swapped", "noid", "missing"):
        with pytest.raises(TDVError):
            e.read_snapshot(str(tmp_path / name))
