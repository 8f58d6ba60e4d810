"""The snapshot header layouts agree: the numpy statement of io.h's
struct gadget_header, a ctypes mirror of it with io.h's own 59 fill bytes,
and the engine's ctypes SnapshotHeader (the C ABI's tdv_snapshot_header,
whose size the library checks at compile time)."""
import ctypes

import numpy as np

import snapshot_format as sf


class _IoH(ctypes.Structure):
    """io.h:23-42 as written there."""
    _fields_ = [("Npart", ctypes.c_uint32 * 6),
                ("Massarr", ctypes.c_double * 6),
                ("Time", ctypes.c_double), ("Redshift", ctypes.c_double),
                ("Flag_Sfr", ctypes.c_int32), ("Flag_Feedback", ctypes.c_int32),
                ("Nall", ctypes.c_uint32 * 6),
                ("Flag_Cooling", ctypes.c_int32), ("Num_Files", ctypes.c_int32),
                ("Boxsize", ctypes.c_double), ("Omega0", ctypes.c_double),
                ("Omega_Lambda", ctypes.c_double),
                ("Hubble_Param", ctypes.c_double),
                ("Flag_Age", ctypes.c_int32), ("Flag_Metals", ctypes.c_int32),
                ("Nall_High_Word", ctypes.c_uint32 * 6),
                ("fill_bytes", ctypes.c_char * 59)]


def test_header_layouts_agree():
    from tandav_amd.engine import SnapshotHeader
    assert ctypes.sizeof(_IoH) == 256        # write_snapshot.c:188's Assert
    assert ctypes.sizeof(SnapshotHeader) == 256
    assert sf.HEADER.itemsize == 256
    for name in sf.HEADER.names:
        off = sf.HEADER.fields[name][1]
        assert getattr(_IoH, name).offset == off, name
        assert getattr(SnapshotHeader, name).offset == off, name


def test_numpy_format_round_trip(tmp_path):
    g = np.random.Generator(np.random.Philox(1))
    pos = g.random((100, 3))
    ids = np.arange(100, dtype=np.uint32)
    h = sf.header([0, 100, 0, 0, 0, 0], Time=0.25, Boxsize=1.0)
    sf.write(tmp_path / "s", h, [("POS ", pos), ("ID  ", ids)])
    blocks, head = sf.read(tmp_path / "s")
    assert list(blocks) == ["HEAD", "POS ", "ID  "]
    assert head["Time"] == 0.25 and list(head["Npart"]) == [0, 100, 0, 0, 0, 0]
    np.testing.assert_array_equal(
        np.frombuffer(blocks["POS "], "<f8").reshape(-1, 3), pos)
    np.testing.assert_array_equal(np.frombuffer(blocks["ID  "], "<u4"), ids)
