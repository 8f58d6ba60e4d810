"""The FP32 walk's tree build (node links by binary searches over the
sorted keys, centres of mass summed bottom-up over the levels) against the
FP64 build, whose every node the parity tests hold bit for bit to the
oracle: the same nodes, DNext, Bitfield, DUp, Npart, node centres and
Tree_Parent exactly; masses and centres of mass (FP64 sums in another
order, stored as float) within one float ulp; the top nodes likewise.
Boxes: the 65,536-particle Hernquist sphere (non-periodic), the periodic
Zel'dovich 32^3 box, the clustered blobs, particles below the key
resolution and exact copies (dumped into level-41 nodes)."""
import numpy as np
import pytest

from tandav_amd import Engine, ics
from test_gpu_dump import _dump_box

pytestmark = pytest.mark.gpu


def _built(pos, m, periodic, precision):
    kw = dict(precision="fp32", compat=False) if precision == "fp32" else {}
    e = Engine(periodic=periodic, softening=1e-3, alpha_rel=0.005, **kw)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.tree_build()
    out = (e.tree(), e.top_nodes(), e.get("tree_parent")["tree_parent"])
    e.close()
    return out


def _within_ulp(a, b, name):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ulp = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    bad = np.abs(a.astype(np.float64) - b.astype(np.float64)) > ulp
    assert not bad.any(), (name, np.argwhere(bad)[:5])


def _cases():
    g = np.random.Generator(np.random.Philox(2))
    pos, _, m = ics.hernquist_sphere(65536, seed=1)
    yield "hernquist", pos, m, False
    pos, _, m = ics.zeldovich_box(32, seed=7)
    yield "zeldovich", pos, m, True
    pos, _, m = ics.clustered_box(1 << 17, nhalo=16, seed=3)
    yield "clustered", pos, m, True
    x, mx = _dump_box()
    yield "dumped", x, mx, True
    pos, _, _ = ics.zeldovich_box(32, seed=7)
    copies = pos[g.choice(len(pos), 7, replace=False)]
    x = np.vstack([pos, copies, copies[:2]])
    yield "copies", x, np.full(len(x), 1.0 / len(x)), True


@pytest.mark.parametrize("case", ["hernquist", "zeldovich", "clustered",
                                  "dumped", "copies"])
def test_fp32_build_matches_fp64_build(case):
    name, pos, m, periodic = next(c for c in _cases() if c[0] == case)
    t64, top64, tp64 = _built(pos, m, periodic, "fp64")
    t32, top32, tp32 = _built(pos, m, periodic, "fp32")
    assert len(t32) == len(t64)
    for f in ("DNext", "Bitfield", "DUp", "Npart", "Pos"):
        np.testing.assert_array_equal(t32[f], t64[f], err_msg=f)
    np.testing.assert_array_equal(tp32, tp64)
    _within_ulp(t32["Mass"], t64["Mass"], "Mass")
    _within_ulp(t32["CoM"], t64["CoM"], "CoM")
    for f in ("Key", "Target", "Level", "First_Part", "Npart", "Pos"):
        np.testing.assert_array_equal(top32[f], top64[f], err_msg=f)
    _within_ulp(top32["Mass"], top64["Mass"], "top Mass")
    _within_ulp(top32["CoM"], top64["CoM"], "top CoM")
