"""The FP32 walk's quadrupoles, formed bottom-up on the device (children's
second moments shifted to the parent's centre of mass, plus particles no
child node covers), against direct FP64 sums over each node's particles on
the CPU: T = sum m (3 y y^T - |y|^2 I) about the node's centre of mass (a
top node's float one), in the walk's units."""
import numpy as np
import pytest

from tandav_amd import Engine, ics

pytestmark = pytest.mark.gpu


def _traceless(pos, m, c):
    y = pos - c
    s = np.einsum("i,ij,ik->jk", m, y, y)
    t = 3 * s - np.trace(s) * np.eye(3)
    return np.array([t[0, 0], t[1, 1], t[0, 1], t[0, 2], t[1, 2]]), np.trace(s)


@pytest.mark.parametrize("case", ["zeldovich", "clustered", "hernquist"])
def test_walk32_quadrupoles_match_direct_sums(case):
    if case == "zeldovich":
        pos, _, m = ics.zeldovich_box(32, seed=7)
        kw = dict(periodic=True, softening=1 / 1280.)
    elif case == "clustered":
        pos, _, m = ics.clustered_box(1 << 15, nhalo=16, seed=3)
        kw = dict(periodic=True, softening=1 / 2560.)
    else:
        pos, _, m = ics.hernquist_sphere(20000, seed=4)
        kw = dict(periodic=False, softening=0.01)
    e = Engine(alpha_rel=0.01, precision="fp32", compat=False,
               ewald_size=0.125, ewald_accuracy=1.0, fp32_order=2,
               fp32_criterion=3, **kw)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    q, tq, first = e.walk32_quadrupoles()
    tree, top = e.tree(), e.top_nodes()
    P = e.get("pos", "mass")
    x, mass = P["pos"], P["mass"]
    if kw["periodic"]:
        S = 2.0 ** 32
    else:
        S = 2.0 ** 31 / e.domain[0]
    qscale = 1.0 * S ** 4                  # G m S^2 masses, S lengths
    assert len(first) == len(tree)
    checked = 0
    for j in range(len(tree)):
        f = first[j]
        if f < 0:
            continue
        npart = tree["Npart"][j]
        want, tr = _traceless(x[f:f + npart], mass[f:f + npart],
                              tree["CoM"][j])
        tol = 2e-6 * tr * qscale + 1e-30
        np.testing.assert_allclose(q[j, :5], want * qscale, rtol=0, atol=tol,
                                   err_msg="node %d" % j)
        checked += 1
    assert checked > 100
    for t in range(len(top)):
        f, npart = top["First_Part"][t], top["Npart"][t]
        want, tr = _traceless(x[f:f + npart], mass[f:f + npart],
                              top["CoM"][t].astype(np.float64))
        np.testing.assert_allclose(tq[t, :5], want * qscale, rtol=0,
                                   atol=2e-6 * tr * qscale + 1e-30,
                                   err_msg="top node %d" % t)
