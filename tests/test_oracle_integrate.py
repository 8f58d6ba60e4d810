"""Pins the integrator's restatement (oracle/integrate.py) and the engine's
host-side comoving tables (tdv_comoving_tables, no GPU):

  * the natural cubic spline against scipy's CubicSpline(bc_type="natural")
    on the reference's own knot layout (the (0, 0) knot included);
  * the comoving kick/drift integrals against scipy's adaptive quadrature
    (tandav_amd/comoving.py, QUADPACK) for flat LCDM and against the
    Einstein-de Sitter closed forms 2/H0 (sqrt(a) - sqrt(a0)) (kick) and
    2/H0 (1/sqrt(a0) - 1/sqrt(a)) (drift) (comoving.c:104-106);
  * timestep2timebin by brute force: the bin is the largest b with
    2^b Step_Min < dt;
  * the float rounding of cosmological_timestep on a hand-worked value;
  * non-comoving kick/drift factors against t(it) = Time.Begin + it Step_Min.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import integrate as I
from tandav_amd import engine
from tandav_amd.comoving import ComovingFactors

pytestmark = pytest.mark.cpu


def _tables(tl):
    try:
        return engine.comoving_tables(tl)
    except engine.TDVError as e:        # library missing: no CPU fallback
        pytest.skip(str(e))


def test_natural_spline_matches_scipy():
    tl = engine.timeline(True, 0.1, 1.0, hubble=2.0, omega_matter=0.3,
                         omega_lambda=0.7)
    a, k, d = _tables(tl)
    for y in (k, d):
        ours = I.NaturalSpline(a, y)
        ref = CubicSpline(a, y, bc_type="natural")
        q = np.concatenate([a, np.linspace(0.0, a[-1], 20011)])
        np.testing.assert_allclose(ours(q), ref(q), rtol=1e-11,
                                   atol=1e-12 * np.abs(y).max())


@pytest.mark.parametrize("om,ol", [(0.3, 0.7), (1.0, 0.0)])
def test_comoving_tables(om, ol):
    h0 = np.sqrt(8 * np.pi / (3 * om))
    tl = engine.timeline(True, 0.05, 1.0, hubble=h0, omega_matter=om,
                         omega_lambda=ol)
    a, k, d = _tables(tl)
    assert a[0] == 0 and k[0] == 0 and d[0] == 0
    beg = 0.95 * 0.05
    np.testing.assert_allclose(a[1:], beg * np.exp(
        np.log(1.05 / beg) / 511 * np.arange(1, 512)), rtol=1e-14)
    ref = ComovingFactors(0.05, 1.0, hubble=h0, omega_matter=om,
                          omega_lambda=ol)
    # scipy's spline tables come from QUADPACK to 1e-8 relative requested;
    # the achieved accuracy on these smooth integrands is far better
    np.testing.assert_allclose(k[1:], ref._kick(a[1:]), rtol=1e-11)
    np.testing.assert_allclose(d[1:], ref._drift(a[1:]), rtol=1e-11)
    if om == 1.0:
        np.testing.assert_allclose(k[1:], 2 / h0 * (np.sqrt(a[1:])
                                                    - np.sqrt(beg)),
                                   rtol=1e-13)
        np.testing.assert_allclose(d[1:], 2 / h0 * (1 / np.sqrt(beg)
                                                    - 1 / np.sqrt(a[1:])),
                                   rtol=1e-13)


def test_timebin_is_largest_power_below_dt():
    tl = I.Timeline(False, 0.0, 3.0)
    g = np.random.Generator(np.random.Philox(4))
    dt = tl.step_min * 2.0 ** g.uniform(0.0, 31.0, 5000)
    dt = np.concatenate([dt, tl.step_min * 2.0 ** np.arange(1, 31)])
    _, e = np.frexp(tl.step_max / dt)
    got = I.NBITS - 1 - e
    for b, t in zip(got, dt):
        ok = [bb for bb in range(31) if 2.0 ** bb * tl.step_min < t]
        assert b == max(ok)


def test_cosmological_timestep_rounds_through_float():
    """dt = accuracy sqrt(2 eps / |a|) is returned as a float
    (timestep.c:495-526): 0.0025 * sqrt(2 * 105/32 * 0.01 / 3) rounds to
    the float 0.00036979728 (0x39b9... in single precision) before the
    bin is taken, so a step just above a bin edge in double but below it
    in float lands one bin lower."""
    tl = I.Timeline(False, 0.0, 1.0)
    f = 0.0025 * np.sqrt(2 * 105.0 / 32.0 * 0.01 / 3.0)
    ga = np.array([[3.0, 0.0, 0.0]])
    b, _, _, ok = I.timebins(tl, ga, np.zeros(1, np.int32), 0.0025, 0.01,
                             1e30, 31)
    assert ok
    dtf = float(np.float32(f))
    _, e = np.frexp(tl.step_max / dtf)
    assert b[0] == I.NBITS - 1 - e
    # a step a hair above the bin edge 2^20 Step_Min in double rounds onto
    # the edge in float; the edge itself is not strictly below the step, so
    # the particle falls to bin 19 (in double it would take bin 20)
    edge = 2.0 ** 20 * tl.step_min
    assert float(np.float32(edge)) == edge
    dt = edge * (1 + 2.0 ** -40)
    eps = 105.0 / 32.0 * 0.01
    acc = 2 * eps / dt ** 2
    assert np.sqrt(2 * eps / acc) > edge
    b, _, _, _ = I.timebins(tl, np.array([[acc, 0, 0]]), np.zeros(1, np.int32),
                            1.0, 0.01, 1e30, 31)
    assert b[0] == 19
    _, e = np.frexp(tl.step_max / np.sqrt(2 * eps / acc))
    assert I.NBITS - 1 - e == 20


def test_non_comoving_factors_closed_form():
    tl = I.Timeline(False, 0.5, 2.5)
    g = np.random.Generator(np.random.Philox(5))
    i0 = g.integers(0, 1 << 30, 1000).astype(np.uint32)
    i1 = (i0 + g.integers(0, 1 << 20, 1000)).astype(np.uint32)
    t0 = 0.5 + i0.astype(np.float64) * tl.step_min
    t1 = 0.5 + i1.astype(np.float64) * tl.step_min
    np.testing.assert_array_equal(tl.kick_step(i0, i1), t1 - t0)
    np.testing.assert_array_equal(tl.drift_step(i0, i1), t1 - t0)
