"""Comoving kick/drift factors (comoving.c restated in tandav_amd.comoving):
the Einstein-de Sitter closed forms kick = 2/H0 (sqrt(a1) - sqrt(a0)) and
drift = 2/H0 (1/sqrt(a0) - 1/sqrt(a1)) (comoving.c:105-106), and direct
integrals of the integrands for a flat LCDM run."""
import numpy as np
import pytest
from scipy.integrate import quad

from tandav_amd.comoving import ComovingFactors, e_hubble


@pytest.mark.parametrize("a0,a1", [(0.1, 0.1005), (0.1, 0.2), (0.3, 0.31),
                                   (0.9, 1.0), (0.1001, 0.10011)])
def test_einstein_de_sitter_closed_form(a0, a1):
    h0 = 0.7
    c = ComovingFactors(0.1, 1.0, hubble=h0, omega_matter=1.0,
                        omega_lambda=0.0)
    kick = 2.0 / h0 * (np.sqrt(a1) - np.sqrt(a0))
    drift = 2.0 / h0 * (1.0 / np.sqrt(a0) - 1.0 / np.sqrt(a1))
    assert c.kick(a0, a1) == pytest.approx(kick, rel=1e-7)
    assert c.drift(a0, a1) == pytest.approx(drift, rel=1e-7)


def test_flat_lcdm_against_direct_integrals():
    om, ol, h0 = 0.3, 0.7, 1.0
    c = ComovingFactors(0.05, 1.0, hubble=h0, omega_matter=om,
                        omega_lambda=ol)

    def hub(a):
        return h0 * e_hubble(a, om + ol, om, ol)

    for a0, a1 in ((0.1, 0.1 * np.exp(0.01)), (0.5, 0.52), (0.06, 0.9)):
        k = quad(lambda a: 1.0 / (hub(a) * a * a), a0, a1, epsrel=1e-12)[0]
        d = quad(lambda a: 1.0 / (hub(a) * a ** 3), a0, a1, epsrel=1e-12)[0]
        assert c.kick(a0, a1) == pytest.approx(k, rel=1e-7)
        assert c.drift(a0, a1) == pytest.approx(d, rel=1e-7)
    # E(a) -> 1 today for a flat model, and the matter term dominates early
    assert e_hubble(1.0, 1.0, om, ol) == pytest.approx(1.0)
    assert e_hubble(1e-3, 1.0, om, ol) == pytest.approx(
        np.sqrt(om / 1e-9 + ol), rel=1e-12)
