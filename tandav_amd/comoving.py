"""Comoving kick and drift factors of the KDK step (the caller's side of
tdv_kick_drift), restating Tandav's comoving.c.

comoving.c:61-69: in code units ("dt = da") the symplectic factors of Quinn,
Katz, Stadel & Lake (1997) are the integrals over the expansion factor of
    drift: 1 / (H(a) a^3),        kick: 1 / (H(a) a^2),
with H(a) = H0 E(a), E(a) = sqrt(Omega_L + (1 - Omega_0) / a^2 + Omega_m /
a^3 + Omega_r / a^4) (cosmology.c:65-74).

comoving.c:109-153: the integrals are tabulated once at TABLESIZE = 512
expansion factors a_i = 0.95 a_begin exp(i d), d = ln(1.05 a_end / (0.95
a_begin)) / 511, i = 1..511, each from 0.95 a_begin (adaptive 41-point
Gauss-Kronrod to 1e-8 relative, GSL's qag; here QUADPACK's qags, the same
method family, through scipy), table entry 0 left at a = 0 with value 0,
and read through natural cubic splines (GSL's cspline); a step's factor is
the difference of the spline at its two ends (comoving.c:30-50).
"""
import numpy as np
from scipy.integrate import quad
from scipy.interpolate import CubicSpline

TABLESIZE = 512


def e_hubble(a, omega_0, omega_matter, omega_lambda, omega_rad=0.0):
    """E(a) = H(a) / H0 (cosmology.c:70-74)."""
    return np.sqrt(omega_lambda + (1.0 - omega_0) / (a * a)
                   + omega_matter / (a * a * a)
                   + omega_rad / (a * a * a * a))


class ComovingFactors:
    """Kick/drift factors for a run from a_begin to a_end."""

    def __init__(self, a_begin, a_end, hubble=1.0, omega_matter=0.3,
                 omega_lambda=0.7, omega_rad=0.0, omega_0=None):
        if omega_0 is None:
            omega_0 = omega_matter + omega_lambda + omega_rad
        self.h0 = hubble

        def hub(a):
            return hubble * e_hubble(a, omega_0, omega_matter, omega_lambda,
                                     omega_rad)

        t_beg, t_end = 0.95 * a_begin, 1.05 * a_end
        d = np.log(t_end / t_beg) / (TABLESIZE - 1)
        a = np.zeros(TABLESIZE)
        drift = np.zeros(TABLESIZE)
        kick = np.zeros(TABLESIZE)
        for i in range(1, TABLESIZE):
            a[i] = t_beg * np.exp(d * i)
            drift[i] = quad(lambda x: 1.0 / (hub(x) * x ** 3), t_beg, a[i],
                            epsabs=0.0, epsrel=1e-8, limit=TABLESIZE)[0]
            kick[i] = quad(lambda x: 1.0 / (hub(x) * x ** 2), t_beg, a[i],
                           epsabs=0.0, epsrel=1e-8, limit=TABLESIZE)[0]
        self.table_a = a
        self._drift = CubicSpline(a, drift, bc_type="natural")
        self._kick = CubicSpline(a, kick, bc_type="natural")

    def kick(self, a_curr, a_next):
        """Particle_Kick_Step (comoving.c:30-39)."""
        return float(self._kick(a_next) - self._kick(a_curr))

    def drift(self, a_curr, a_next):
        """Particle_Drift_Step (comoving.c:41-50)."""
        return float(self._drift(a_next) - self._drift(a_curr))
