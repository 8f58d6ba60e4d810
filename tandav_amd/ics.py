"""Synthetic initial conditions for the force-path workloads.

Every generator draws from numpy's Philox-4x32-10 counter-based bit
generator (``numpy.random.Philox``) seeded with the workload's seed, so the
same arrays are produced on any host.  Positions, velocities and masses are
FP64; units are G = 1.

* ``hernquist_sphere``  -- isolated Hernquist (1990) sphere, M = 1, a = 1,
  radii by inverting M(<r)/M = r^2/(r+a)^2 truncated at 100 a, isotropic
  directions, isotropic Gaussian velocities at v_c(r)/sqrt(3).
* ``zeldovich_box``     -- unit periodic box, particles on a regular n^3
  grid displaced by the Zel'dovich field of a Gaussian random density field
  with P(k) ~ k^-2.5 (modes 0 < |k| <= k_Nyquist), normalised so that the
  rms of |psi| equals the requested fraction of the grid spacing.
* ``clustered_box``     -- periodic box of Hernquist halos (power-law mass
  function dN/dM ~ M^-1.9 over three decades, r_s ~ M^(1/3), uniform
  centres) plus a uniform background.
"""
import os

import numpy as np


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def hernquist_sphere(n=65536, seed=1, a=1.0, mass=1.0, rmax_over_a=100.0,
                     G=1.0):
    g = rng(seed)
    qmax = (rmax_over_a / (rmax_over_a + 1.0)) ** 2
    q = g.uniform(0.0, qmax, n)
    sq = np.sqrt(q)
    r = a * sq / (1.0 - sq)
    mu = g.uniform(-1.0, 1.0, n)
    phi = g.uniform(0.0, 2.0 * np.pi, n)
    st = np.sqrt(1.0 - mu * mu)
    pos = np.stack([r * st * np.cos(phi), r * st * np.sin(phi), r * mu], 1)
    m_enc = mass * r * r / (r + a) ** 2
    vc = np.sqrt(G * m_enc / r)
    vel = g.normal(0.0, 1.0, (n, 3)) * (vc / np.sqrt(3.0))[:, None]
    m = np.full(n, mass / n)
    return pos, vel, m


def _fft_workers():
    return max(1, min(16, os.cpu_count() or 1))


def _zeldovich_displacement(ngrid, seed, slope, rms_frac):
    import scipy.fft
    g = rng(seed)
    white = g.normal(0.0, 1.0, (ngrid, ngrid, ngrid))
    dk = scipy.fft.rfftn(white, workers=_fft_workers())
    del white
    k1 = np.fft.fftfreq(ngrid) * ngrid          # integer wave numbers
    kz = np.fft.rfftfreq(ngrid) * ngrid
    kx = k1[:, None, None]
    ky = k1[None, :, None]
    kzz = kz[None, None, :]
    k2 = kx * kx + ky * ky + kzz * kzz
    knyq = ngrid // 2
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where((k2 > 0) & (k2 <= knyq * knyq),
                       np.power(k2, slope / 4.0), 0.0)   # sqrt(P) = k^(s/2)
        dk *= amp
        inv_k2 = np.where(k2 > 0, 1.0 / k2, 0.0)
    psi = []
    for kd in (kx, ky, kzz):
        comp = scipy.fft.irfftn(1j * kd * inv_k2 * dk, s=(ngrid,) * 3,
                                axes=(0, 1, 2), workers=_fft_workers())
        psi.append(comp.reshape(-1))
    psi = np.stack(psi, 1)
    rms = np.sqrt(np.mean(np.sum(psi * psi, axis=1)))
    return psi * (rms_frac / ngrid / rms)


def zeldovich_box(ngrid=128, seed=7, slope=-2.5, rms_frac=0.1, box=1.0,
                  vel_factor=0.0):
    """Regular grid + Zel'dovich displacement psi, wrapped into [0, box);
    velocities vel_factor * psi (the growing mode: a H f for comoving
    velocities), zero by default."""
    psi = _zeldovich_displacement(ngrid, seed, slope, rms_frac)
    n = ngrid ** 3
    vel = psi * (vel_factor * box) if vel_factor else np.zeros((n, 3))
    i = (np.arange(ngrid, dtype=np.float64) + 0.5) / ngrid
    pos = psi
    pos[:, 0] += np.repeat(i, ngrid * ngrid)
    pos[:, 1] += np.tile(np.repeat(i, ngrid), ngrid)
    pos[:, 2] += np.tile(i, ngrid * ngrid)
    if box != 1.0:
        pos *= box
    pos = np.mod(pos, box, out=pos)
    pos[pos >= box] -= box          # np.mod can round up to box
    return pos, vel, np.full(n, 1.0 / n)


def clustered_box(n=1 << 23, nhalo=256, seed=11, background=0.2,
                  slope=-1.9, decades=3.0, box=1.0, rs_max=0.02,
                  rmax_over_rs=20.0):
    """Hernquist halos with a power-law mass function in a periodic box."""
    g = rng(seed)
    nbg = int(round(background * n))
    nh = n - nbg
    # dN/dM ~ M^s on [1, 10^decades]: inverse CDF of M^(s+1)
    e = slope + 1.0
    u = g.uniform(0.0, 1.0, nhalo)
    mlo, mhi = 1.0, 10.0 ** decades
    mh = (mlo ** e + u * (mhi ** e - mlo ** e)) ** (1.0 / e)
    counts = np.floor(nh * mh / mh.sum()).astype(np.int64)
    counts[np.argmax(counts)] += nh - counts.sum()
    rs = rs_max * (mh / mh.max()) ** (1.0 / 3.0)
    centres = g.uniform(0.0, box, (nhalo, 3))
    parts = []
    qmax = (rmax_over_rs / (rmax_over_rs + 1.0)) ** 2
    for h in range(nhalo):
        c = int(counts[h])
        if c == 0:
            continue
        sq = np.sqrt(g.uniform(0.0, qmax, c))
        r = rs[h] * sq / (1.0 - sq)
        mu = g.uniform(-1.0, 1.0, c)
        phi = g.uniform(0.0, 2.0 * np.pi, c)
        st = np.sqrt(1.0 - mu * mu)
        parts.append(centres[h] + np.stack(
            [r * st * np.cos(phi), r * st * np.sin(phi), r * mu], 1))
    parts.append(g.uniform(0.0, box, (nbg, 3)))
    pos = np.mod(np.concatenate(parts, 0), box)
    pos[pos >= box] -= box
    return pos, np.zeros((n, 3)), np.full(n, 1.0 / n)
