"""The per-element key function on the host (tdv_peano_keys_host, which the
binding's Peano_Key and Reversed_Peano_Key call instead of a device round
trip) against the oracle's keys, bit for bit: random points over a domain
with a negative origin, the 17^3 corner grid (the quantiser's extremes),
and the argument checks.  No device is touched."""
import numpy as np
import pytest

import oracle
from tandav_amd import engine

pytestmark = pytest.mark.cpu


def test_host_keys_match_oracle():
    g = np.random.Generator(np.random.Philox(31))
    x = g.uniform(-3.0, 5.0, (50000, 3))
    origin = np.array([-3.0, -3.0, -3.0])
    for rev in (False, True):
        got = engine.peano_keys_host(x[:, 0], x[:, 1], x[:, 2], origin,
                                     8.0001, rev)
        want = oracle.peano_keys(x[:, 0], x[:, 1], x[:, 2], origin, 8.0001,
                                 rev)
        np.testing.assert_array_equal(got, want)
    i = np.arange(17) / 16.0
    gx, gy, gz = np.meshgrid(i, i, i, indexing="ij")
    p = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) * 0.9999999999
    for rev in (False, True):
        np.testing.assert_array_equal(
            engine.peano_keys_host(p[:, 0], p[:, 1], p[:, 2], [0, 0, 0], 1.0,
                                   rev),
            oracle.peano_keys(p[:, 0], p[:, 1], p[:, 2], [0, 0, 0], 1.0, rev))


def test_host_keys_on_dyadic_boundaries():
    """Points on and next to the cell walls of levels 1-42 (the bits the
    transform's undo steps compare), where a step left out or one too many
    would change the key."""
    g = np.random.Generator(np.random.Philox(32))
    lv = g.integers(1, 43, 60000)
    k = np.floor(g.random(60000) * 2.0 ** np.minimum(lv, 52)).astype(np.float64)
    wall = k / 2.0 ** lv
    x = np.stack([wall, g.random(60000), g.random(60000)], 1)
    for c in range(3):
        xx = np.roll(x, c, axis=1)
        for nudge in (0.0, 1e-15, -1e-15):
            p = np.clip(xx + nudge, 0.0, 0.999999999999)
            for rev in (False, True):
                np.testing.assert_array_equal(
                    engine.peano_keys_host(p[:, 0], p[:, 1], p[:, 2],
                                           [0, 0, 0], 1.0, rev),
                    oracle.peano_keys(p[:, 0], p[:, 1], p[:, 2], [0, 0, 0],
                                      1.0, rev))


def test_host_keys_reject_bad_size():
    with pytest.raises(engine.TDVError):
        engine.peano_keys_host([0.5], [0.5], [0.5], [0, 0, 0], 0.0)
    assert engine.peano_keys_host([], [], [], [0, 0, 0], 1.0).shape == (0, 2)
