"""The reference-side binding (tandav_amd/glue/tandav_glue.c) built into a
minimal host program (tests/glue/host.c): it links, exports every Tandav
symbol of the force path, and its pure host functions agree with the
oracle.  No GPU is touched here."""
import subprocess
import sys
import os

import numpy as np
import pytest

import oracle

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "glue"))
import build_host  # noqa: E402
import host as glue_host  # noqa: E402

pytestmark = pytest.mark.cpu


@pytest.fixture(scope="module", autouse=True)
def built():
    build_host.build()


@pytest.mark.parametrize("periodic", [False, True])
def test_binding_exports_tandav_symbols(periodic):
    out = subprocess.run(["nm", "-D", "--defined-only",
                          build_host.lib_path(periodic)],
                         capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    want = glue_host.EXPORTS_PERIODIC if periodic else glue_host.EXPORTS
    missing = [s for s in want if s not in names]
    assert not missing, missing
    # the engine does the work: the binding links against its C ABI
    und = subprocess.run(["nm", "-D", "--undefined-only",
                          build_host.lib_path(periodic)],
                         capture_output=True, text=True, check=True).stdout
    assert "tdv_compute_acceleration" in und
    assert "tdv_sort_particles" in und


def test_reverse_peano_key_matches_reversed_keys():
    """Reverse_Peano_Key (peano.c:278-296), which Tandav never calls, run
    on standard keys: bit for bit the oracle's reversed keys."""
    h = glue_host.Host(periodic=False)
    g = np.random.Generator(np.random.Philox(3))
    x, y, z = g.uniform(0, 1, (3, 4096))
    std = oracle.peano_keys(x, y, z, np.zeros(3), 1.0)
    rev = oracle.peano_keys(x, y, z, np.zeros(3), 1.0, reversed_=True)
    np.testing.assert_array_equal(h.reverse_peano_key(std), rev)


def test_cmp_peano_keys_is_unsigned_128_bit_order():
    h = glue_host.Host(periodic=False)
    g = np.random.Generator(np.random.Philox(4))
    a = g.integers(0, 2**64, (500, 2), dtype=np.uint64)
    b = a.copy()
    b[:200, 1] = g.integers(0, 2**64, 200, dtype=np.uint64)  # hi differs
    b[200:400, 0] = g.integers(0, 2**64, 200, dtype=np.uint64)  # lo only
    for i in range(len(a)):
        ia = int(a[i, 1]) << 64 | int(a[i, 0])
        ib = int(b[i, 1]) << 64 | int(b[i, 0])
        assert h.cmp_keys(a[i], b[i]) == (ia > ib) - (ia < ib)
