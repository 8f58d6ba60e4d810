"""tdv_sort_keys_noalloc, the key sort that allocates nothing (its buffers
are a static per-device scratch array), against tdv_sort_keys, the same
sort in a context's own buffers: bit-identical permutations on calls in a
row (the second smaller, the scratch keeping the first call's contents),
the tie fallback, the largest accepted size, the size limit and duplicate
keys."""
import ctypes

import numpy as np
import pytest

from tandav_amd import Engine, engine

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20              # TDV_NOALLOC_MAX_KEYS
ERR_ARG, ERR_DUP = 1, 4


def _keys(n, seed, long_runs=False):
    g = np.random.Generator(np.random.Philox(seed))
    k = g.integers(0, 2 ** 63, (n, 2), dtype=np.uint64)
    k[:, 1] |= np.uint64(1 << 63) * (g.integers(0, 2, n).astype(np.uint64))
    if long_runs:                # > 4096 equal high words: 128-bit fallback
        k[: n // 2, 1] = k[0, 1]
    return k


def _noalloc(keys2, stream=None):
    """tdv_sort_keys_noalloc on device copies of keys2 (n x 2 words)."""
    from hip_mem import DeviceBuffer
    L = engine.load_library()
    n = keys2.shape[0]
    kd = DeviceBuffer.of(np.ascontiguousarray(keys2, dtype=np.uint64))
    pd = DeviceBuffer.of(np.full(max(n, 1), 0xFFFFFFFF, np.uint32))
    rc = L.tdv_sort_keys_noalloc(ctypes.c_size_t(n), kd.ptr, pd.ptr,
                                 stream.handle if stream else None)
    got = pd.to_host(np.uint32, n)
    kd.free()
    pd.free()
    return rc, got


def test_noalloc_sort_matches_context_sort():
    from hip_mem import Stream
    e = Engine()
    st = Stream()
    for n, seed, runs, strm in ((300001, 31, True, None),
                                (4097, 32, False, st), (MAX_KEYS, 33, False, st),
                                (70000, 34, True, None), (1, 35, False, st)):
        k = _keys(n, seed, runs)
        print("noalloc sort n=%d seed=%d long_runs=%s stream=%s"
              % (n, seed, runs, strm is not None), flush=True)
        rc, got = _noalloc(k, strm)
        assert rc == 0, rc
        want = e.sort_keys(k)
        np.testing.assert_array_equal(got, want)
        # ascending 128-bit order, a permutation
        srt = k[got.astype(np.int64)]
        assert np.array_equal(np.sort(got), np.arange(n, dtype=np.uint32))
        hi, lo = srt[:, 1], srt[:, 0]
        assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1])
                                            & (lo[1:] > lo[:-1])))
    st.destroy()
    e.close()


def test_noalloc_sort_limits_and_duplicates():
    L = engine.load_library()
    # beyond the scratch: refused before anything is launched
    rc = L.tdv_sort_keys_noalloc(ctypes.c_size_t(MAX_KEYS + 1), None, None,
                                 None)
    assert rc == ERR_ARG
    assert L.tdv_sort_keys_noalloc(ctypes.c_size_t(0), None, None, None) == 0
    k = _keys(5000, 36)
    k[4000] = k[17]
    rc, _ = _noalloc(k)
    assert rc == ERR_DUP
    # the flag does not outlive the call
    k = _keys(3000, 37)
    rc, got = _noalloc(k)
    assert rc == 0
    e = Engine()
    np.testing.assert_array_equal(got, e.sort_keys(k))
    e.close()
