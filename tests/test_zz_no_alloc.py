"""The allocation-free 2-NN entry point (akz_match2nn_device_noalloc) against
the allocating one (akz_match2nn_device) through the C ABI: identical
DMatches on two calls in a row, the second smaller, and the oracle on the
smaller one.  Sizes beyond the fixed scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest

import akaze_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def rows64(desc):
    out = np.zeros((len(desc), 64), np.uint8)
    out[:, :61] = desc
    out[:, 60] &= 0x3F
    return out


def device_copy(arr):
    buf = A.DeviceBuffer(max(arr.nbytes, 1))
    buf.upload(arr)
    return buf


def run_both(L, m, q, t):
    nq, nt = len(q), len(t)
    dq, dt = device_copy(rows64(q)), device_copy(rows64(t))
    d_ref = A.DeviceBuffer(16 * 2 * nq)
    d_new = A.DeviceBuffer(16 * 2 * nq)
    assert L.akz_match2nn_device(m, dq.ptr, nq, dt.ptr, nt, d_ref.ptr) == 0
    assert L.akz_matcher_wait(m) == 0
    assert L.akz_match2nn_device_noalloc(dq.ptr, nq, dt.ptr, nt, d_new.ptr, None) == 0
    assert L.akz_device_synchronize() == 0
    ref = d_ref.download(np.zeros(2 * nq, O.DM_DTYPE)).reshape(-1, 2)
    new = d_new.download(np.zeros(2 * nq, O.DM_DTYPE)).reshape(-1, 2)
    return ref, new


def test_noalloc_matches_allocating_entry_point_twice():
    L = A.lib()
    m = ctypes.c_void_p()
    assert L.akz_matcher_create(ctypes.byref(m)) == 0
    try:
        rng = np.random.default_rng(77)
        cases = [(5000, 9000), (700, 1200)]   # the second call is smaller
        for k, (nq, nt) in enumerate(cases):
            print("noalloc call %d: nq=%d nt=%d" % (k, nq, nt), flush=True)
            t = rng.integers(0, 256, (nt, 61), dtype=np.uint8)
            q = rng.integers(0, 256, (nq, 61), dtype=np.uint8)
            q[: nq // 4] = t[rng.integers(0, nt, nq // 4)]   # exact matches
            t[5] = t[6]                                       # a tie in distance
            ref, new = run_both(L, m, q, t)
            assert np.array_equal(ref.view(np.uint8), new.view(np.uint8))
            if k == 1:
                q61 = q.copy()
                q61[:, 60] &= 0x3F
                t61 = t.copy()
                t61[:, 60] &= 0x3F
                orc = O.match2nn(q61, t61)
                assert np.array_equal(new["trainIdx"], orc["trainIdx"])
                assert np.array_equal(new["distance"], orc["distance"])
    finally:
        L.akz_matcher_destroy(m)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    L = A.lib()
    one = ctypes.c_void_p(1)   # never dereferenced: the size check comes first
    st = L.akz_match2nn_device_noalloc(one, (1 << 20) + 1, one, 10, one, None)
    assert st == 1
    assert L.akz_match2nn_device_noalloc(None, 10, None, 10, None, None) == 1
    assert L.akz_match2nn_device_noalloc(one, 0, one, 10, one, None) == 0
