"""gs_radix_sort_indices_noalloc (the allocation-free radixSortIndices on
device arrays, scratch in a static device array) against the handle-based
gs_radix_sort_indices, through the C ABI: two calls in a row, the second
smaller, on a stream of the test's own and on the null stream, and the size
limit refused before any work.  Device buffers come from the same HIP runtime
the library links (hipMalloc through ctypes)."""
import ctypes

import numpy as np
import pytest

import galaxy_hip
from gpu_util import banner

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    galaxy_hip.lib()  # loads libamdhip64 as the library's dependency
    h = ctypes.CDLL("libamdhip64.so.7")
    for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy",
                 "hipStreamSynchronize", "hipDeviceSynchronize"):
        getattr(h, name).restype = ctypes.c_int
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return h


def _ok(rc, what):
    assert rc == 0, f"{what}: HIP error {rc}"


def _noalloc_sort(hip, codes, stream):
    n = codes.size
    dc, dp = ctypes.c_void_p(), ctypes.c_void_p()
    _ok(hip.hipMalloc(ctypes.byref(dc), 4 * n), "hipMalloc")
    _ok(hip.hipMalloc(ctypes.byref(dp), 4 * n), "hipMalloc")
    try:
        _ok(hip.hipMemcpy(dc, codes.ctypes.data, 4 * n, H2D), "hipMemcpy")
        junk = np.full(n, 0xDEADBEEF, np.uint32)  # the output starts as garbage
        _ok(hip.hipMemcpy(dp, junk.ctypes.data, 4 * n, H2D), "hipMemcpy")
        rc = galaxy_hip.lib().gs_radix_sort_indices_noalloc(dc, n, dp, stream)
        assert rc == galaxy_hip.GS_OK
        _ok(hip.hipStreamSynchronize(stream), "hipStreamSynchronize")
        out = np.empty(n, np.uint32)
        _ok(hip.hipMemcpy(out.ctypes.data, dp, 4 * n, D2H), "hipMemcpy")
        return out
    finally:
        hip.hipFree(dc)
        hip.hipFree(dp)


def test_noalloc_sort_equals_handle_sort():
    hip = _hip()
    rng = np.random.default_rng(17)
    big = rng.integers(0, 1 << 30, 1_000_000, dtype=np.uint32)
    big[:300_000] = big[300_000:600_000]  # heavy ties: stability decides
    from galaxy_hip import ic
    data = ic.make("collision_4m", n=300_001)
    small, _ = galaxy_hip.compute_morton_codes(data.reshape(-1), 300_001, 8)
    st = ctypes.c_void_p()
    _ok(hip.hipStreamCreate(ctypes.byref(st)), "hipStreamCreate")
    try:
        for codes, stream, label in ((big, st, "1,000,000 keys, own stream"),
                                     (small, st, "300,001 Morton codes, own stream"),
                                     (big[:65_537], None, "65,537 keys, null stream")):
            banner(f"noalloc sort {label}")
            want = galaxy_hip.radix_sort_indices(codes)
            got = _noalloc_sort(hip, np.ascontiguousarray(codes), stream)
            assert np.array_equal(got, want), label
            assert np.array_equal(want, np.argsort(codes, kind="stable").astype(np.uint32))
    finally:
        hip.hipStreamDestroy(st)


def test_noalloc_sort_refuses_sizes_beyond_its_scratch():
    hip = _hip()
    dc = ctypes.c_void_p()
    _ok(hip.hipMalloc(ctypes.byref(dc), 16), "hipMalloc")
    try:
        L = galaxy_hip.lib()
        assert L.gs_radix_sort_indices_noalloc(dc, (1 << 20) + 1, dc, None) == galaxy_hip.GS_E_INVALID
        assert L.gs_radix_sort_indices_noalloc(None, 4, dc, None) == galaxy_hip.GS_E_INVALID
        assert L.gs_radix_sort_indices_noalloc(dc, 0, dc, None) == galaxy_hip.GS_OK
        _ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    finally:
        hip.hipFree(dc)
