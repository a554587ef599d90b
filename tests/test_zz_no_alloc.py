"""fluid_scan_u32_noalloc (include/fluid_sim.h) against fluid_scan_u32, through
the C ABI: calls in a row on device buffers (the second
with a smaller input, so a stale look-back word or ticket of the first would
show), each bit-exact with the allocating entry point and with numpy's
exclusive cumsum; the caller's counts unchanged; sizes past the scratch's
2^24 cells refused with a message before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

import fluid_amd as fa

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libfluid_sim.so itself runs on (its soname resolves to
    the copy already loaded with the library), for device buffers and
    streams; torch ships a runtime of its own, which cannot share the device
    once this one holds it."""
    fa.lib()
    h = C.CDLL("libamdhip64.so.7")
    for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize"):
        getattr(h, name).restype = C.c_int
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    return h


def _noalloc(counts, own_stream=False):
    h = _hip()
    d_counts, d_starts, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    n = counts.size
    assert h.hipMalloc(C.byref(d_counts), n * 4) == 0 and h.hipMalloc(C.byref(d_starts), (n + 1) * 4) == 0
    try:
        assert h.hipMemcpy(d_counts, counts.ctypes.data, n * 4, 1) == 0  # host to device
        fill = np.full(n + 1, 0xFFFFFFFF, np.uint32)
        assert h.hipMemcpy(d_starts, fill.ctypes.data, (n + 1) * 4, 1) == 0
        if own_stream:
            assert h.hipStreamCreate(C.byref(st)) == 0
        fa._lib.check(fa.lib().fluid_scan_u32_noalloc(d_counts, n, d_starts, st), "fluid_scan_u32_noalloc")
        assert h.hipStreamSynchronize(st) == 0
        out = np.empty(n + 1, np.uint32)
        after = np.empty(n, np.uint32)
        assert h.hipMemcpy(out.ctypes.data, d_starts, (n + 1) * 4, 2) == 0  # device to host
        assert h.hipMemcpy(after.ctypes.data, d_counts, n * 4, 2) == 0
        return out, after
    finally:
        if st.value:
            h.hipStreamDestroy(st)
        h.hipFree(d_counts)
        h.hipFree(d_starts)


def test_noalloc_scan_matches_the_allocating_scan_twice():
    for cells, seed in ((3_000_017, 1), (100_003, 2), (2048, 3), (1, 4)):
        rng = np.random.default_rng(seed)
        counts = rng.integers(0, 40, size=cells, dtype=np.uint32)
        counts[rng.integers(0, cells, size=max(1, cells // 10))] = 0
        print(f"noalloc scan: {cells} cells", flush=True)
        got, after = _noalloc(counts)
        ref = fa.scan_u32(counts)
        assert np.array_equal(got, ref), cells
        expect = np.zeros(cells + 1, np.uint64)
        expect[1:] = np.cumsum(counts, dtype=np.uint64)
        assert np.array_equal(got.astype(np.uint64), expect), cells
        assert np.array_equal(after, counts), cells


def test_noalloc_scan_on_its_own_stream():
    counts = np.random.default_rng(5).integers(0, 9, size=777_777, dtype=np.uint32)
    got, _ = _noalloc(counts, own_stream=True)
    assert np.array_equal(got, fa.scan_u32(counts))


def test_noalloc_scan_refuses_past_its_scratch():
    L = fa.lib()
    rc = L.fluid_scan_u32_noalloc(None, (1 << 24) + 1, None, None)
    assert rc != 0
    rc = L.fluid_scan_u32_noalloc(1, (1 << 24) + 1, 1, None)  # never dereferenced: refused first
    assert rc == -1 and b"at most 16777216 cells" in L.fluid_sim_last_error()
