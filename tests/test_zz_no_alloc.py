"""The allocation-free scan (gpsm_scan_u32_noalloc) against the allocating
one (gpsm_scan_u32) through the C ABI: two calls in a row (the second
smaller, so stale block sums of the first would show), an aliased in-place
call, the size limit and the empty input.  The device buffers are the
caller's: allocated here with hipMalloc of the HIP runtime libgpsm.so is
linked against (torch ships a runtime of its own, whose allocations this
process's library runtime does not share)."""
import ctypes as C

import numpy as np
import pytest

from gpsm_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    _lib.lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            p = line.split()[-1]
            if "libamdhip64.so" in p and "/torch/" not in p:
                path = p
                break
    assert path, "libgpsm.so's HIP runtime is not loaded"
    rt = C.CDLL(path)
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


class DevArray:
    def __init__(self, rt, n):
        self.rt, self.p = rt, C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(4 * n, 4)) == 0
        self.n = n

    def put(self, a):
        assert self.rt.hipMemcpy(self.p, a.ctypes.data, 4 * len(a), H2D) == 0

    def get(self):
        out = np.empty(self.n, dtype=np.uint32)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, 4 * self.n, D2H) == 0
        return out

    def free(self):
        self.rt.hipFree(self.p)


def noalloc_scan(ctx, d_in, d_out, n):
    tot = C.c_uint64(0)
    rc = _lib.lib().gpsm_scan_u32_noalloc(ctx.h, d_in, d_out, n, C.byref(tot))
    return rc, tot.value


def test_noalloc_scan_equals_allocating_scan(ctx, hip):
    rng = np.random.default_rng(5)
    for n in (3 * 4096 * 1000 + 17, 70_001):  # second call smaller than the first
        a = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        want, want_tot = ctx.scan(a)
        din, dout = DevArray(hip, n), DevArray(hip, n)
        din.put(a)
        dout.put(np.full(n, 0xFFFFFFFF, dtype=np.uint32))
        rc, tot = noalloc_scan(ctx, din.p, dout.p, n)
        assert rc == 0
        assert np.array_equal(dout.get(), want) and tot == want_tot, n
        din.free()
        dout.free()
    # in place (d_out aliases d_in)
    a = rng.integers(0, 1000, size=123_457, dtype=np.uint32)
    want, want_tot = ctx.scan(a)
    d = DevArray(hip, len(a))
    d.put(a)
    rc, tot = noalloc_scan(ctx, d.p, d.p, len(a))
    assert rc == 0 and tot == want_tot
    assert np.array_equal(d.get(), want)
    d.free()


def test_noalloc_scan_limits(ctx):
    # beyond the fixed scratch: refused before any launch (no buffer is touched)
    rc, _ = noalloc_scan(ctx, C.c_void_p(16), C.c_void_p(16), _lib.SCAN_NOALLOC_MAX + 1)
    assert rc == _lib.ESCRATCH
    rc, tot = noalloc_scan(ctx, None, None, 0)
    assert rc == 0 and tot == 0
