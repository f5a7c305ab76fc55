"""The allocation-free scan (gpsm_scan_u32_noalloc) against the allocating
one (gpsm_scan_u32) through the C ABI: device buffers from torch, two calls
in a row (the second smaller, so stale block sums of the first would show),
then an aliased in-place call, the size limit and the empty input."""
import ctypes as C

import numpy as np
import pytest

from gpsm_amd import _lib

pytestmark = pytest.mark.gpu


def noalloc_scan(ctx, d_in, d_out, n):
    tot = C.c_uint64(0)
    rc = _lib.lib().gpsm_scan_u32_noalloc(ctx.h, C.c_void_p(d_in), C.c_void_p(d_out), n, C.byref(tot))
    return rc, tot.value


def test_noalloc_scan_equals_allocating_scan(ctx):
    import torch
    rng = np.random.default_rng(5)
    for n in (3 * 4096 * 1000 + 17, 70_001):  # second call smaller than the first
        a = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        want, want_tot = ctx.scan(a)
        t_in = torch.from_numpy(a.view(np.int32)).to("cuda")
        t_out = torch.full_like(t_in, -1)
        torch.cuda.synchronize()
        rc, tot = noalloc_scan(ctx, t_in.data_ptr(), t_out.data_ptr(), n)
        assert rc == 0
        got = t_out.cpu().numpy().view(np.uint32)
        assert np.array_equal(got, want) and tot == want_tot, n
    # in place (d_out aliases d_in)
    a = rng.integers(0, 1000, size=123_457, dtype=np.uint32)
    want, want_tot = ctx.scan(a)
    t = torch.from_numpy(a.view(np.int32)).to("cuda")
    torch.cuda.synchronize()
    rc, tot = noalloc_scan(ctx, t.data_ptr(), t.data_ptr(), len(a))
    assert rc == 0 and tot == want_tot
    assert np.array_equal(t.cpu().numpy().view(np.uint32), want)


def test_noalloc_scan_limits(ctx):
    # beyond the fixed scratch: refused before any launch (no buffer is touched)
    rc, _ = noalloc_scan(ctx, 16, 16, _lib.SCAN_NOALLOC_MAX + 1)
    assert rc == _lib.ESCRATCH
    rc, tot = noalloc_scan(ctx, 0, 0, 0)
    assert rc == 0 and tot == 0
