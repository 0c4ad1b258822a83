"""The allocation-free batch entry point (sarbp_batch_focus_noalloc) through the C ABI (GPU).

Its requests' pixel arrays and outputs are device buffers (allocated here by the test itself, with
the HIP runtime the library uses); its work list lives in the library's fixed device scratch.  Two
calls in a row, the second with fewer and smaller requests, each equal bit for bit to the same
requests focused by sarbp_batch_create + sarbp_batch_run + sarbp_batch_download; another work
mapping gives the same bits; a request list beyond the scratch is refused before any device work.
"""
import ctypes

import numpy as np
import pytest

from sarbp import _lib, synth
from sarbp.batch import Batch, patch_requests
from sarbp.core import Dataset

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.load()  # the library's HIP runtime is loaded first; the same one is reused here
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


class _DeviceRequests:
    """Device copies of requests' pixel arrays plus output buffers (the test's own allocations)."""

    def __init__(self, hip, reqs):
        self.hip, self.ptrs = hip, []
        self.cs = (_lib.Request * len(reqs))()
        self.outs = (ctypes.c_void_p * len(reqs))()
        self.cnts = (ctypes.c_void_p * len(reqs))()
        for k, r in enumerate(reqs):
            dev = [self._put(a) for a in (r.x, r.y, r.z, r.mask, r.index, r.sint)]
            self.cs[k] = _lib.Request(*dev, r.n, int(r.dimx), 1 if r.precision == "split" else 0, int(r.tile_x))
            self.outs[k] = self._alloc(16 * r.n)
            self.cnts[k] = self._alloc(4 * r.n)

    def _alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(1, nbytes)) == 0
        self.ptrs.append(p.value)
        return p.value

    def _put(self, a):
        a = np.ascontiguousarray(a)
        p = self._alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def result(self, k, n):
        out, cnt = np.empty(n, np.complex128), np.empty(n, np.int32)
        assert self.hip.hipMemcpy(out.ctypes.data, self.outs[k], 16 * n, D2H) == 0
        assert self.hip.hipMemcpy(cnt.ctypes.data, self.cnts[k], 4 * n, D2H) == 0
        return out, cnt

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(ds, dev, nreq, tile_px=0, chunk=0):
    return _lib.load().sarbp_batch_focus_noalloc(ds.handle, dev.cs, nreq, dev.outs, dev.cnts, tile_px, chunk)


def test_noalloc_equals_batch_two_calls(gpu):
    hip = _hip()
    sc = synth.full_scene(seed=6, nx=512, ny=250, naz=5121, ntarget=100)
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    calls = ((6, (64, 200), (16, 80), 500, 4000), (3, (16, 48), (8, 24), 300, 1200))  # the second is smaller
    for call, (nreq, w, h, s0, s1) in enumerate(calls):
        reqs = patch_requests(sc, nreq, seed=20 + call, w_range=w, h_range=h, sint_range=(s0, s1))
        print("case no-alloc call", call, [(r.dimx, r.n // r.dimx, r.precision) for r in reqs], flush=True)
        b = Batch(ds, reqs)
        b.run(64, 32)
        want = [b.result(k, with_counts=True) for k in range(nreq)]
        b.close()
        dev = _DeviceRequests(hip, reqs)
        try:
            for tp, ch in ((0, 0), (128, 64)):
                assert _noalloc(ds, dev, nreq, tp, ch) == _lib.OK, _lib.load().sarbp_last_error()
                for k, r in enumerate(reqs):
                    got, cnt = dev.result(k, r.n)
                    assert np.array_equal(cnt, want[k][1]) and np.array_equal(got, want[k][0]), (call, tp, k)
        finally:
            dev.free()
    ds.close()


def test_noalloc_refuses_work_beyond_scratch(gpu):
    hip = _hip()
    sc = synth.tiny_scene(seed=2)
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    reqs = patch_requests(sc, 1, seed=1, w_range=(8, 8), h_range=(4, 4))
    dev = _DeviceRequests(hip, reqs)
    try:
        # a request claiming 64 M pixels: its 1 M work items exceed the scratch; refused on the host,
        # before anything reads its buffers
        dev.cs[0].n, dev.cs[0].dimx = 1 << 26, 8192
        assert _noalloc(ds, dev, 1) == _lib.EINVAL
        assert b"scratch" in _lib.load().sarbp_last_error()
        dev.cs[0].n, dev.cs[0].dimx = reqs[0].n, reqs[0].dimx
        assert _noalloc(ds, dev, 1, 96, 32) == _lib.EINVAL  # not a valid work mapping
        assert _noalloc(ds, dev, 1) == _lib.OK
    finally:
        dev.free()
    ds.close()
