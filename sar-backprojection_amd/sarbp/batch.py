"""Many requests (pixel sets of uneven size, aperture length and precision) focused in one
persistent launch against one device-resident dataset (sarbp_batch_* in include/sarbp.h)."""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib
from ._lib import check
from .core import Dataset, count_pp, precision_code


@dataclasses.dataclass
class Request:
    """One sub-scene: F-order flattened pixel arrays of a dimx x (n/dimx) grid."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    mask: np.ndarray
    index: np.ndarray
    sint: np.ndarray
    dimx: int = 0
    precision: str = "f64"
    tile_x: int = 0

    def __post_init__(self):
        c = lambda a, dt: np.ascontiguousarray(np.asarray(a, dtype=dt).ravel())  # noqa: E731
        self.x, self.y, self.z = c(self.x, np.float64), c(self.y, np.float64), c(self.z, np.float64)
        self.mask, self.index, self.sint = c(self.mask, np.uint8), c(self.index, np.int32), c(self.sint, np.int32)
        n = self.x.size
        if not all(a.size == n for a in (self.y, self.z, self.mask, self.index, self.sint)):
            raise ValueError("request pixel arrays must all have the same size")
        precision_code(self.precision)

    @property
    def n(self) -> int:
        return int(self.x.size)

    def c_struct(self) -> _lib.Request:
        return _lib.Request(self.x.ctypes.data, self.y.ctypes.data, self.z.ctypes.data, self.mask.ctypes.data,
                            self.index.ctypes.data, self.sint.ctypes.data, self.n, int(self.dimx),
                            precision_code(self.precision), int(self.tile_x))

    def count_pp(self, naz: int) -> int:
        return count_pp(self.mask, self.index, self.sint, naz)


class Batch:
    _close_order = 1  # servers, then batches/scenes, then datasets
    def __init__(self, ds: Dataset, requests):
        self.ds = ds
        self.requests = list(requests)
        arr = (_lib.Request * max(1, len(self.requests)))(*[r.c_struct() for r in self.requests])
        h = ctypes.c_void_p()
        check(_lib.load().sarbp_batch_create(ds.handle, arr, len(self.requests), ctypes.byref(h)), "sarbp_batch_create")
        self._h = h
        _lib.track(self)
        ds._adopt(self)

    def run(self, tile_px: int = 0, pulse_chunk: int = 0) -> float:
        """One persistent launch over every request; returns the kernel time in ms."""
        ms = ctypes.c_float(0)
        check(_lib.load().sarbp_batch_run(self._h, int(tile_px), int(pulse_chunk), ctypes.byref(ms)),
              "sarbp_batch_run")
        return float(ms.value)

    def autotune(self, reps: int = 2):
        """Returns (tile_px, pulse_chunk, ms) of the fastest combination; later runs use it."""
        t, c, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
        check(_lib.load().sarbp_batch_autotune(self._h, int(reps), ctypes.byref(t), ctypes.byref(c), ctypes.byref(ms)),
              "sarbp_batch_autotune")
        return int(t.value), int(c.value), float(ms.value)

    def result(self, k: int, with_counts=False):
        r = self.requests[k]
        out = np.zeros(r.n, np.complex128)
        cnt = np.zeros(r.n, np.int32) if with_counts else None
        check(_lib.load().sarbp_batch_download(self._h, int(k), out.ctypes.data,
                                               cnt.ctypes.data if with_counts else None), "sarbp_batch_download")
        return (out, cnt) if with_counts else out

    def count_pp(self) -> int:
        return sum(r.count_pp(self.ds.naz) for r in self.requests)

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().sarbp_batch_destroy(h)
            self._h = None

    def __del__(self):
        try:  # at interpreter exit the module's globals may already be gone
            if _lib.alive():
                self.close()
        except Exception:
            pass


def patch_requests(scene, nreq: int, seed: int = 0, w_range=(128, 1024), h_range=(32, 250),
                   sint_range=(500, 4000), precisions=("split", "f64")):
    """Rectangular sub-scenes of a synthetic scene: width/height uniform in the given ranges
    (clipped to the grid), one Sint per request uniform in ``sint_range``, precision alternating."""
    rng = np.random.default_rng(seed)
    nx, ny = scene.shape
    reqs = []
    for k in range(nreq):
        w = int(min(nx, rng.integers(w_range[0], w_range[1] + 1)))
        h = int(min(ny, rng.integers(h_range[0], h_range[1] + 1)))
        x0 = int(rng.integers(0, nx - w + 1))
        y0 = int(rng.integers(0, ny - h + 1))
        s = int(rng.integers(sint_range[0], sint_range[1] + 1))
        sub = scene.sub(x0, x0 + w, y0, y0 + h)
        f = lambda a: np.asarray(a).flatten(order="F")  # noqa: E731
        reqs.append(Request(f(sub.x), f(sub.y), f(sub.z), f(sub.mask), f(sub.index), np.full(w * h, s, np.int32),
                            dimx=w, precision=precisions[k % len(precisions)]))
    return reqs
