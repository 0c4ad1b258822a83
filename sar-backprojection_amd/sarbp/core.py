"""Python handles over the C ABI: device-resident datasets (RawData + trajectory) and pixel sets."""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import PREC_F64, PREC_SPLIT, check

PRECISIONS = {"f64": PREC_F64, "fp64": PREC_F64, "split": PREC_SPLIT, "split-fp32": PREC_SPLIT}


def precision_code(precision) -> int:
    if isinstance(precision, str):
        try:
            return PRECISIONS[precision.lower()]
        except KeyError:
            raise ValueError(f"unknown precision {precision!r}; use 'f64' or 'split'") from None
    if precision not in (PREC_F64, PREC_SPLIT):
        raise ValueError(f"unknown precision {precision!r}")
    return int(precision)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def raw_colmajor(raw) -> np.ndarray:
    """RawData (Nra, Naz) as a complex64 array in column-major order (no copy if it already is)."""
    raw = np.asarray(raw)
    if raw.ndim != 2:
        raise ValueError("Raw_Data must be a 2D array")
    if raw.dtype != np.complex64:
        raw = raw.astype(np.complex64)
    return np.asfortranarray(raw)


class Dataset:
    _close_order = 2  # servers, then batches/scenes, then datasets
    """RawData, trajectory and the three scalars, resident in HBM (sarbp_dataset_create)."""

    def __init__(self, raw, tx, ty, tz, near_range, dra, lambda_val):
        L = _lib.load()
        _lib.require_device()
        rawf = raw_colmajor(raw)
        self.nra, self.naz = (int(v) for v in rawf.shape)
        tx, ty, tz = _c(tx, np.float64), _c(ty, np.float64), _c(tz, np.float64)
        if not (tx.size == ty.size == tz.size == self.naz):
            raise ValueError(f"trajectory length ({tx.size}, {ty.size}, {tz.size}) must equal Naz={self.naz}")
        self.near_range, self.dra, self.lambda_val = float(near_range), float(dra), float(lambda_val)
        h = ctypes.c_void_p()
        check(L.sarbp_dataset_create(rawf.ctypes.data, self.nra, self.naz, tx.ctypes.data, ty.ctypes.data,
                                     tz.ctypes.data, self.near_range, self.dra, self.lambda_val,
                                     ctypes.byref(h)), "sarbp_dataset_create")
        self._h = h
        _lib.track(self)

    @classmethod
    def streamed(cls, fill, nra, naz, tx, ty, tz, near_range, dra, lambda_val):
        """Upload RawData in pulse blocks: ``fill(a0, a1) -> complex64 (nra, a1-a0)`` array."""
        L = _lib.load()
        _lib.require_device()
        self = cls.__new__(cls)
        self.nra, self.naz = int(nra), int(naz)
        tx, ty, tz = _c(tx, np.float64), _c(ty, np.float64), _c(tz, np.float64)
        if not (tx.size == ty.size == tz.size == self.naz):
            raise ValueError("trajectory length must equal Naz")
        self.near_range, self.dra, self.lambda_val = float(near_range), float(dra), float(lambda_val)
        errors = []

        def _fill(user, a0, a1, dst):
            try:
                blk = raw_colmajor(fill(int(a0), int(a1)))
                if blk.shape != (self.nra, a1 - a0):
                    raise ValueError(f"fill returned shape {blk.shape}, expected {(self.nra, a1 - a0)}")
                ctypes.memmove(dst, blk.ctypes.data, blk.nbytes)
                return 0
            except Exception as e:  # reported after the C call returns
                errors.append(e)
                return 1

        cb = _lib.FILL_FN(_fill)
        h = ctypes.c_void_p()
        st = L.sarbp_dataset_create_streamed(cb, None, self.nra, self.naz, tx.ctypes.data, ty.ctypes.data,
                                             tz.ctypes.data, self.near_range, self.dra, self.lambda_val,
                                             ctypes.byref(h))
        if errors:
            raise errors[0]
        check(st, "sarbp_dataset_create_streamed")
        self._h = h
        _lib.track(self)
        return self

    @property
    def handle(self):
        return self._h

    @property
    def device_bytes(self) -> int:
        v = ctypes.c_int64(0)
        check(_lib.load().sarbp_dataset_bytes(self._h, ctypes.byref(v)))
        return int(v.value)

    def _adopt(self, child):
        """Scenes and batches of this dataset: closed with it (the library refuses to destroy a
        dataset whose scenes or batches still exist)."""
        if not hasattr(self, "_children"):
            self._children = weakref.WeakSet()
        self._children.add(child)

    def close(self):
        """Release the device memory, after closing the dataset's scenes and batches.  Raises
        SarbpError (ESTATE) while a server still has requests against this dataset that are not done
        or failed; the dataset then stays valid."""
        h = getattr(self, "_h", None)
        if h:
            for c in list(getattr(self, "_children", ())):
                c.close()
            check(_lib.load().sarbp_dataset_destroy(h), "sarbp_dataset_destroy")
            self._h = None

    def __del__(self):
        try:  # at interpreter exit the module's globals may already be gone
            if _lib.alive():
                self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def scene(self, x, y, z, mask, index, sint, dimx=0) -> "Scene":
        return Scene(self, x, y, z, mask, index, sint, dimx)

    def focus_flat(self, x, y, z, mask, index, sint, precision="f64", with_counts=False, dimx=0):
        """Focus F-order flattened pixels of a grid whose fast dimension is ``dimx`` (0: a plain
        list); returns complex128[n] (and int32 counts)."""
        x, y, z = _c(x, np.float64).ravel(), _c(y, np.float64).ravel(), _c(z, np.float64).ravel()
        mask, index, sint = _c(mask, np.uint8).ravel(), _c(index, np.int32).ravel(), _c(sint, np.int32).ravel()
        n = x.size
        if not (y.size == z.size == mask.size == index.size == sint.size == n):
            raise ValueError("pixel arrays must all have the same size")
        out = np.zeros(n, np.complex128)
        cnt = np.zeros(n, np.int32)
        check(_lib.load().sarbp_focus(self._h, x.ctypes.data, y.ctypes.data, z.ctypes.data, mask.ctypes.data,
                                      index.ctypes.data, sint.ctypes.data, n, int(dimx), precision_code(precision),
                                      out.ctypes.data, cnt.ctypes.data), "sarbp_focus")
        return (out, cnt) if with_counts else out

    def split_error_estimate(self, x, y, z, mask, index, sint, dimx=0, tile_x=0) -> float:
        """A-priori RMS phase error (rad) of the split path for these pixels (F-order grid of fast
        dimension dimx, 0 = plain list) against FP64, over pixels above 1e-5 x max amplitude; the
        estimate the server compares with its fallback limit (1e9: rows too wide for the split form)."""
        c = lambda a, dt: np.ascontiguousarray(np.asarray(a, dtype=dt).ravel())  # noqa: E731
        x, y, z = c(x, np.float64), c(y, np.float64), c(z, np.float64)
        mask, index, sint = c(mask, np.uint8), c(index, np.int32), c(sint, np.int32)
        n = x.size
        if not all(a.size == n for a in (y, z, mask, index, sint)):
            raise ValueError("pixel arrays must all have the same size")
        req = _lib.Request(x.ctypes.data, y.ctypes.data, z.ctypes.data, mask.ctypes.data, index.ctypes.data,
                           sint.ctypes.data, n, int(dimx), PREC_SPLIT, int(tile_x))
        est = ctypes.c_double()
        check(_lib.load().sarbp_split_error_estimate(self.handle, ctypes.byref(req), ctypes.byref(est)),
              "sarbp_split_error_estimate")
        return float(est.value)

    def eval_pairs(self, x, y, z, pulse) -> dict:
        """The FP64 path's per-(pixel, pulse) arithmetic on the device (sarbp_eval_pairs): R2, R,
        bin (-1 outside [0, Nra-1)), phase, sin, cos for equal-length x, y, z and pulse arrays."""
        x, y, z = _c(x, np.float64).ravel(), _c(y, np.float64).ravel(), _c(z, np.float64).ravel()
        pulse = _c(pulse, np.int32).ravel()
        n = x.size
        if not (y.size == z.size == pulse.size == n):
            raise ValueError("pair arrays must all have the same size")
        out = {k: np.empty(n, np.float64) for k in ("r2", "r", "ph", "sin", "cos")}
        out["bin"] = np.empty(n, np.int32)
        check(_lib.load().sarbp_eval_pairs(self._h, x.ctypes.data, y.ctypes.data, z.ctypes.data, pulse.ctypes.data, n,
                                           out["r2"].ctypes.data, out["r"].ctypes.data, out["bin"].ctypes.data,
                                           out["ph"].ctypes.data, out["sin"].ctypes.data, out["cos"].ctypes.data),
              "sarbp_eval_pairs")
        return out


class Scene:
    _close_order = 1  # servers, then batches/scenes, then datasets
    """A device-resident pixel set (sarbp_scene_create) focused repeatedly against one dataset."""

    def __init__(self, ds: Dataset, x, y, z, mask, index, sint, dimx=0):
        self.ds = ds
        self.dimx = int(dimx)
        self.x, self.y, self.z = _c(x, np.float64).ravel(), _c(y, np.float64).ravel(), _c(z, np.float64).ravel()
        self.mask = _c(mask, np.uint8).ravel()
        self.index = _c(index, np.int32).ravel()
        self.sint = _c(sint, np.int32).ravel()
        self.n = self.x.size
        if not (self.y.size == self.z.size == self.mask.size == self.index.size == self.sint.size == self.n):
            raise ValueError("pixel arrays must all have the same size")
        h = ctypes.c_void_p()
        check(_lib.load().sarbp_scene_create(ds.handle, self.x.ctypes.data, self.y.ctypes.data, self.z.ctypes.data,
                                             self.mask.ctypes.data, self.index.ctypes.data, self.sint.ctypes.data,
                                             self.n, self.dimx, ctypes.byref(h)), "sarbp_scene_create")
        self._h = h
        _lib.track(self)
        ds._adopt(self)

    def set_tiling(self, precision="f64", kernel=1, tile_x=16, accum=0):
        """Work mapping for one precision (sarbp_scene_set_tiling)."""
        check(_lib.load().sarbp_scene_set_tiling(self._h, precision_code(precision), int(kernel), int(tile_x),
                                                  int(accum)), "sarbp_scene_set_tiling")

    def set_bin_min(self, bin_min: int):
        """Lowest contributing range bin, process_point's ind_line_ra_min (sarbp_scene_set_bin_min)."""
        check(_lib.load().sarbp_scene_set_bin_min(self._h, int(bin_min)), "sarbp_scene_set_bin_min")

    def count_pp(self) -> int:
        return count_pp(self.mask, self.index, self.sint, self.ds.naz)

    def focus(self, precision="f64") -> float:
        """Run one focusing pass; returns the kernel time in ms (HIP events)."""
        ms = ctypes.c_float(0.0)
        check(_lib.load().sarbp_scene_focus(self.ds.handle, self._h, precision_code(precision), ctypes.byref(ms)),
              "sarbp_scene_focus")
        return float(ms.value)

    def result(self, with_counts=False):
        out = np.zeros(self.n, np.complex128)
        cnt = np.zeros(self.n, np.int32) if with_counts else None
        check(_lib.load().sarbp_scene_download(self._h, out.ctypes.data, cnt.ctypes.data if with_counts else None),
              "sarbp_scene_download")
        return (out, cnt) if with_counts else out

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().sarbp_scene_destroy(h)
            self._h = None

    def __del__(self):
        try:  # at interpreter exit the module's globals may already be gone
            if _lib.alive():
                self.close()
        except Exception:
            pass


def count_pp(mask, index, sint, naz) -> int:
    """Pixel-pulse pairs a focusing pass evaluates (sarbp_count_pp)."""
    mask, index, sint = _c(mask, np.uint8).ravel(), _c(index, np.int32).ravel(), _c(sint, np.int32).ravel()
    return int(_lib.load().sarbp_count_pp(mask.ctypes.data, index.ctypes.data, sint.ctypes.data, mask.size, int(naz)))
