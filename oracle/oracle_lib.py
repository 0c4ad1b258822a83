"""ctypes loader for the C oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use this module, as
the checker or as the timed CPU baseline.  Restates sar_focuser_puntual.py:21-155 and 259-268.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libsarbp_oracle.so")
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_p = ctypes.c_void_p


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.sarbp_oracle_focus.argtypes = [_p] * 6 + [_i64] + [_p] * 4 + [_i64, _i64, _d, _d, _d, _p, _p,
                                                                     ctypes.c_int]
        L.sarbp_oracle_focus.restype = None
        L.sarbp_oracle_point.argtypes = [_d, _d, _d, ctypes.c_uint8, _i32, _i32, _p, _p, _p, _p, _i64,
                                         _i64, _d, _d, _d, ctypes.POINTER(_d), ctypes.POINTER(_d)]
        L.sarbp_oracle_point.restype = _i32
        L.sarbp_oracle_point_ex.argtypes = [_d, _d, _d, ctypes.c_uint8, _i32, _i32, _p, _p, _p, _p, _i64,
                                            _i64, _d, _d, _d, _i32, ctypes.POINTER(_d), ctypes.POINTER(_d)]
        L.sarbp_oracle_point_ex.restype = _i32
        L.sarbp_oracle_trace.argtypes = [_d, _d, _d, _i32, _i32, _p, _p, _p, _i64, _i64, _d, _d, _p,
                                         _p, _p, _i64]
        L.sarbp_oracle_trace.restype = _i64
        L.sarbp_oracle_pairs.argtypes = [_p, _p, _p, _p, _i64, _p, _p, _p, _i64, _d, _d, _d] + [_p] * 6 + \
            [ctypes.c_int]
        L.sarbp_oracle_pairs.restype = None
        L.sarbp_oracle_count_pp.argtypes = [_p, _p, _p, _i64, _i64]
        L.sarbp_oracle_count_pp.restype = _i64
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _raw_f32(raw):
    """complex64 (Nra, Naz) -> interleaved float32 buffer in column-major order."""
    raw = np.asarray(raw)
    if raw.dtype != np.complex64:
        raw = raw.astype(np.complex64)
    return np.asfortranarray(raw)


def focus_flat(x, y, z, mask, index, sint, tx, ty, tz, raw, near, dra, lam, nthreads=0):
    """Flat (F-order) pixels -> (complex128[n], int32 n[n])."""
    x, y, z = _c(x, np.float64), _c(y, np.float64), _c(z, np.float64)
    mask, index, sint = _c(mask, np.uint8), _c(index, np.int32), _c(sint, np.int32)
    tx, ty, tz = _c(tx, np.float64), _c(ty, np.float64), _c(tz, np.float64)
    rawf = _raw_f32(raw)
    Nra, Naz = rawf.shape
    n = x.size
    out = np.zeros(n, np.complex128)
    cnt = np.zeros(n, np.int32)
    lib().sarbp_oracle_focus(x.ctypes.data, y.ctypes.data, z.ctypes.data, mask.ctypes.data,
                             index.ctypes.data, sint.ctypes.data, n, tx.ctypes.data, ty.ctypes.data,
                             tz.ctypes.data, rawf.ctypes.data, Nra, Naz, float(near), float(dra),
                             float(lam), out.ctypes.data, cnt.ctypes.data, int(nthreads))
    return out, cnt


def point(x, y, z, mask, index, sint, tx, ty, tz, raw, near, dra, lam, ind_line_ra_min=0):
    """process_point (sar_focuser_puntual.py:21-155) for one pixel, with its ind_line_ra_min
    argument: (complex128, contributor count)."""
    tx, ty, tz = _c(tx, np.float64), _c(ty, np.float64), _c(tz, np.float64)
    rawf = _raw_f32(raw)
    Nra, Naz = rawf.shape
    re, im = _d(), _d()
    n = lib().sarbp_oracle_point_ex(float(x), float(y), float(z), int(mask), int(index), int(sint), tx.ctypes.data,
                                    ty.ctypes.data, tz.ctypes.data, rawf.ctypes.data, Nra, Naz, float(near),
                                    float(dra), float(lam), int(ind_line_ra_min), ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value), int(n)


def focus(xxg1, yyg1, zzg1, mask, index, wgsx, wgsy, wgsz, raw, near, dra, lam, Sint, nthreads=0):
    """FocalizzatoreBpPx argument order; 2-D inputs -> (complex128 (dimx,dimy), n (dimx,dimy))."""
    shp = np.shape(xxg1)
    f = lambda a: np.asarray(a).flatten(order="F")
    out, cnt = focus_flat(f(xxg1), f(yyg1), f(zzg1), f(mask), f(index), f(Sint), wgsx, wgsy, wgsz,
                          raw, near, dra, lam, nthreads)
    return out.reshape(shp, order="F"), cnt.reshape(shp, order="F")


def trace(x, y, z, index, sint, tx, ty, tz, Nra, Naz, near, dra, cap=1 << 16):
    tx, ty, tz = _c(tx, np.float64), _c(ty, np.float64), _c(tz, np.float64)
    pul = np.zeros(cap, np.int64)
    bins = np.zeros(cap, np.int64)
    rng = np.zeros(cap, np.float64)
    k = lib().sarbp_oracle_trace(x, y, z, index, sint, tx.ctypes.data, ty.ctypes.data, tz.ctypes.data,
                                 Nra, Naz, near, dra, pul.ctypes.data, bins.ctypes.data,
                                 rng.ctypes.data, cap)
    return pul[:k], bins[:k], rng[:k]


def count_pp(mask, index, sint, Naz) -> int:
    mask, index, sint = _c(mask, np.uint8).ravel(), _c(index, np.int32).ravel(), _c(sint, np.int32).ravel()
    return int(lib().sarbp_oracle_count_pp(mask.ctypes.data, index.ctypes.data, sint.ctypes.data,
                                           mask.size, Naz))


def pairs(x, y, z, pulse, tx, ty, tz, Nra, near, dra, lam, nthreads=0):
    """Per-(pixel, pulse) R2, R, bin (-1 outside [0, Nra-1)), phase, sin, cos of
    sar_focuser_puntual.py:82-121 for equal-length arrays x, y, z (float64) and pulse (int32)."""
    x, y, z = _c(x, np.float64), _c(y, np.float64), _c(z, np.float64)
    pulse = _c(pulse, np.int32)
    tx, ty, tz = _c(tx, np.float64), _c(ty, np.float64), _c(tz, np.float64)
    n = x.size
    out = {k: np.empty(n, np.float64) for k in ("r2", "r", "ph", "sin", "cos")}
    out["bin"] = np.empty(n, np.int32)
    lib().sarbp_oracle_pairs(x.ctypes.data, y.ctypes.data, z.ctypes.data, pulse.ctypes.data, n, tx.ctypes.data,
                             ty.ctypes.data, tz.ctypes.data, int(Nra), float(near), float(dra), float(lam),
                             out["r2"].ctypes.data, out["r"].ctypes.data, out["bin"].ctypes.data,
                             out["ph"].ctypes.data, out["sin"].ctypes.data, out["cos"].ctypes.data, int(nthreads))
    return out
