"""Drop-in for the entry points of gpu_small_dataset.py, computed on the MI355X.

``FocalizzatoreBpCuPy``, ``FocalizzatoreBpCuPy_Batched`` and ``load_raw_data`` keep the
reference's names, argument order and return values (gpu_small_dataset.py:113, 229, 368).  The
focusing follows the Numba path's semantics (half-to-even range bins, last bin excluded), not the
CUDA kernel's (gpu_small_dataset.py:75, 78), so results equal ``FocalizzatoreBpPx``'s.
"""
from __future__ import annotations

import os

import numpy as np

from . import formats
from .core import Dataset
from .sar_focuser_puntual import _flatten_inputs


def _check_2d(Raw_Data, xxg1):
    if np.ndim(Raw_Data) != 2:
        raise ValueError("Raw_Data must be a 2D array")
    if np.ndim(xxg1) != 2:
        raise ValueError("xxg1 must be a 2D array")


def FocalizzatoreBpCuPy(xxg1, yyg1, zzg1, mask, index, wgsx, wgsy, wgsz, Raw_Data, near_range, dra, lambda_val,
                        Sint, precision="f64"):
    """All pixels in one pass (gpu_small_dataset.py:113-223); complex128 of shape xxg1.shape."""
    _check_2d(Raw_Data, xxg1)
    dimx, dimy = np.shape(xxg1)
    xf, yf, zf, mf, inf, sf = _flatten_inputs(xxg1, yyg1, zzg1, mask, index, Sint)
    with Dataset(Raw_Data, wgsx, wgsy, wgsz, near_range, dra, lambda_val) as ds:  # the arrays as they are now
        out = ds.focus_flat(xf, yf, zf, mf, inf, sf, precision, dimx=dimx)
    return np.reshape(out, (dimx, dimy), order="F")


def FocalizzatoreBpCuPy_Batched(xxg1, yyg1, zzg1, mask, index, wgsx, wgsy, wgsz, Raw_Data, near_range, dra,
                                lambda_val, Sint, batch_size=500000, precision="f64"):
    """Pixels in batches of ``batch_size`` (gpu_small_dataset.py:229-362), RawData resident once.
    Batches are rounded up to whole grid columns so tiles stay two-dimensional."""
    _check_2d(Raw_Data, xxg1)
    if int(batch_size) <= 0:
        raise ValueError("batch_size must be positive")
    dimx, dimy = np.shape(xxg1)
    dim = dimx * dimy
    xf, yf, zf, mf, inf, sf = _flatten_inputs(xxg1, yyg1, zzg1, mask, index, Sint)
    step = -(-int(batch_size) // dimx) * dimx
    out = np.zeros(dim, np.complex128)
    with Dataset(Raw_Data, wgsx, wgsy, wgsz, near_range, dra, lambda_val) as ds:  # RawData resident once
        for i in range(0, dim, step):
            j = min(dim, i + step)
            out[i:j] = ds.focus_flat(xf[i:j], yf[i:j], zf[i:j], mf[i:j], inf[i:j], sf[i:j], precision, dimx=dimx)
    return np.reshape(out, (dimx, dimy), order="F")


def load_raw_data(data_dir, filename, expected_Nra, expected_Naz):
    """Tolerant RawData loader (gpu_small_dataset.py:368-451): returns (RawData, Nra, Naz)."""
    return formats.load_raw_tolerant(os.path.join(data_dir, filename), int(expected_Nra), int(expected_Naz))
