"""NumPy restatement of the reference back-projection -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import this
module, and only as a checker.  The product package never imports it.

It is a second, independent statement of the same algorithm as ``oracle/sarbp_oracle.c``,
written in the array style of the reference (per-pixel vector temporaries) so that the two can
check each other:

* ``process_point``       -- sar_focuser_puntual.py:21-155
* ``focus``               -- FocalizzatoreBpPx, sar_focuser_puntual.py:160-330 (flatten in F
                             order, Sint == -9999 masks the pixel, serial pixel loop, F reshape)
* ``aperture``            -- sar_focuser_puntual.py:57-66
* ``count_pp``            -- the unit of work (pixel-pulse pairs evaluated)

Constants (sar_focuser_puntual.py:35-40): PI is float32(pi) stored as a double, C = 2.9979246e8,
ALFA = 3.335987619777693e11, BETA = 4*PI*ALFA/C^2.
"""
from __future__ import annotations

import math

import numpy as np

C_F64 = 2.9979246e8
ALFA_F64 = 3.335987619777693e11
PI_F64 = 3.141592741012573242
BETA_F64 = 4.0 * PI_F64 * ALFA_F64 / (C_F64 * C_F64)


def aperture(index: int, sint: int) -> tuple[int, int]:
    """(start pulse, J) of one pixel -- sar_focuser_puntual.py:57-62.

    J uses the signed Sint, the half-width its absolute value."""
    J = int(math.ceil(sint / 2.0)) * 2 + 1
    h = int(math.ceil(abs(sint) / 2.0))
    return int(index) - h, J


def process_point(x, y, z, mask, index, sint, tx, ty, tz, raw, near_range, dra, lambda_val, ind_line_ra_min=0):
    """One pixel; ``raw`` is a complex64 array of shape (Nra, Naz).  Returns (complex, n).
    ``ind_line_ra_min``: lowest range bin kept by the filter (sar_focuser_puntual.py:106); the
    accumulation then skips bins outside [0, Nra) (:138)."""
    Nra, Naz = raw.shape
    if mask != 1 or sint == -9999:
        return 0j, 0
    a0, J = aperture(index, sint)
    if J <= 0:
        return 0j, 0
    a = a0 + np.arange(J, dtype=np.int64)
    a = a[(a < Naz - 1) & (a > 0)]
    if a.size == 0:
        return 0j, 0
    dx = x - tx[a]
    dy = y - ty[a]
    dz = z - tz[a]
    r2 = dx * dx + dy * dy + dz * dz
    r = np.sqrt(r2)
    rb = np.rint((r - near_range) / dra)  # half-to-even, as np.round
    keep = (rb < Nra - 1) & (rb >= ind_line_ra_min) & (rb >= 0)
    if not np.any(keep):
        return 0j, 0
    r = r[keep]
    r2 = r2[keep]
    rb = rb[keep].astype(np.int64)
    a = a[keep]
    phase = (-4.0 * PI_F64 / lambda_val * r) + (BETA_F64 * r2)
    # exp(-j*phase) = (cos phase, -sin phase); scalar math.cos/sin are glibc's.
    fr = np.array([math.cos(p) for p in phase])
    fi = -np.array([math.sin(p) for p in phase])
    d = raw[rb, a]
    dre = d.real.astype(np.float64)
    dim = d.imag.astype(np.float64)
    t1r = r2 * dre
    t1i = r2 * dim
    tr = t1r * fr - t1i * fi
    ti = t1r * fi + t1i * fr
    acc_re = 0.0
    acc_im = 0.0
    for k in range(tr.size):  # ascending-j sequential sum
        acc_re += float(tr[k])
        acc_im += float(ti[k])
    n = int(tr.size)
    return complex(acc_re / n, acc_im / n), n


def focus(xxg1, yyg1, zzg1, mask, index, wgsx, wgsy, wgsz, raw, near_range, dra, lambda_val, Sint):
    """FocalizzatoreBpPx semantics on 2-D inputs.  Returns (complex128 (dimx,dimy), n (dimx,dimy))."""
    dimx, dimy = xxg1.shape
    xf = np.asarray(xxg1, np.float64).flatten(order="F")
    yf = np.asarray(yyg1, np.float64).flatten(order="F")
    zf = np.asarray(zzg1, np.float64).flatten(order="F")
    mf = np.asarray(mask, np.uint8).flatten(order="F")
    inf = np.asarray(index, np.int32).flatten(order="F")
    sf = np.asarray(Sint, np.int32).flatten(order="F")
    mf[sf == -9999] = 0
    out = np.zeros(xf.size, np.complex128)
    cnt = np.zeros(xf.size, np.int32)
    for i in range(xf.size):
        out[i], cnt[i] = process_point(xf[i], yf[i], zf[i], mf[i], inf[i], sf[i], wgsx, wgsy, wgsz,
                                       raw, near_range, dra, lambda_val)
    return out.reshape((dimx, dimy), order="F"), cnt.reshape((dimx, dimy), order="F")


def count_pp(mask, index, sint, Naz) -> int:
    """Pixel-pulse pairs whose geometry is evaluated (range-bin test not applied)."""
    m = np.asarray(mask).ravel()
    idx = np.asarray(index, np.int64).ravel()
    s = np.asarray(sint, np.int64).ravel()
    ok = (m == 1) & (s != -9999)
    J = np.ceil(s / 2.0).astype(np.int64) * 2 + 1
    h = np.ceil(np.abs(s) / 2.0).astype(np.int64)
    lo = np.maximum(idx - h, 1)
    hi = np.minimum(idx - h + J - 1, Naz - 2)
    n = np.where(ok & (J > 0), np.maximum(hi - lo + 1, 0), 0)
    return int(n.sum())
