"""Writes the small golden fixture tests/golden/tiny_16x8_naz257_nra64.npz.

The reference ships no tests, fixtures or golden vectors, and its own code cannot run here, so
the expected outputs come from the NumPy restatement of process_point / FocalizzatoreBpPx
(oracle/bp_numpy.py, sar_focuser_puntual.py:21-155, 160-330), which tests/test_oracle.py pins
against closed forms, exact ties and the C restatement.  The fixture freezes those outputs:
the C oracle must keep reproducing them bit for bit, and the GPU paths must meet their bounds
against them.  Inputs: a 16 x 8 pixel scene over 257 pulses and 64 range bins in the full
workload's geometry, with Sint = -9999 / 0 / 1 / negative values, masked pixels, apertures
clipped at both track ends and pixels whose bins leave the swath.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "sar-backprojection_amd")]

from oracle import bp_numpy as B  # noqa: E402
from sarbp import synth  # noqa: E402

OUT = os.path.join(HERE, "tiny_16x8_naz257_nra64.npz")


def main():
    sc = synth.tiny_scene(seed=7, nx=16, ny=8, naz=257, nra=64, sint_lo=10, sint_hi=90, ntarget=3)
    out, cnt = B.focus(sc.x, sc.y, sc.z, sc.mask, sc.index, sc.tx, sc.ty, sc.tz, sc.raw, sc.near, sc.dra, sc.lam,
                       sc.sint)
    np.savez(OUT, x=sc.x, y=sc.y, z=sc.z, mask=sc.mask, index=sc.index, sint=sc.sint, tx=sc.tx, ty=sc.ty,
             tz=sc.tz, raw=sc.raw, near=sc.near, dra=sc.dra, lam=sc.lam, out=out, ncontrib=cnt)
    print(f"wrote {OUT}: {int((cnt > 0).sum())} of {cnt.size} pixels with contributors, "
          f"{int(cnt.sum())} pixel-pulse terms")


if __name__ == "__main__":
    main()
