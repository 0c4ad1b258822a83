"""Preview of the focused image (main_gpu.py:635-743) on the CPU: the column window and the
display range follow the reference's rules, worked out by hand for each branch, and
generate_preview writes the linear, log-scale and full-resolution figures from a stored image."""
import json
import os

import numpy as np
import pytest

from sarbp import main_gpu
from sarbp.main_gpu import preview_scale, preview_window


@pytest.mark.parametrize("dim_az,size,processed,want", [
    (800, 1000, None, (0, 800)),        # no wider than the preview: every column
    (1000, 1000, 400, (0, 1000)),
    (5000, 1000, None, (2000, 3000)),   # no attribute: centred on the image
    (5000, 1000, 3000, (1000, 2000)),   # more than the preview processed: centred on them
    (5000, 1000, 1001, (0, 1000)),      # ... clamped at column 0
    (5000, 1000, 9000, (2000, 3000)),   # ... never beyond the image
    (5000, 1000, 600, (0, 600)),        # fewer processed: exactly those
    (5000, 1000, 1000, (0, 1000)),
    (4999, 1000, None, (1999, 2999)),   # integer halves
])
def test_preview_window(dim_az, size, processed, want):
    assert preview_window(dim_az, size, processed) == want


def test_preview_scale():
    a = np.concatenate([np.zeros(50), np.arange(1, 101, dtype=np.float64)])  # median 25.5 of 150 values
    vmin, vmax = preview_scale(a)
    assert vmax == min(15 * 25.5, 100.0) == 100.0
    assert vmin == pytest.approx(1.0 + 0.01 * 99)  # 1st percentile of 1..100 (linear interpolation)
    b = np.full(10, 2.0)
    b[0] = 100.0
    assert preview_scale(b) == (2.0, 30.0)  # 15 x median caps the bright outlier
    assert preview_scale(np.zeros(7)) == (0.0, 0.0)
    c = np.array([np.nan, np.inf, 3.0])
    assert preview_scale(c) == (3.0, 3.0)


def test_generate_preview_files(tmp_path, monkeypatch):
    pytest.importorskip("matplotlib")
    monkeypatch.chdir(tmp_path)
    f = main_gpu.SARSuperOptimizedFocuser(output_file=str(tmp_path / "slc.npy"))
    rng = np.random.default_rng(0)
    img = (rng.normal(size=(40, 1500)) + 1j * rng.normal(size=(40, 1500))).astype(np.complex128)
    img[:, 700:] = 0
    img[3, 5] = np.nan
    np.save(tmp_path / "slc.npy", np.asfortranarray(img))
    with open(tmp_path / "slc.npy.attrs.json", "w") as fh:
        json.dump({"processed_azimuth": 700}, fh)
    assert f._read_attrs()["processed_azimuth"] == 700
    fn = f.generate_preview(preview_size=500)
    assert fn == os.path.join("checkpoints", "SlcCuPy_SuperOptimized_Preview.png")
    for name in ("SlcCuPy_SuperOptimized_Preview.png", "SlcCuPy_SuperOptimized_LogScale_Preview.png",
                 "SlcCuPy_SuperOptimized_Full.png"):
        assert os.path.getsize(os.path.join("checkpoints", name)) > 1000
