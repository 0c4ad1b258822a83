"""Error and quality reports (sarbp.metrics) against values fixed in closed form.

The reports follow compare_phase.py:41-198 and choerence_measure.py:165-229; each case below builds
images whose phase and magnitude differences are known exactly (a constant rotation, a linear ramp of
rotations, a known amplitude scale, planted zero-amplitude pixels), so a wrong mask, sign, wrap or
normalisation changes a checked number.
"""
import json
import os

import numpy as np
import pytest

from sarbp import formats, metrics


def _img(shape=(40, 25), seed=0, amp_lo=0.5):
    rng = np.random.default_rng(seed)
    a = rng.uniform(amp_lo, 2.0, shape)
    p = rng.uniform(-np.pi, np.pi, shape)
    return a * np.exp(1j * p)


def test_constant_rotation():
    ref = _img()
    th = 0.3
    st = metrics.compare_phase(ref * np.exp(1j * th), ref)
    n = ref.size
    assert st["valid_pixels"] == n and st["invalid_pixels"] == 0
    for k in ("mean_rad", "median_rad", "rms_rad", "min_rad", "max_rad"):
        assert st[k] == pytest.approx(th, abs=1e-12), k
    assert st["std_rad"] < 1e-12 and st["range_rad"] < 1e-12 and st["iqr_rad"] < 1e-12
    assert st["mean_deg"] == pytest.approx(th * 180 / np.pi, abs=1e-9)
    # 0.3 rad lies between pi/16 (0.196) and pi/8 (0.393)
    assert st["large_errors"] == {"pi_4": 0, "pi_8": 0, "pi_16": n, "pi_32": n}
    # sign: file1 minus file2
    assert metrics.compare_phase(ref, ref * np.exp(1j * th))["mean_rad"] == pytest.approx(-th, abs=1e-12)


def test_rotation_beyond_pi_is_wrapped():
    ref = _img(seed=1)
    st = metrics.compare_phase(ref * np.exp(1j * 3.5), ref)
    assert st["mean_rad"] == pytest.approx(3.5 - 2 * np.pi, abs=1e-12)
    assert st["large_errors"]["pi_4"] == ref.size
    assert metrics.wrap_phase(3.5) == pytest.approx(3.5 - 2 * np.pi)
    assert metrics.wrap_phase(-3.5) == pytest.approx(2 * np.pi - 3.5)
    # the two wraps of the reference scripts differ only at the branch point: [-pi, pi) vs (-pi, pi]
    assert metrics.wrap_phase(np.pi) == pytest.approx(-np.pi)
    assert metrics.phase_difference(np.array([-1.0 + 0j]), np.array([1.0 + 0j]))[0] == pytest.approx(np.pi)


def test_ramp_of_rotations_statistics():
    # 101 rotations -1, -0.98, ..., 1 rad: percentiles, IQR, RMS and counts in closed form
    th = np.linspace(-1.0, 1.0, 101)
    ref = np.full(101, 1.0 + 0j)
    st = metrics.compare_phase(ref * np.exp(1j * th), ref)
    assert st["mean_rad"] == pytest.approx(0.0, abs=1e-12)
    assert st["median_rad"] == pytest.approx(0.0, abs=1e-12)
    p = st["percentiles"]
    assert (p["p1"], p["p5"], p["p25"], p["p75"], p["p95"], p["p99"]) == pytest.approx(
        (-0.98, -0.9, -0.5, 0.5, 0.9, 0.98), abs=1e-12)
    assert st["iqr_rad"] == pytest.approx(1.0, abs=1e-12)
    # sum_{j=-50..50} (j/50)^2 = 2 * (50*51*101/6) / 2500 = 34.34
    assert st["rms_rad"] == pytest.approx(np.sqrt(34.34 / 101), rel=1e-12)
    assert st["std_rad"] == pytest.approx(np.sqrt(34.34 / 101), rel=1e-12)
    assert st["range_rad"] == pytest.approx(2.0, abs=1e-12)
    # |j/50| > pi/4 for |j| >= 40: 11 per side; > pi/8 for |j| >= 20: 31 per side; > pi/16 for |j| >= 10:
    # 41 per side; > pi/32 for |j| >= 5: 46 per side
    assert st["large_errors"] == {"pi_4": 22, "pi_8": 62, "pi_16": 82, "pi_32": 92}


def test_zero_amplitude_mask():
    ref = _img(seed=2)
    test = ref * np.exp(1j * 0.1)
    test.flat[:7] = 1e-7          # below 1e-6: masked
    ref.flat[10:13] = 1e-6 + 0j   # exactly 1e-6: masked (strict >)
    st = metrics.compare_phase(test, ref)
    assert st["valid_pixels"] == ref.size - 10 and st["invalid_pixels"] == 10
    assert st["total_pixels"] == st["valid_pixels"]
    assert st["mean_rad"] == pytest.approx(0.1, abs=1e-12)
    st2 = metrics.compare_phase(test, ref, mask_zero_amp=False)
    assert st2["valid_pixels"] == ref.size
    assert metrics.compare_phase(np.zeros(5, complex), np.ones(5, complex)) is None
    with pytest.raises(ValueError):
        metrics.compare_phase(np.ones(4, complex), np.ones(5, complex))


def test_coherence_report_known_scale_and_rotation():
    ref = _img(seed=3)
    eps, th = -2e-3, 0.05
    test = ref * (1 + eps) * np.exp(1j * th)
    ref[0, 0] = 1e-6 * np.abs(ref).max()  # below 1e-5 x max|ref|: no phase statistics there
    test[0, 0] = ref[0, 0] * (1 + eps) * np.exp(1j * th)
    r = metrics.coherence_report(ref, test)
    a = np.abs(ref)
    assert r["mean_signed_diff_mag"] == pytest.approx(eps * a.mean(), rel=1e-9)  # test - ref
    assert r["mean_abs_diff_mag"] == pytest.approx(-eps * a.mean(), rel=1e-9)
    assert r["max_abs_diff_mag"] == pytest.approx(-eps * a.max(), rel=1e-9)
    assert r["rmse_mag"] == pytest.approx(-eps * np.sqrt(np.mean(a ** 2)), rel=1e-9)
    assert r["std_signed_diff_mag"] == pytest.approx(-eps * a.std(), rel=1e-9)
    assert r["phase_threshold"] == pytest.approx(1e-5 * a.max())
    assert r["valid_phase_pixels"] == ref.size - 1
    assert r["mean_signed_phase_diff_rad"] == pytest.approx(th, abs=1e-12)
    assert r["max_abs_phase_diff_rad"] == pytest.approx(th, abs=1e-12)
    assert r["std_signed_phase_diff_rad"] < 1e-12
    z = metrics.coherence_report(np.zeros(4, complex), np.zeros(4, complex))
    assert z["valid_phase_pixels"] == 0 and "mean_abs_phase_diff_rad" not in z


def test_split_error_report_known_errors():
    ref = _img(seed=4)
    test = ref * (1 + 1e-5) * np.exp(1j * 1e-4)
    r = metrics.split_error_report(test, ref)
    for k in ("phase_rms_rad_abs1e-6", "phase_max_rad_abs1e-6", "phase_rms_rad_rel1e-5max", "phase_max_rad_rel1e-5max"):
        assert r[k] == pytest.approx(1e-4, rel=1e-6), k
    assert r["amp_rel_rmse"] == pytest.approx(1e-5, rel=1e-6)
    assert r["amp_rel_max_rel1e-5max"] == pytest.approx(1e-5, rel=1e-6)
    for k in ("phase_rms_rad_rel1e-2max", "phase_max_rad_rel1e-2max"):
        assert r[k] == pytest.approx(1e-4, rel=1e-6), k
    assert r["amp_rel_rms_rel1e-2max"] == pytest.approx(1e-5, rel=1e-6)
    assert r["amp_rel_max_rel1e-2max"] == pytest.approx(1e-5, rel=1e-6)
    a = np.abs(ref)
    assert r["pixels_rel1e-2max"] == int((a > 1e-2 * a.max()).sum()) > 0
    # an error planted on one dark pixel (1e-3 x its amplitude) shows in the 1e-5 x max tails only
    dark_ref = ref.copy()
    dark_ref.flat[7] *= 1e-3
    bad = dark_ref.copy()
    bad.flat[7] *= 1.5 * np.exp(0.5j)
    r2 = metrics.split_error_report(bad, dark_ref)
    assert r2["phase_max_rad_rel1e-5max"] == pytest.approx(0.5) and r2["amp_rel_max_rel1e-5max"] == pytest.approx(0.5)
    assert r2["phase_max_rad_rel1e-2max"] == 0.0 and r2["amp_rel_max_rel1e-2max"] == 0.0
    same = metrics.split_error_report(ref, ref)
    assert same["phase_rms_rad_abs1e-6"] == 0.0 and same["amp_rel_rmse"] == 0.0


def test_target_peaks():
    nx, ny = 60, 30
    x0, y0, dx, dy = 100.0, -20.0, 1.0, 4.0
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pos = [(10, 5), (30, 15), (50, 24), (2, 2)]  # the last one's window leaves the grid
    ref = np.full((nx, ny), 0.01 + 0j)
    test = ref.copy()
    for k, (px, py) in enumerate(pos):
        g = np.exp(-((ix - px) ** 2 + (iy - py) ** 2) / 2.0)
        ref += 5 * g
        shift = 1 if k == 1 else 0  # the second target's test peak is one pixel off
        test += 5 * (1.02 if k == 2 else 1.0) * np.exp(-((ix - px - shift) ** 2 + (iy - py) ** 2) / 2.0)
    txy = [(x0 + px * dx + 0.3, y0 + py * dy - 1.0) for px, py in pos]
    r = metrics.target_peaks(test, ref, txy, x0, y0, dx, dy)
    assert r["targets"] == 4 and r["targets_above_noise"] == 3
    assert r["same_peak_pixel_fraction"] == pytest.approx(2 / 3)
    # peaks 5 x 1.02 + 0.01 against 5 + 0.01 (the background adds to both)
    assert r["peak_height_rel_err_max"] == pytest.approx(0.1 / 5.01, rel=1e-9)


def test_load_image_formats_and_cli(tmp_path):
    img = _img((12, 7), seed=5)
    base = str(tmp_path / "Slc")
    formats.write_outputs(img, base)
    formats.write_slcmatr(img, str(tmp_path))
    np.asarray(img, np.complex64).flatten(order="F").tofile(str(tmp_path / "truth.dat"))
    c64 = img.astype(np.complex64)
    assert np.array_equal(metrics.load_image(base + "_InterleavedF32.dat", img.shape, "interleaved"), c64)
    assert np.array_equal(metrics.load_image(base + "_SplitF32.dat", img.shape, "split"), c64)
    assert np.array_equal(metrics.load_image(str(tmp_path / "truth.dat"), img.shape, "truth"), c64)
    assert np.array_equal(metrics.load_image(base + ".npy", img.shape, "npy"), img)
    assert np.array_equal(metrics.load_image(str(tmp_path), img.shape, "slcmatr"), img)
    with pytest.raises(ValueError):
        metrics.load_image(base + "_SplitF32.dat", (12, 8), "split")
    with pytest.raises(ValueError):
        metrics.load_image(base + ".npy", img.shape, "bogus")
    st = metrics.compare_phase_files(base + "_InterleavedF32.dat", base + ".npy", img.shape, "interleaved", "npy")
    assert st["valid_pixels"] == img.size and st["rms_rad"] < 1e-6  # float32 rounding only
    out = str(tmp_path / "cmp")
    rc = metrics._main(["--file1", base + "_SplitF32.dat", "--file2", str(tmp_path / "truth.dat"),
                        "--shape", "12,7", "--fmt1", "split", "--fmt2", "truth", "--out", out])
    assert rc == 0
    with open(out + "_statistics.json") as f:
        saved = json.load(f)
    assert saved["valid_pixels"] == img.size and saved["rms_rad"] == 0.0
    assert os.path.exists(out + "_statistics.json")


def test_figures_are_written(tmp_path):
    pytest.importorskip("matplotlib")
    ref = _img((30, 20), seed=6)
    test = ref * np.exp(1j * 0.2)
    for fn in (metrics.plot_compare_phase(test, ref, str(tmp_path / "p.png")),
               metrics.plot_coherence(ref, test, str(tmp_path / "c.png"))):
        with open(fn, "rb") as f:
            head = f.read(8)
        assert head == b"\x89PNG\r\n\x1a\n" and os.path.getsize(fn) > 10000
    base = str(tmp_path / "Slc")
    formats.write_outputs(ref, base)
    out = str(tmp_path / "cmp")
    assert metrics._main(["--file1", base + "_SplitF32.dat", "--file2", base + ".npy", "--shape", "30,20",
                          "--fmt1", "split", "--fmt2", "npy", "--out", out, "--plot"]) == 0
    assert os.path.exists(out + "_phase_diff.png") and os.path.exists(out + "_comparison.png")
