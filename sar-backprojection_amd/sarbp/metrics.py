"""Error and quality reports for focused images, with the definitions of the reference's comparison
scripts, computed on arrays (or on files in the output formats the focusers write).

- :func:`compare_phase`: wrapped phase difference statistics of two complex images
  (compare_phase.py:41-198): difference ``angle(exp(j(angle(a) - angle(b))))``, pixels where both
  amplitudes exceed 1e-6, mean / median / std / RMS / range / IQR / percentiles and counts above
  pi/4 .. pi/32, in a dict with the same keys.
- :func:`coherence_report`: the magnitude and wrapped-phase difference statistics of
  choerence_measure.py:165-229: ``wrap_phase`` = (d + pi) mod 2 pi - pi, signed differences
  test - reference, phase statistics over the pixels whose amplitudes both exceed
  1e-5 x the reference's largest amplitude.
- :func:`split_error_report`: the figures used to accept the split-precision path against the FP64
  path: RMS and max wrapped phase error under both masks above, relative amplitude RMSE
  (||test| - |ref||_2 / ||ref||_2) and the largest per-pixel relative amplitude error.
- :func:`target_peaks`: point-target responses of two images of one scene: same peak pixel,
  relative peak-height difference.
- :func:`plot_compare_phase` and :func:`plot_coherence`: the figures of the two scripts (phases,
  amplitudes, difference maps, histogram), written as PNG with matplotlib (Agg).

    python -m sarbp.metrics --file1 A_InterleavedF32.dat --file2 B.npy --shape 4000,250 \
        --fmt1 interleaved --fmt2 npy [--no-mask] [--out prefix [--plot]]
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

from . import formats

PHASE_MAGNITUDE_THRESHOLD_FRACTION = 1e-5  # choerence_measure.py:20
WELL_CONDITIONED_FRACTION = 1e-2  # per-pixel error tails of the split report: |a| > this x max|ref|
ZERO_AMPLITUDE = 1e-6                       # compare_phase.py:55


def wrap_phase(phase_diff):
    """Phase difference brought into [-pi, pi) (choerence_measure.py:165-167)."""
    return (np.asarray(phase_diff) + np.pi) % (2 * np.pi) - np.pi


def phase_difference(a, b):
    """``angle(exp(j (angle(a) - angle(b))))``, in (-pi, pi] (compare_phase.py:50-52)."""
    return np.angle(np.exp(1j * (np.angle(a) - np.angle(b))))


def compare_phase(arr1, arr2, mask_zero_amp: bool = True):
    """Phase-difference statistics of two complex images of one shape (compare_phase.py:41-198).

    Returns the reference's statistics dict, or None when no pixel is valid
    (compare_phase.py:63-65)."""
    arr1, arr2 = np.asarray(arr1), np.asarray(arr2)
    if arr1.shape != arr2.shape:
        raise ValueError(f"shape mismatch: {arr1.shape} vs {arr2.shape}")
    d = phase_difference(arr1, arr2)
    if mask_zero_amp:
        m = (np.abs(arr1) > ZERO_AMPLITUDE) & (np.abs(arr2) > ZERO_AMPLITUDE)
        d = np.where(m, d, np.nan)
    valid = np.isfinite(d)
    v = d[valid]
    if v.size == 0:
        return None
    mean, median, std = float(np.mean(v)), float(np.median(v)), float(np.std(v))
    p1, p5, q25, q75, p95, p99 = (float(np.percentile(v, q)) for q in (1, 5, 25, 75, 95, 99))
    lo, hi = float(np.min(v)), float(np.max(v))
    rms = float(np.sqrt(np.mean(v ** 2)))
    av = np.abs(v)
    deg = 180 / np.pi
    return {
        "total_pixels": int(np.sum(valid)),  # as the reference counts it: the valid pixels
        "valid_pixels": int(v.size),
        "invalid_pixels": int(np.sum(~valid)),
        "mean_rad": mean, "mean_deg": mean * deg,
        "median_rad": median, "median_deg": median * deg,
        "std_rad": std, "std_deg": std * deg,
        "rms_rad": rms, "rms_deg": rms * deg,
        "min_rad": lo, "max_rad": hi, "range_rad": hi - lo,
        "iqr_rad": q75 - q25,
        "percentiles": {"p1": p1, "p5": p5, "p25": q25, "p75": q75, "p95": p95, "p99": p99},
        "large_errors": {f"pi_{k}": int(np.sum(av > np.pi / k)) for k in (4, 8, 16, 32)},
    }


def coherence_report(ref, test, threshold_fraction: float = PHASE_MAGNITUDE_THRESHOLD_FRACTION):
    """Magnitude and wrapped-phase differences of ``test`` against ``ref``
    (choerence_measure.py:185-229; there ref is the IDL image and test the Python one)."""
    ref = np.asarray(ref).astype(np.complex128)
    test = np.asarray(test)
    if ref.shape != test.shape:
        raise ValueError(f"shape mismatch: {ref.shape} vs {test.shape}")
    mag_r, mag_t = np.abs(ref), np.abs(test)
    dmag = mag_t - mag_r
    admag = np.abs(dmag)
    dph = wrap_phase(np.angle(test) - np.angle(ref))
    mmax = float(np.max(mag_r)) if np.any(mag_r) else 1.0
    thr = threshold_fraction * mmax
    valid = (mag_r > thr) & (mag_t > thr)
    out = {
        "pixels": int(ref.size),
        "mean_abs_diff_mag": float(np.mean(admag)),
        "std_abs_diff_mag": float(np.std(admag)),
        "max_abs_diff_mag": float(np.max(admag)),
        "mean_signed_diff_mag": float(np.mean(dmag)),
        "std_signed_diff_mag": float(np.std(dmag)),
        "rmse_mag": float(np.sqrt(np.mean(dmag ** 2))),
        "phase_threshold": thr,
        "valid_phase_pixels": int(np.sum(valid)),
    }
    if out["valid_phase_pixels"]:
        s, a = dph[valid], np.abs(dph[valid])
        out.update({
            "mean_abs_phase_diff_rad": float(np.mean(a)),
            "std_abs_phase_diff_rad": float(np.std(a)),
            "max_abs_phase_diff_rad": float(np.max(a)),
            "mean_signed_phase_diff_rad": float(np.mean(s)),
            "std_signed_phase_diff_rad": float(np.std(s)),
        })
    return out


def split_error_report(test, ref):
    """Split-precision result ``test`` against the FP64 result ``ref`` of the same pixels: RMS and max
    wrapped phase error over |a| > 1e-6 (compare_phase.py:55) and over |a| > 1e-5 x max|ref|
    (choerence_measure.py:195-197), the relative amplitude RMSE (over the whole image: the reading
    of a relative amplitude error the split path's 1e-4 target uses) and the largest per-pixel
    relative amplitude error over the second mask.  Per-pixel tails are also given over the
    well-conditioned pixels, |a| > 1e-2 x max|ref| (``*_rel1e-2max``): a pixel's phase and relative
    amplitude errors grow as 1/|a| where its sum nearly cancels, so the largest values over the
    lower thresholds sit on near-zero pixels."""
    test, ref = np.asarray(test), np.asarray(ref)
    a_t, a_r = np.abs(test), np.abs(ref)
    out = {}
    rmax = float(a_r.max()) if a_r.size else 0.0
    for name, thr in (("abs1e-6", ZERO_AMPLITUDE), ("rel1e-5max", PHASE_MAGNITUDE_THRESHOLD_FRACTION * rmax)):
        m = (a_t > thr) & (a_r > thr)
        d = phase_difference(test[m], ref[m])
        out[f"phase_rms_rad_{name}"] = float(np.sqrt(np.mean(d ** 2))) if m.any() else None
        out[f"phase_max_rad_{name}"] = float(np.abs(d).max()) if m.any() else None
    m = a_r > PHASE_MAGNITUDE_THRESHOLD_FRACTION * rmax
    den = float(np.sum(a_r ** 2))
    out["amp_rel_rmse"] = float(np.sqrt(np.sum((a_t - a_r) ** 2) / den)) if den > 0 else None
    out["amp_rel_max_rel1e-5max"] = float((np.abs(a_t[m] - a_r[m]) / a_r[m]).max()) if m.any() else None
    m = a_r > WELL_CONDITIONED_FRACTION * rmax
    if m.any():
        d = phase_difference(test[m], ref[m])
        rel = np.abs(a_t[m] - a_r[m]) / a_r[m]
        out.update({"pixels_rel1e-2max": int(m.sum()),
                    "phase_rms_rad_rel1e-2max": float(np.sqrt(np.mean(d ** 2))),
                    "phase_max_rad_rel1e-2max": float(np.abs(d).max()),
                    "amp_rel_rms_rel1e-2max": float(np.sqrt(np.mean(rel ** 2))),
                    "amp_rel_max_rel1e-2max": float(rel.max())})
    else:
        out.update({"pixels_rel1e-2max": 0, "phase_rms_rad_rel1e-2max": None, "phase_max_rad_rel1e-2max": None,
                    "amp_rel_rms_rel1e-2max": None, "amp_rel_max_rel1e-2max": None})
    return out


def target_peaks(test, ref, targets_xy, x0, y0, dx, dy, half_x=3, half_y=2, noise_factor=3.0):
    """Point-target responses of two images (shape (DimRg, DimAz), grid x = x0 + i dx, y = y0 + j dy)
    of one scene: for each target position (x, y) the peak of |image| in a (2 half_x + 1) x
    (2 half_y + 1) window around its nearest pixel; targets whose reference peak is at most
    ``noise_factor`` x the median amplitude, or whose window leaves the grid, are skipped."""
    a_t, a_r = np.abs(np.asarray(test)), np.abs(np.asarray(ref))
    if a_t.shape != a_r.shape or a_t.ndim != 2:
        raise ValueError("two 2-D images of one shape are needed")
    nx, ny = a_r.shape
    med = float(np.median(a_r))
    same, rel, used = 0, [], 0
    targets_xy = np.asarray(targets_xy, dtype=np.float64).reshape(-1, 2)
    for tx, ty in targets_xy:
        ix, iy = int(round((tx - x0) / dx)), int(round((ty - y0) / dy))
        if ix - half_x < 0 or ix + half_x >= nx or iy - half_y < 0 or iy + half_y >= ny:
            continue
        wt = a_t[ix - half_x:ix + half_x + 1, iy - half_y:iy + half_y + 1]
        wr = a_r[ix - half_x:ix + half_x + 1, iy - half_y:iy + half_y + 1]
        if wr.max() <= noise_factor * med:
            continue
        used += 1
        same += int(np.argmax(wt) == np.argmax(wr))
        rel.append(abs(float(wt.max()) / float(wr.max()) - 1.0))
    rel = np.asarray(rel)
    return {"targets": int(targets_xy.shape[0]), "targets_above_noise": used,
            "same_peak_pixel_fraction": same / max(used, 1),
            "peak_height_rel_err_rms": float(np.sqrt(np.mean(rel ** 2))) if used else None,
            "peak_height_rel_err_max": float(rel.max()) if used else None}


def _pyplot():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    return plt


def plot_compare_phase(arr1, arr2, out_png, mask_zero_amp: bool = True):
    """The 2 x 3 figure of compare_phase.py:123-173: phase of each image, the masked wrapped phase
    difference map, the amplitude of each image and the histogram of the difference."""
    plt = _pyplot()
    arr1, arr2 = np.asarray(arr1), np.asarray(arr2)
    d = phase_difference(arr1, arr2)
    if mask_zero_amp:
        d = np.where((np.abs(arr1) > ZERO_AMPLITUDE) & (np.abs(arr2) > ZERO_AMPLITUDE), d, np.nan)
    fig = plt.figure(figsize=(18, 10))
    panels = ((np.angle(arr1), "twilight", "Phase file 1", "Phase (rad)", {}),
              (np.angle(arr2), "twilight", "Phase file 2", "Phase (rad)", {}),
              (d, "twilight", "Phase difference map", "Phase diff (rad)", {"vmin": -np.pi, "vmax": np.pi}),
              (np.abs(arr1), "gray", "Amplitude file 1", "Amplitude", {}),
              (np.abs(arr2), "gray", "Amplitude file 2", "Amplitude", {}))
    for k, (img, cmap, title, label, kw) in enumerate(panels):
        plt.subplot(2, 3, k + 1)
        plt.imshow(img, cmap=cmap, aspect="auto", **kw)
        plt.colorbar(label=label)
        plt.title(title)
    plt.subplot(2, 3, 6)
    v = d[np.isfinite(d)].ravel()
    rng = None  # the data's extent, or a small window around a (numerically) constant difference
    if v.size == 0 or float(v.max() - v.min()) <= 1e-12:
        c = float(v.mean()) if v.size else 0.0
        rng = (c - 1e-6, c + 1e-6)
    plt.hist(v, bins=100, range=rng, color="C0", alpha=0.7)
    plt.xlabel("Phase diff (rad)")
    plt.ylabel("Count")
    plt.title("Phase diff histogram")
    plt.tight_layout()
    fig.savefig(out_png, dpi=100)
    plt.close(fig)
    return out_png


def plot_coherence(ref, test, out_png, threshold_fraction: float = PHASE_MAGNITUDE_THRESHOLD_FRACTION):
    """The 2 x 3 figure of choerence_measure.py:231-301: magnitudes of the reference and the test
    image (1st-99th percentile scale), |magnitude difference|, both phases and the |wrapped phase
    difference| over the pixels above the magnitude threshold."""
    plt = _pyplot()
    ref, test = np.asarray(ref).astype(np.complex128), np.asarray(test)
    mag_r, mag_t = np.abs(ref), np.abs(test)
    dph = np.abs(wrap_phase(np.angle(test) - np.angle(ref)))
    thr = threshold_fraction * (float(np.max(mag_r)) if np.any(mag_r) else 1.0)
    valid = (mag_r > thr) & (mag_t > thr)
    pos = mag_r[mag_r > 1e-9]
    lo = float(np.percentile(pos if pos.size else mag_r, 1))
    hi = float(np.percentile(mag_r, 99))
    fig, ax = plt.subplots(2, 3, figsize=(18, 10), sharex=True, sharey=True)
    ax[0, 0].imshow(mag_r, cmap="gray", vmin=lo, vmax=hi, aspect="auto")
    ax[0, 0].set_title("Reference magnitude")
    ax[0, 1].imshow(mag_t, cmap="gray", vmin=lo, vmax=hi, aspect="auto")
    ax[0, 1].set_title("Test magnitude")
    im = ax[0, 2].imshow(np.abs(mag_t - mag_r), cmap="magma", aspect="auto")
    ax[0, 2].set_title("|Magnitude difference|")
    fig.colorbar(im, ax=ax[0, 2], fraction=0.046, pad=0.04)
    ax[1, 0].imshow(np.angle(ref), cmap="hsv", vmin=-np.pi, vmax=np.pi, aspect="auto")
    ax[1, 0].set_title("Reference phase")
    ax[1, 1].imshow(np.angle(test), cmap="hsv", vmin=-np.pi, vmax=np.pi, aspect="auto")
    ax[1, 1].set_title("Test phase")
    shown = np.where(valid, dph, np.nan)
    vmax = float(np.percentile(dph[valid], 99.5)) if valid.any() else np.pi
    im = ax[1, 2].imshow(shown, cmap="viridis", vmin=0, vmax=vmax or np.pi, aspect="auto")
    ax[1, 2].set_title("|Wrapped phase difference| (valid pixels)")
    fig.colorbar(im, ax=ax[1, 2], fraction=0.046, pad=0.04)
    fig.tight_layout()
    fig.savefig(out_png, dpi=100)
    plt.close(fig)
    return out_png


def load_image(path: str, shape, fmt: str):
    """A complex image in one of the formats the focusers and the reference tools use:
    ``interleaved`` / ``split`` float32 (compare_phase.py:8-31), ``truth`` complex64
    (compare_phase.py:33-37), ``npy`` (``numpy.load``, no pickles), ``slcmatr`` (a directory with
    ``SlcMatr_real.dat`` / ``SlcMatr_imag.dat``, float64).  Sizes are checked against ``shape``."""
    shape = tuple(int(s) for s in shape)
    n = shape[0] * shape[1]
    if fmt in ("interleaved", "split"):
        size = os.path.getsize(path) // 4
        if size != 2 * n:
            raise ValueError(f"File size mismatch for {path} (expected {2 * n}, got {size})")
        return (formats.read_interleaved_f32 if fmt == "interleaved" else formats.read_split_f32)(path, shape)
    if fmt == "truth":
        size = os.path.getsize(path) // 8
        if size != n:
            raise ValueError(f"File size mismatch for {path} (expected {n}, got {size})")
        return np.fromfile(path, dtype=np.complex64).reshape(shape, order="F")
    if fmt == "npy":
        a = np.load(path, allow_pickle=False)
        if a.shape != shape:
            raise ValueError(f"{path}: shape {a.shape}, expected {shape}")
        return a
    if fmt == "slcmatr":
        return formats.read_slcmatr(path, shape)
    raise ValueError(f"Unknown format: {fmt}")


def compare_phase_files(file1, file2, shape, fmt1="interleaved", fmt2="truth", mask_zero_amp=True):
    """:func:`compare_phase` on two files (compare_phase.py:41-48 loading, then the statistics)."""
    return compare_phase(load_image(file1, shape, fmt1), load_image(file2, shape, fmt2), mask_zero_amp)


def _main(argv=None):
    ap = argparse.ArgumentParser(description="Compare phase and magnitude of two focused SAR images.")
    ap.add_argument("--file1", required=True)
    ap.add_argument("--file2", required=True, help="the reference image (e.g. truth)")
    ap.add_argument("--shape", default="4000,32000")
    fmts = ["interleaved", "split", "truth", "npy", "slcmatr"]
    ap.add_argument("--fmt1", default="interleaved", choices=fmts)
    ap.add_argument("--fmt2", default="truth", choices=fmts)
    ap.add_argument("--no-mask", action="store_true", help="do not mask zero amplitudes")
    ap.add_argument("--out", default=None, help="prefix of the <prefix>_statistics.json file")
    ap.add_argument("--plot", action="store_true", help="also write <prefix>_phase_diff.png and <prefix>_comparison.png")
    a = ap.parse_args(argv)
    shape = tuple(map(int, a.shape.split(",")))
    img1, img2 = load_image(a.file1, shape, a.fmt1), load_image(a.file2, shape, a.fmt2)
    rep = {"compare_phase": compare_phase(img1, img2, not a.no_mask), "coherence": coherence_report(img2, img1)}
    if a.out:
        with open(f"{a.out}_statistics.json", "w") as f:
            json.dump(rep["compare_phase"], f, indent=2)
        if a.plot:
            rep["figures"] = [plot_compare_phase(img1, img2, f"{a.out}_phase_diff.png", not a.no_mask),
                              plot_coherence(img2, img1, f"{a.out}_comparison.png")]
    print(json.dumps(rep, indent=2))
    return 0 if rep["compare_phase"] is not None else 1


if __name__ == "__main__":
    raise SystemExit(_main())
