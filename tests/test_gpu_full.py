"""The full workload (4000 x 250 pixels, 40001 pulses, Nra 4762, 1.5 GB of RawData) through the C ABI,
checked by properties that hold exactly at any size (GPU).

- Every pulse of every aperture contributes: the 4762-bin swath starts at 4 km with 0.5 m bins and
  so covers every slant range of the scene (4.2-5.0 km).  The contributor counts of both paths
  therefore add up to the pixel-pulse pairs the oracle counts from the integer inputs alone.
- Scaling RawData by 2 and rotating it by i are exact in float32, and every product and sum of both
  paths commutes with them exactly (a power-of-two scale, and a swap of the components with a sign),
  so the results must be 2x and i x the originals bit for bit: a mixed-up component, a dropped or
  doubled term, or a result that depends on anything but the samples breaks this.
- The split result stays within its stated tolerances of the FP64 result: RMS wrapped phase error
  <= 1e-3 rad and relative amplitude RMSE <= 1e-4 over pixels above 1e-5 x max, and every point
  target above the noise peaks on the same pixel.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O
from sarbp import metrics, synth
from sarbp.core import Dataset

pytestmark = pytest.mark.gpu

PRECISIONS = ("f64", "split")


@pytest.fixture(scope="module")
def full():
    print("case full workload 4000x250 px x 40001 pulses: synthesising", flush=True)
    sc = synth.full_scene(seed=1)
    print("case full workload: uploading RawData", flush=True)
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    scene = ds.scene(*sc.flat(), dimx=sc.dimx)
    res = {}
    for p in PRECISIONS:
        scene.focus(p)
        res[p] = scene.result(with_counts=True)
        print("case full workload:", p, "focused", flush=True)
    yield sc, res
    scene.close()
    ds.close()


def _focus_with(sc, raw, precision):
    ds = Dataset(raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    scene = ds.scene(*sc.flat(), dimx=sc.dimx)
    scene.focus(precision)
    out = scene.result()
    scene.close()
    ds.close()
    return out


def test_full_counts_cover_every_aperture_pulse(gpu, full):
    sc, res = full
    f = lambda a: np.asarray(a).flatten(order="F")  # noqa: E731
    pp = O.count_pp(f(sc.mask), f(sc.index), f(sc.sint), sc.Naz)
    c64 = res["f64"][1]
    print("pixel-pulse pairs", pp, "counted", int(c64.sum()), flush=True)
    assert int(c64.sum()) == pp
    assert np.array_equal(res["split"][1], c64)
    off = f(sc.mask) != 1
    assert np.all(c64[off] == 0) and np.all(res["f64"][0][off] == 0)


def test_full_scaling_and_rotation_are_exact(gpu, full):
    sc, res = full
    for p in PRECISIONS:
        ref = res[p][0]
        print("case full workload: RawData x 2,", p, flush=True)
        got = _focus_with(sc, np.asfortranarray(sc.raw * np.complex64(2)), p)
        assert np.array_equal(got, 2 * ref), p
        print("case full workload: RawData x i,", p, flush=True)
        rot = np.empty_like(sc.raw)
        rot.real, rot.imag = -sc.raw.imag, sc.raw.real
        got = _focus_with(sc, rot, p)
        want = np.empty_like(ref)
        want.real, want.imag = -ref.imag, ref.real
        assert np.array_equal(got, want), p


def test_full_split_within_tolerance_of_f64(gpu, full):
    sc, res = full
    e = metrics.split_error_report(res["split"][0], res["f64"][0])
    print("split vs f64 at full size", e, flush=True)
    assert e["phase_rms_rad_rel1e-5max"] <= 1e-3
    assert e["amp_rel_rmse"] <= 1e-4
    shape = sc.shape
    pk = metrics.target_peaks(res["split"][0].reshape(shape, order="F"), res["f64"][0].reshape(shape, order="F"),
                              sc.targets[:, :2], float(sc.x[0, 0]), float(sc.y[0, 0]),
                              float(sc.x[1, 0] - sc.x[0, 0]), float(sc.y[0, 1] - sc.y[0, 0]))
    print("target peaks", pk, flush=True)
    assert pk["targets_above_noise"] > 1000 and pk["same_peak_pixel_fraction"] == 1.0
