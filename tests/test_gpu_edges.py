"""Every aperture and range-bin edge case through every GPU path, against the oracle (GPU).

The grid of sarbp.synth.edge_grid_scene crosses Sint in [-5, 12], -9999 and two long apertures with
aperture centres Index = -3..3 and Naz-4..Naz+2 (apertures clipped at both ends of the track,
sar_focuser_puntual.py:57-66), in bands of pixels whose range bins step across bin 0, across the
excluded last bin Nra-1 (:106) and inside the swath.  Each path does its own integer work -- the
FP64 tile kernel tests each lane's aperture and bin per pulse; the split kernel counts contributors
in closed form and corrects them for bins that leave the swath, skipping the per-lane tests in
groups inside every aperture; the batch kernel and the serving loop cut the tiles' pulse ranges into
items and segments -- so every path's contributor counts must equal the oracle's bit for bit, FP64
values within 1e-10 relative, and each request's batch and served results must equal the same
request focused alone.  process_point's ind_line_ra_min (:27, :106) moves the lowest contributing
bin; the scenes' bin floor is checked against the oracle's restatement of it.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O
from sarbp import synth
from sarbp.batch import Batch, Request
from sarbp.core import Dataset
from sarbp.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[False, True], ids=["track-start", "track-end"])
def grid(request):
    sc = synth.edge_grid_scene(end=request.param)
    ref, rcnt = O.focus(*sc.args())
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    print("case", sc.name, "pixels", sc.x.size, "contributors", int(rcnt.sum()), flush=True)
    yield sc, ds, ref.ravel(order="F"), rcnt.ravel(order="F")
    ds.close()


def _check(prec, got, cnt, ref, rcnt, what):
    assert np.array_equal(cnt, rcnt), (what, np.nonzero(cnt != rcnt)[0][:10])
    nz = rcnt > 0
    assert np.all(got[~nz] == 0), what
    if not nz.any():
        return
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    if prec == "f64":
        assert rel.max() <= 1e-10, (what, float(rel.max()))
    else:  # the split form's per-term accuracy, on sums of one to a few noise samples that can cancel
        a = np.abs(ref[nz])
        err = np.abs(got[nz] - ref[nz])
        assert np.median(rel) <= 1e-4, (what, float(np.median(rel)))
        assert np.all(err <= 1e-4 * (a + np.median(a))), (what, float((err / (a + np.median(a))).max()))


def test_edges_scene_paths(gpu, grid):
    sc, ds, ref, rcnt = grid
    scene = ds.scene(*sc.flat(), dimx=sc.dimx)
    for prec in ("f64", "split"):
        scene.focus(prec)
        got, cnt = scene.result(with_counts=True)
        _check(prec, got, cnt, ref, rcnt, ("scene", prec))
    scene.set_tiling("f64", 0, 16, 0)  # one pixel per lane
    scene.focus("f64")
    got, cnt = scene.result(with_counts=True)
    _check("f64", got, cnt, ref, rcnt, ("scene lanes", "f64"))
    scene.close()


def _requests(sc):
    f = sc.flat()
    return [Request(*f, dimx=sc.dimx, precision=p) for p in ("f64", "split")]


def _alone(ds, r):
    s = ds.scene(r.x, r.y, r.z, r.mask, r.index, r.sint, dimx=r.dimx)
    s.focus(r.precision)
    out = s.result(with_counts=True)
    s.close()
    return out


def test_edges_batch_kernel(gpu, grid):
    sc, ds, ref, rcnt = grid
    reqs = _requests(sc)
    b = Batch(ds, reqs)
    for tp, ch in ((64, 8), (128, 32), (256, 64)):
        b.run(tp, ch)
        for k, r in enumerate(reqs):
            got, cnt = b.result(k, with_counts=True)
            _check(r.precision, got, cnt, ref, rcnt, ("batch", tp, ch, r.precision))
            a, ac = _alone(ds, r)
            assert np.array_equal(got, a) and np.array_equal(cnt, ac), ("batch vs alone", tp, ch, r.precision)
    b.close()


@pytest.mark.parametrize("max_segments", [1, 16])
def test_edges_serving(gpu, grid, max_segments):
    """Served in many short rounds (8-pulse chunks) or with the long apertures' pulses in segments."""
    sc, ds, ref, rcnt = grid
    # rounds of ~3e4 pixel-pulse pairs (the server's initial rate estimate x 0.1 us): 72-pulse spans
    srv = Server(quantum_ms=1e-4, pulse_chunk=8, max_segments=max_segments)
    hs = [srv.submit(ds, r) for r in _requests(sc)]
    for h in hs:
        assert h.wait(timeout=60)
        got, cnt = h.result(with_counts=True)
        inf = h.info()
        print("served", h.req.precision, "segments", inf["segments"], "rounds", inf["rounds"], flush=True)
        _check(h.req.precision, got, cnt, ref, rcnt, ("served", max_segments, h.req.precision))
        if max_segments == 1:
            assert inf["rounds"] > 1
    srv.close()


@pytest.mark.parametrize("bin_min", [-3, 1, 2, 62, 63, 200])
def test_edges_bin_floor(gpu, grid, bin_min):
    """The scene's lowest contributing bin (process_point's ind_line_ra_min) against the oracle."""
    sc, ds, _, _ = grid
    x, y, z, m, idx, s = sc.flat()
    refs = [O.point(x[i], y[i], z[i], m[i], idx[i], s[i], sc.tx, sc.ty, sc.tz, sc.raw, sc.near, sc.dra, sc.lam,
                    ind_line_ra_min=bin_min) for i in range(x.size)]
    ref = np.array([r[0] for r in refs])
    rcnt = np.array([r[1] for r in refs], np.int32)
    scene = ds.scene(x, y, z, m, idx, s, dimx=sc.dimx)
    scene.set_bin_min(bin_min)
    try:
        for prec in ("f64", "split"):
            scene.focus(prec)
            got, cnt = scene.result(with_counts=True)
            _check(prec, got, cnt, ref, rcnt, ("bin floor", bin_min, prec))
    finally:
        scene.close()
