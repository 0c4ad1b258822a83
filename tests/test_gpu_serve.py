"""Serving loop: continuous batching, preemption/resume, admission budget, FP64 fallback (GPU).

The bar for continuous batching: every request's result is bit-identical whether it ran alone or
in a mix with others; with each tile's pulses in one sequence it also equals the same request
focused by the batch kernel alone, and with the pulses of a small request split into segments its
FP64 result stays within the oracle's parity bound.
"""
import time

import numpy as np
import pytest

from oracle import oracle_lib as O
from sarbp import _lib, metrics, synth
from sarbp.batch import Batch, Request, patch_requests
from sarbp.core import Dataset
from sarbp.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    sc = synth.full_scene(seed=7, nx=512, ny=250, naz=5121, nra=4762, ntarget=100)
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    ds2 = Dataset(np.asfortranarray(sc.raw * np.complex64(1j)), sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    yield sc, ds, ds2
    ds.close()
    ds2.close()


def _serve_alone_and_mixed(srv, dsets, reqs):
    alone = []
    for d, r in zip(dsets, reqs):
        h = srv.submit(d, r)
        assert h.wait(timeout=60)
        alone.append(h.result(with_counts=True))
    hs = [srv.submit(d, r) for d, r in zip(dsets, reqs)]
    for h in hs:
        assert h.wait(timeout=60)
    return alone, hs


def test_alone_equals_mix_equals_batch(gpu, world):
    """Pulses in one sequence per tile (max_segments=1): the same bits alone, in a mix and in the
    one-shot batch kernel."""
    sc, ds, ds2 = world
    reqs = patch_requests(sc, 8, seed=3, w_range=(16, 80), h_range=(8, 40), sint_range=(500, 4000))
    dsets = [ds if k % 3 else ds2 for k in range(len(reqs))]
    print("case serve mix of", len(reqs), flush=True)
    srv = Server(quantum_ms=2.0, max_segments=1)
    alone, hs = _serve_alone_and_mixed(srv, dsets, reqs)
    rounds = [h.info()["rounds"] for h in hs]
    assert all(h.info()["segments"] == 1 for h in hs)
    for (a, ac), h in zip(alone, hs):
        b, bc = h.result(with_counts=True)
        assert np.array_equal(ac, bc) and np.array_equal(a, b)
    # the one-shot batch kernel on the same requests gives the same bits
    for d in (ds, ds2):
        idx = [k for k in range(len(reqs)) if dsets[k] is d]
        b = Batch(d, [reqs[k] for k in idx])
        b.run(64, 64)
        for j, k in enumerate(idx):
            assert np.array_equal(b.result(j), alone[k][0])
        b.close()
    st = srv.stats()
    assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
    print("rounds per request in the mix", rounds, st, flush=True)
    srv.close()


def test_segmented_requests(gpu, world):
    """Small requests with their tiles' pulses split into segments (the default): the same bits
    alone and in a mix; contributor counts equal to the sequential run's; FP64 results within the
    parity bound of the oracle, split results within the split tolerances of the sequential ones."""
    sc, ds, ds2 = world
    reqs = patch_requests(sc, 8, seed=4, w_range=(16, 80), h_range=(8, 40), sint_range=(500, 4000))
    dsets = [ds if k % 3 else ds2 for k in range(len(reqs))]
    print("case serve segmented mix of", len(reqs), flush=True)
    srv = Server(quantum_ms=2.0)
    alone, hs = _serve_alone_and_mixed(srv, dsets, reqs)
    segs = [h.info()["segments"] for h in hs]
    rounds = [h.info()["rounds"] for h in hs]
    print("segments", segs, "rounds", rounds, flush=True)
    assert max(segs) > 1 and all(r == 1 for r, s in zip(rounds, segs) if s > 1)
    for (a, ac), h in zip(alone, hs):
        b, bc = h.result(with_counts=True)
        assert np.array_equal(ac, bc) and np.array_equal(a, b)
    srv.close()
    seq = Server(quantum_ms=2.0, max_segments=1)
    for k, (d, r) in enumerate(zip(dsets, reqs)):
        h = seq.submit(d, r)
        assert h.wait(timeout=60)
        ref, rc = h.result(with_counts=True)
        got, gc = alone[k]
        assert np.array_equal(gc, rc)
        if r.precision == "f64":
            raw = sc.raw if d is ds else np.asfortranarray(sc.raw * np.complex64(1j))
            o, oc = O.focus_flat(r.x, r.y, r.z, r.mask, r.index, r.sint, sc.tx, sc.ty, sc.tz, raw, sc.near, sc.dra,
                                 sc.lam)
            assert np.array_equal(gc, oc)
            m = oc > 0
            rel = np.abs(got[m] - o[m]) / np.abs(o[m])
            print("request", k, "segments", segs[k], "f64 max rel err vs oracle", float(rel.max()),
                  "vs sequential", float((np.abs(got[m] - ref[m]) / np.abs(ref[m])).max()), flush=True)
            assert rel.max() <= 1e-10
        else:
            e = metrics.split_error_report(got, ref)
            print("request", k, "segments", segs[k], "split vs sequential", e, flush=True)
            assert e["phase_rms_rad_rel1e-5max"] <= 1e-4 and e["amp_rel_rmse"] <= 1e-5
    seq.close()


def test_preempt_and_resume_is_bit_identical(gpu, world):
    sc, ds, _ = world
    (req,) = patch_requests(sc, 1, seed=5, w_range=(200, 200), h_range=(64, 64), sint_range=(3800, 3800),
                            precisions=("split",))
    srv = Server(quantum_ms=0.001, pulse_chunk=64, max_segments=1)  # one 64-pulse chunk per round: many rounds
    ref = srv.submit(ds, req)
    assert ref.wait(timeout=120)
    h = srv.submit(ds, req)
    t0 = time.time()
    while h.info()["rounds"] < 3 and time.time() - t0 < 60:
        time.sleep(0.001)
    h.preempt()
    assert not h.wait(timeout=60)  # returns at the round boundary where the state was saved
    inf = h.info()
    print("case preempt", inf, flush=True)
    assert inf["state"] == _lib.REQ_PREEMPTED and inf["preemptions"] == 1
    h.resume()
    assert h.wait(timeout=120)
    assert np.array_equal(h.result(), ref.result())
    assert h.info()["rounds"] > inf["rounds"]
    srv.close()


def test_budget_admission_and_rejection(gpu, world):
    sc, ds, _ = world
    reqs = patch_requests(sc, 4, seed=8, w_range=(64, 64), h_range=(32, 32), sint_range=(1000, 1000))
    per_req = 64 * 32 * (8 * 3 + 1 + 4 * 2 + 16 + 4 + 32) + 8 * 256
    srv = Server(hbm_budget_bytes=int(per_req * 1.5), quantum_ms=1.0, max_segments=1)  # one request at a time
    hs = [srv.submit(ds, r) for r in reqs]
    for h in hs:
        assert h.wait(timeout=120)
    admits = sorted((h.info()["admit_s"], h.info()["gpu_done_s"], h.info()["done_s"]) for h in hs)
    for (a0, g0, d0), (a1, g1, d1) in zip(admits, admits[1:]):
        # the next request is admitted only after the previous one left the device (its last round
        # ended); its results may still be on the way to the caller's buffers
        assert a1 >= g0 - 1e-6
        assert g0 <= d0
    big = Request(*(np.tile(a, 4) for a in (reqs[0].x, reqs[0].y, reqs[0].z, reqs[0].mask, reqs[0].index,
                                           reqs[0].sint)), dimx=64)
    with pytest.raises(_lib.SarbpError):
        srv.submit(ds, big)
    srv.close()


def test_split_falls_back_to_fp64_when_estimate_exceeds_limit(gpu, world):
    sc, ds, _ = world
    # a coarse 20 m grid: tile rows span hundreds of metres, the split estimate is far above 1e-3 rad
    sub = sc.sub(0, 512, 0, 250)
    pick = (slice(0, 512, 20), slice(0, 250, 20))
    f = lambda a: np.asfortranarray(a[pick]).flatten(order="F")  # noqa: E731
    nx = len(range(0, 512, 20))
    req = Request(f(sub.x), f(sub.y), f(sub.z), f(sub.mask), f(sub.index), f(sub.sint), dimx=nx, precision="split")
    fine = patch_requests(sc, 1, seed=9, w_range=(64, 64), h_range=(16, 16), precisions=("split",))[0]
    srv = Server()
    h1, h2 = srv.submit(ds, req), srv.submit(ds, fine)
    assert h1.wait(60) and h2.wait(60)
    i1, i2 = h1.info(), h2.info()
    print("estimates", i1["phase_err_est_rad"], i2["phase_err_est_rad"], flush=True)
    assert i1["precision"] == _lib.PREC_F64 and i1["phase_err_est_rad"] > 1e-3
    assert ds.split_error_estimate(req.x, req.y, req.z, req.mask, req.index, req.sint, dimx=nx) == i1["phase_err_est_rad"]
    assert i2["precision"] == _lib.PREC_SPLIT and i2["phase_err_est_rad"] < 1e-3
    srv.close()


def test_split_estimate_bounds_measured_phase_error(gpu, world):
    """The a-priori estimate behind the FP64 fallback against the measured RMS phase error of the
    split result (sarbp.metrics.split_error_report against the FP64 result of the same pixels), on
    grids 1 m to 6 m apart along x: the tile rows' span, hence the range-offset term, grows with the
    spacing."""
    sc, ds, _ = world
    srv = Server(phase_err_limit_rad=1e9)  # no fallback: every split request runs split
    for step in (1, 2, 4, 6):
        sub = sc.sub(0, 512, 100, 164)
        pick = (slice(0, 512, step), slice(0, 64))
        f = lambda a: np.asfortranarray(a[pick]).flatten(order="F")  # noqa: E731
        nx = len(range(0, 512, step))
        args = [f(a) for a in (sub.x, sub.y, sub.z, sub.mask, sub.index, sub.sint)]
        print("case estimate grid step", step, "m, dimx", nx, flush=True)
        h1 = srv.submit(ds, Request(*args, dimx=nx, precision="split"))
        h2 = srv.submit(ds, Request(*args, dimx=nx, precision="f64"))
        assert h1.wait(60) and h2.wait(60)
        est = h1.info()["phase_err_est_rad"]
        assert est == ds.split_error_estimate(*args, dimx=nx)  # the same estimate for one-shot callers
        assert h1.info()["precision"] == _lib.PREC_SPLIT
        e = metrics.split_error_report(h1.result(), h2.result())
        print("step", step, "estimate", est, "measured", e, flush=True)
        assert e["phase_rms_rad_rel1e-5max"] <= est
        assert e["phase_rms_rad_rel1e-5max"] <= 1e-3
    srv.close()


def test_preempt_after_short_tiles_finished(gpu, world):
    """Tiles of unequal union length: the short ones finish (and write their means) rounds before the
    long ones.  Preempted after that and resumed, the request must still return every pixel, equal
    to the uninterrupted run."""
    sc, ds, _ = world
    # 256 x 160 px (enough pixels that a round's pulse span is one 8-pulse chunk): the left half
    # with 300-pulse apertures (tile unions ~450 pulses, done after ~60 rounds), the right half with
    # 4000-pulse ones (~520 rounds)
    sub = sc.sub(100, 356, 40, 200)
    f = lambda a: np.asarray(a).flatten(order="F")  # noqa: E731
    sint = np.where(np.arange(256)[:, None] < 128, 300, 4000) * np.ones((1, 160), np.int64)
    for prec in ("split", "f64"):
        req = Request(f(sub.x), f(sub.y), f(sub.z), f(sub.mask), f(sub.index), f(sint.astype(np.int32)), dimx=256,
                      precision=prec)
        srv = Server(quantum_ms=0.001, pulse_chunk=8, max_segments=1)  # 8 pulses per round
        ref = srv.submit(ds, req)
        assert ref.wait(timeout=120)
        h = srv.submit(ds, req)
        t0 = time.time()
        while h.info()["rounds"] < 70 and time.time() - t0 < 60:
            time.sleep(0.0002)
        h.preempt()
        assert not h.wait(timeout=60)
        inf = h.info()
        print("case preempt after short tiles,", prec, inf, flush=True)
        assert inf["state"] == _lib.REQ_PREEMPTED and 60 < inf["rounds"] < 400
        h.resume()
        assert h.wait(timeout=120)
        got, cnt = h.result(with_counts=True)
        want, wcnt = ref.result(with_counts=True)
        assert np.array_equal(cnt, wcnt) and np.array_equal(got, want), prec
        assert np.all(cnt[(f(sint) == 300) & (f(sub.mask) == 1)] > 0)  # the short tiles' means were kept
        srv.close()


def test_request_lifetime(gpu, world):
    """The server keeps a request's buffers and dataset alive until it is done, even when the caller
    drops its handle; a dataset with unfinished server requests refuses to be destroyed."""
    import gc
    sc, _, _ = world
    own = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    reqs = patch_requests(sc, 3, seed=11, w_range=(64, 128), h_range=(32, 64), sint_range=(2000, 4000))
    srv = Server(quantum_ms=0.001, pulse_chunk=8, max_segments=1)
    hs = [srv.submit(own, r) for r in reqs]
    hs[-1].preempt()  # preempted (queued or running): it may be resumed, so it holds the dataset
    assert not hs[-1].wait(timeout=60)
    with pytest.raises(_lib.SarbpError) as ei:
        own.close()
    assert ei.value.status == _lib.ESTATE
    rid = hs[0].id
    del hs[0]  # the caller drops a handle before completion
    gc.collect()
    for h in hs[:-1]:
        assert h.wait(timeout=120)
    while srv.reap():
        time.sleep(0.01)
        if srv._inflight and all(h.state() == _lib.REQ_PREEMPTED for h in srv._inflight.values()):
            break
    assert list(srv._inflight) == [hs[-1].id] and rid not in srv._inflight
    with pytest.raises(_lib.SarbpError):
        own.close()
    hs[-1].resume()
    assert hs[-1].wait(timeout=120)
    own.close()  # every request is done: the dataset goes
    srv.close()


def test_split_request_beyond_lds_is_refused_alone(gpu, world):
    """A split request whose tile rows x pulse chunk exceed the batch kernel's LDS is refused at
    submission; the other requests run; a server whose chunk cannot hold the default split tile is
    refused at creation."""
    sc, ds, _ = world
    with pytest.raises(ValueError):
        Server(pulse_chunk=512)  # 512 pulses x 4 rows x 48 B > 64 KiB
    srv = Server(pulse_chunk=256)  # 4 rows: 48 KiB; 16 rows (tile_x 4): 192 KiB
    good = patch_requests(sc, 2, seed=12, w_range=(32, 64), h_range=(16, 32))
    tall = Request(good[0].x, good[0].y, good[0].z, good[0].mask, good[0].index, good[0].sint, dimx=good[0].dimx,
                   precision="split", tile_x=4)
    with pytest.raises(ValueError):
        srv.submit(ds, tall)
    hs = [srv.submit(ds, r) for r in good]
    for h in hs:
        assert h.wait(timeout=60)
    assert srv.stats()["failed"] == 0
    srv.close()


def test_release_request_records(gpu, world):
    """A finished request's record is dropped by sarbp_server_release (its id becomes unknown); an
    unfinished one's release is refused; a Handle dropped after completion releases its record."""
    import gc
    sc, ds, _ = world
    reqs = patch_requests(sc, 3, seed=13, w_range=(64, 128), h_range=(32, 64), sint_range=(2000, 3000))
    srv = Server(quantum_ms=0.001, pulse_chunk=8, max_segments=1)  # many rounds: still running below
    L = _lib.load()
    hs = [srv.submit(ds, r) for r in reqs]
    st = L.sarbp_server_release(srv._h, hs[0].id)
    assert st == _lib.ESTATE, st  # queued or running
    assert all(h.wait(timeout=120) for h in hs)
    assert srv.stats()["records"] == 3
    assert L.sarbp_server_release(srv._h, hs[1].id) == _lib.OK
    with pytest.raises(ValueError, match="unknown request id"):
        hs[1].info()
    assert L.sarbp_server_release(srv._h, hs[1].id) == _lib.EINVAL  # unknown id now
    assert srv.stats()["records"] == 2
    want = hs[2].result().copy()
    del hs[2]
    gc.collect()
    assert srv.stats()["records"] == 1  # released with its handle
    assert np.all(np.isfinite(want))
    srv.close()
