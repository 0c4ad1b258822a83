"""Persistent batch kernel: many uneven requests in one launch (GPU).

Each request's result must equal the same request focused alone (bit-for-bit: the per-pixel
arithmetic does not depend on the other requests or on the work mapping), and the FP64 requests
must meet the oracle's parity bound.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O
from sarbp import synth
from sarbp.batch import Batch, patch_requests
from sarbp.core import Dataset

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    return synth.full_scene(seed=6, nx=512, ny=250, naz=5121, nra=4762, ntarget=100)


def alone(ds, r):
    sc = ds.scene(r.x, r.y, r.z, r.mask, r.index, r.sint, dimx=r.dimx)
    sc.focus(r.precision)
    out = sc.result(with_counts=True)
    sc.close()
    return out


def test_batch_equals_alone_and_oracle(gpu, scene):
    reqs = patch_requests(scene, 10, seed=1, w_range=(16, 96), h_range=(8, 40), sint_range=(500, 4000))
    ds = Dataset(scene.raw, scene.tx, scene.ty, scene.tz, scene.near, scene.dra, scene.lam)
    b = Batch(ds, reqs)
    print("case batch of", len(reqs), [(r.dimx, r.n // r.dimx, r.precision, int(r.sint[0])) for r in reqs], flush=True)
    b.run(64, 64)
    for k, r in enumerate(reqs):
        got, cnt = b.result(k, with_counts=True)
        ref, rcnt = alone(ds, r)
        assert np.array_equal(cnt, rcnt), k
        assert np.array_equal(got, ref), k
        if r.precision == "f64" and k < 4:
            o, oc = O.focus_flat(r.x, r.y, r.z, r.mask, r.index, r.sint, scene.tx, scene.ty, scene.tz, scene.raw,
                                 scene.near, scene.dra, scene.lam)
            assert np.array_equal(cnt, oc)
            nz = oc > 0
            assert (np.abs(got[nz] - o[nz]) / np.abs(o[nz])).max() <= 1e-10
    # every work mapping gives the same bits
    base = [b.result(k) for k in range(len(reqs))]
    for tp, ch in ((128, 64), (256, 16), (64, 256), (128, 32)):
        print("mapping", tp, ch, flush=True)
        b.run(tp, ch)
        for k in range(len(reqs)):
            assert np.array_equal(b.result(k), base[k]), (tp, ch, k)
    t, c, ms = b.autotune(reps=1)
    assert t in (64, 128, 256) and c % 8 == 0 and ms > 0
    b.close()
    ds.close()


def test_batch_rejects_bad_mapping(gpu, scene):
    reqs = patch_requests(scene, 2, seed=2, w_range=(16, 32), h_range=(8, 16))
    ds = Dataset(scene.raw, scene.tx, scene.ty, scene.tz, scene.near, scene.dra, scene.lam)
    b = Batch(ds, reqs)
    with pytest.raises(ValueError):
        b.run(256, 1024)  # 4 x 1024 records exceed the LDS limit
    with pytest.raises(ValueError):
        b.run(96, 64)
    b.close()
    ds.close()
