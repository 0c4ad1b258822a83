"""Scenes of one dataset focused from several host threads at once (GPU).

The C ABI allows different handles to be used from different threads; two scenes share their
dataset's stream and timing events.  Results must equal the single-threaded ones bit for bit."""
import threading

import numpy as np
import pytest

from sarbp import synth
from sarbp.core import Dataset

pytestmark = pytest.mark.gpu


def test_two_threads_two_scenes_one_dataset(gpu):
    sc = synth.tiny_scene(seed=5, nx=48, ny=24)
    print(f"case tiny threads shape={sc.shape}", flush=True)
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    try:
        scenes = [ds.scene(*sc.flat(), dimx=sc.dimx) for _ in range(2)]
        precs = ("f64", "split")
        expect = []
        for s, p in zip(scenes, precs):
            s.focus(p)
            expect.append(s.result(with_counts=True))
        got = [[], []]
        errors = []

        def work(k):
            try:
                for _ in range(20):
                    ms = scenes[k].focus(precs[k])
                    assert ms >= 0.0
                    got[k].append(scenes[k].result(with_counts=True))
            except Exception as e:  # reported in the main thread
                errors.append(e)

        th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        assert not errors, errors
        for k in range(2):
            assert len(got[k]) == 20
            for out, cnt in got[k]:
                assert np.array_equal(out, expect[k][0]) and np.array_equal(cnt, expect[k][1])
        for s in scenes:
            s.close()
    finally:
        ds.close()
