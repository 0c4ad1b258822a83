"""Host side of the Poisson serving benchmark (sarbp.serve.poisson_run) with a stand-in server whose
requests take a known time: arrivals follow the seeded schedule, every request is harvested once,
the warm-up is excluded from the percentiles, and p50/p99/max, the pixel-pulse rate and the
precision counts come out of the request records as defined (latency = done_s - submit_s).
"""
import numpy as np
import pytest

from sarbp import _lib, serve


class _Handle:
    def __init__(self, server, k, t_submit, precision):
        self.server, self.k, self.t_submit, self.precision = server, k, t_submit, precision

    def wait(self, timeout=None):
        return True

    def state(self):
        return _lib.REQ_DONE

    def info(self):
        lat = self.server.latency(self.k)
        return {"submit_s": self.t_submit, "done_s": self.t_submit + lat, "pp": 1000 + self.k,
                "precision": self.precision, "state": _lib.REQ_DONE}


class _Server:
    """Request k takes (k % 10 + 1) ms; it 'arrives' at the wall time of its submission."""

    def __init__(self):
        self.n = 0
        self.t = []

    @staticmethod
    def latency(k):
        return (k % 10 + 1) * 1e-3

    def submit(self, ds, req):
        import time
        t = time.perf_counter()
        self.t.append(t)
        h = _Handle(self, self.n, t, _lib.PREC_SPLIT if req == "s" else _lib.PREC_F64)
        self.n += 1
        return h

    def stats(self):
        return {"submitted": self.n}


def test_poisson_run_statistics():
    srv = _Server()
    pool = [(0, "s"), (0, "d")]
    res = serve.poisson_run(srv, [None], pool, rate_hz=400.0, duration_s=0.5, seed=5, warmup_s=0.1)
    n = res["requests"]
    assert n == srv.n and n > 100  # ~200 arrivals expected
    # the seeded schedule: the k-th arrival at the sum of the first k exponential gaps
    rng = np.random.default_rng(5)
    offs, t = [], 0.0
    while t < 0.5:
        offs.append(t)
        t += rng.exponential(1.0 / 400.0)
    assert n == len(offs)
    lat = [(k % 10 + 1) * 1e-3 for k in range(n) if offs[k] >= 0.1]
    assert res["latency_p50_ms"] == pytest.approx(np.percentile(lat, 50) * 1e3)
    assert res["latency_p99_ms"] == pytest.approx(np.percentile(lat, 99) * 1e3)
    assert res["latency_max_ms"] == pytest.approx(max(lat) * 1e3)
    assert res["precision_counts"] == {"split": (n + 1) // 2, "f64": n // 2}
    assert res["pp_total"] == sum(1000 + k for k in range(n))
    last_done = max(srv.t[k] + (k % 10 + 1) * 1e-3 for k in range(n))
    assert res["pp_per_s"] == pytest.approx(res["pp_total"] / (last_done - srv.t[0]), rel=1e-9)


def test_percentile_of_empty_is_none():
    assert serve.percentile([], 50) is None
    assert serve.percentile([1.0, 3.0], 50) == 2.0
