"""Host side of the Poisson serving benchmark (sarbp.serve.poisson_run) with a stand-in server whose
requests take a known time and whose clock has its own origin: arrivals follow the seeded schedule,
every request is harvested once, the warm-up is excluded from the percentiles, the latency from the
entry to submit is done_s - submit_s, the latency from the scheduled arrival maps the driver's clock
onto the server's, and the pixel-pulse rate, round count and precision counts come out of the
records as defined.
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
        c = self.server.CLOCK  # the server's clock starts elsewhere than the driver's
        return {"submit_s": self.t_submit + c, "done_s": self.t_submit + c + lat, "pp": 1000 + self.k,
                "precision": self.precision, "state": _lib.REQ_DONE}


class _Server:
    """Request k takes (k % 10 + 1) ms from the entry to submit; every 7th submission is delayed
    by 3 ms before that entry (a driver behind its schedule)."""
    CLOCK = -1234.5

    def __init__(self):
        self.n = 0
        self.t = []
        self.calls = []

    @staticmethod
    def latency(k):
        return (k % 10 + 1) * 1e-3

    def submit(self, ds, req):
        import time
        self.calls.append(time.perf_counter())
        if self.n % 7 == 3:
            time.sleep(3e-3)
        t = time.perf_counter()
        self.t.append(t)
        h = _Handle(self, self.n, t, _lib.PREC_SPLIT if req == "s" else _lib.PREC_F64)
        self.n += 1
        return h

    def stats(self):
        return {"submitted": self.n, "rounds": 2 * self.n, "gpu_busy_s": 1e-3 * self.n}


def test_poisson_run_statistics():
    srv = _Server()
    pool = [(0, "s"), (0, "d")]
    res = serve.poisson_run(srv, [None], pool, rate_hz=400.0, duration_s=0.5, seed=5, warmup_s=0.1,
                            submit_threads=1)
    n = res["requests"]
    assert n == srv.n and n > 100  # ~200 arrivals expected
    # the seeded schedule: the k-th arrival at the sum of the first k exponential gaps
    rng = np.random.default_rng(5)
    offs, t = [], 0.0
    while t < 0.5:
        offs.append(t)
        t += rng.exponential(1.0 / 400.0)
    assert n == len(offs)
    keep = [k for k in range(n) if offs[k] >= 0.1]
    lat = [(k % 10 + 1) * 1e-3 for k in keep]
    assert res["submit_latency_p50_ms"] == pytest.approx(np.percentile(lat, 50) * 1e3)
    assert res["submit_latency_p99_ms"] == pytest.approx(np.percentile(lat, 99) * 1e3)
    # from the scheduled arrival: the stand-in's latency plus the time from the schedule to the entry
    # to submit (the driver's own lag and the 3 ms delays); the clock offset is estimated from the
    # fastest call, so each value is exact to within that call's overhead
    # the schedule's origin: no call precedes its scheduled time and most are made on time (to within
    # microseconds), so the earliest call relative to its arrival gives it
    t0 = min(c - o for c, o in zip(srv.calls, offs))
    full = [srv.t[k] - (t0 + offs[k]) + (k % 10 + 1) * 1e-3 for k in keep]
    assert res["latency_p50_ms"] == pytest.approx(np.percentile(full, 50) * 1e3, abs=0.3)
    assert res["latency_p99_ms"] == pytest.approx(np.percentile(full, 99) * 1e3, abs=0.3)
    assert res["latency_max_ms"] == pytest.approx(max(full) * 1e3, abs=0.3)
    assert res["latency_max_ms"] >= 3.0 + 1.0  # a delayed submission is counted from its arrival
    assert res["rounds"] == 2 * n and res["rounds_per_request"] == pytest.approx(2.0)
    assert res["precision_counts"] == {"split": (n + 1) // 2, "f64": n // 2}
    assert res["pp_total"] == sum(1000 + k for k in range(n))
    last_done = max(srv.t[k] + (k % 10 + 1) * 1e-3 for k in range(n))
    assert res["pp_per_s"] == pytest.approx(res["pp_total"] / (last_done - srv.t[0]), rel=1e-9)
    assert res["gpu_busy_frac"] == pytest.approx(1e-3 * n / (last_done - srv.t[0]), rel=1e-9)


class _SlowServer:
    """Each submission takes 2 ms in the caller's thread (the library copies the pixels there);
    requests finish 1 ms after the entry to submit."""

    def __init__(self):
        import threading
        self.lock = threading.Lock()
        self.seen = {}  # arrival index -> driver time at the entry to submit

    def submit(self, ds, req):
        import time
        t = time.perf_counter()
        with self.lock:
            assert req not in self.seen
            self.seen[req] = t
        time.sleep(2e-3)
        return _Handle2(req, t)

    def stats(self):
        return {"rounds": 0, "gpu_busy_s": 0.0}


class _Handle2:
    def __init__(self, j, t):
        self.j, self.t = j, t

    def wait(self, timeout=None):
        return True

    def state(self):
        return _lib.REQ_DONE

    def info(self):
        return {"submit_s": self.t, "done_s": self.t + 1e-3, "pp": self.j, "precision": _lib.PREC_F64}


def test_poisson_run_threads_keep_up():
    """At 1000 arrivals/s and 2 ms per submission one driver thread falls behind by ~0.5 s over the
    run; eight threads submit every arrival exactly once, none before its scheduled time, and keep
    the lag from the schedule to a few ms."""
    import time
    rng = np.random.default_rng(3)
    offs, t = [], 0.0
    while t < 0.5:
        offs.append(t)
        t += rng.exponential(1.0 / 1000.0)
    pool = [(0, j) for j in range(len(offs))]  # request j is arrival j
    srv = _SlowServer()
    t_before = time.perf_counter()
    res = serve.poisson_run(srv, [None], pool, rate_hz=1000.0, duration_s=0.5, seed=3, warmup_s=0.0,
                            submit_threads=8)
    assert res["requests"] == len(offs) and sorted(srv.seen) == list(range(len(offs)))
    assert res["pp_total"] == sum(range(len(offs)))
    t0 = min(srv.seen.values()) - offs[0]
    assert t0 >= t_before - 1e-3
    assert all(srv.seen[j] >= t0 + offs[j] - 2e-3 for j in range(len(offs)))
    assert res["arrival_lag_p99_ms"] < 50.0
    srv1 = _SlowServer()
    one = serve.poisson_run(srv1, [None], pool, rate_hz=1000.0, duration_s=0.5, seed=3, warmup_s=0.0,
                            submit_threads=1)
    assert one["requests"] == len(offs) and one["arrival_lag_p99_ms"] > 200.0


def test_percentile_of_empty_is_none():
    assert serve.percentile([], 50) is None
    assert serve.percentile([1.0, 3.0], 50) == 2.0
