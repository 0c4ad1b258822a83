"""Serving: continuous batching of focusing requests on device-resident datasets (sarbp_server_*).

:class:`Server` wraps the library's scheduling thread; :func:`poisson_run` drives it with Poisson
arrivals of sub-scene requests and reports p50/p99 latency (completion - arrival, queueing
included) and aggregate pixel-pulse pairs per second.

    python -m sarbp.serve --rates 5,20,50 --duration 60
    python -m sarbp.serve --rates 50,100,200,400,800,1600 --duration 20 --pool-per-dataset 256   # load sweep
    python -m sarbp.serve --sweep          # isolated requests by size: latency phases
"""
from __future__ import annotations

import argparse
import collections
import ctypes
import json
import sys
import threading
import time

import numpy as np

from . import _lib
from ._lib import EAGAIN, check
from .batch import Request


class Handle:
    """A submitted request.  It holds the dataset, the request's arrays and the result buffers the
    library writes into; the server keeps every handle of an unfinished request alive, so dropping
    one before completion is safe.  When the last reference to the handle of a finished request
    goes, the server's record of it is released (sarbp_server_release)."""

    def __init__(self, server, rid, ds, req: Request, out, cnt, t_submit):
        self.server, self.id, self.ds, self.req = server, rid, ds, req
        self.out, self.cnt, self.t_submit = out, cnt, t_submit

    def wait(self, timeout=None) -> bool:
        """True when the request is done; False on timeout or when it was preempted (see state());
        raises when it failed."""
        st = _lib.load().sarbp_server_wait(self.server._h, self.id, -1.0 if timeout is None else float(timeout))
        if st == EAGAIN:
            return False
        if st == _lib.ESTATE and self.state() == _lib.REQ_PREEMPTED:
            return False
        self.server._settled(self.id)
        check(st, "sarbp_server_wait")
        return True

    def state(self) -> int:
        v = ctypes.c_int32()
        check(_lib.load().sarbp_server_poll(self.server._h, self.id, ctypes.byref(v)), "sarbp_server_poll")
        if v.value in (_lib.REQ_DONE, _lib.REQ_FAILED):
            self.server._settled(self.id)
        return int(v.value)

    def info(self) -> dict:
        i = _lib.RequestInfo()
        check(_lib.load().sarbp_server_info(self.server._h, self.id, ctypes.byref(i)), "sarbp_server_info")
        return {k: getattr(i, k) for k, _ in _lib.RequestInfo._fields_}

    def preempt(self):
        check(_lib.load().sarbp_server_preempt(self.server._h, self.id), "sarbp_server_preempt")

    def resume(self):
        check(_lib.load().sarbp_server_resume(self.server._h, self.id), "sarbp_server_resume")

    def result(self, with_counts=False):
        return (self.out, self.cnt) if with_counts else self.out

    def __del__(self):
        try:  # at interpreter exit the module's globals may already be gone
            h = getattr(self.server, "_h", None)
            if h and _lib.alive():
                _lib.load().sarbp_server_release(h, self.id)  # refused (ignored) unless finished
        except Exception:
            pass


class Server:
    """sarbp_server_*: zero arguments take the library defaults (200 GB budget, 20 ms rounds, 256
    active requests, 32-pulse chunks, 1e-3 rad split fallback limit, up to 16 pulse segments per
    tile for requests too small to fill the device; max_segments=1 runs every tile's pulses in one
    sequence, as the one-shot batch kernel does)."""
    _close_order = 0  # servers, then batches/scenes, then datasets
    def __init__(self, hbm_budget_bytes=0, quantum_ms=0.0, max_active=0, pulse_chunk=0, phase_err_limit_rad=0.0,
                 max_segments=0):
        _lib.require_device()
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), float(quantum_ms), int(max_active), int(pulse_chunk),
                                float(phase_err_limit_rad), int(max_segments))
        h = ctypes.c_void_p()
        check(_lib.load().sarbp_server_create(ctypes.byref(cfg), ctypes.byref(h)), "sarbp_server_create")
        self._h = h
        self._inflight = {}  # id -> Handle of every request not yet seen done or failed
        self._since_reap = 0
        self._lock = threading.Lock()  # submit may be called from several threads
        _lib.track(self)

    def submit(self, ds, req: Request) -> Handle:
        out = np.zeros(req.n, np.complex128)
        cnt = np.zeros(req.n, np.int32)
        rid = ctypes.c_int64()
        st = req.c_struct()
        t = time.perf_counter()
        check(_lib.load().sarbp_server_submit(self._h, ds.handle, ctypes.byref(st), out.ctypes.data, cnt.ctypes.data,
                                              ctypes.byref(rid)), "sarbp_server_submit")
        hd = Handle(self, int(rid.value), ds, req, out, cnt, t)
        with self._lock:
            self._inflight[hd.id] = hd
            self._since_reap += 1
            reap = self._since_reap >= 64
        if reap:
            self.reap()
        return hd

    def _settled(self, rid):
        with self._lock:
            self._inflight.pop(rid, None)

    def reap(self) -> int:
        """Release the server's references to requests that have finished (done or failed);
        returns the number still in flight."""
        with self._lock:
            self._since_reap = 0
            ids = list(self._inflight)
        L = _lib.load()
        v = ctypes.c_int32()
        for rid in ids:
            if L.sarbp_server_poll(self._h, rid, ctypes.byref(v)) == _lib.OK and v.value in (_lib.REQ_DONE,
                                                                                            _lib.REQ_FAILED):
                self._settled(rid)
        with self._lock:
            return len(self._inflight)

    def stats(self) -> dict:
        s = _lib.ServerStats()
        check(_lib.load().sarbp_server_get_stats(self._h, ctypes.byref(s)), "sarbp_server_get_stats")
        return {k: getattr(s, k) for k, _ in _lib.ServerStats._fields_}

    def close(self):
        """Finish the current round, fail what is still queued or running, stop the server."""
        h = getattr(self, "_h", None)
        if h:
            _lib.load().sarbp_server_destroy(h)
            self._h = None
        self._inflight = {}  # the library no longer writes into any request's buffers

    def __del__(self):
        try:  # at interpreter exit the module's globals may already be gone
            if _lib.alive():
                self.close()
        except Exception:
            pass


def percentile(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else None


def poisson_run(server: Server, datasets, pool, rate_hz: float, duration_s: float, seed: int = 0,
                warmup_s: float = 2.0, log=None, submit_threads: int = 4):
    """Submit requests drawn from ``pool`` (list of (dataset index, Request)) at the times of a seeded
    Poisson schedule of rate ``rate_hz`` for ``duration_s``; wait for all.

    ``submit_threads`` driver threads take the scheduled arrivals in order, each sleeping until just
    before its next arrival, yielding until the arrival and submitting it (the library copies the pixels in the caller's thread, so one
    thread caps the rate a driver can offer).  Latency is completion minus ARRIVAL, the request's
    time in the Poisson schedule: it includes any delay before the driver got to submit it (a driver
    that falls behind the schedule shows up as latency), the submission itself and the queueing.
    Completion times are on the server's clock (``done_s``); the driver's clock is mapped onto it by
    the smallest observed difference between a request's ``submit_s`` (server clock, entry to
    submit) and the driver's time just before the call.  The same statistics from the entry to
    submit are reported as ``submit_latency_*``.  Finished requests are harvested as the run goes
    (their result buffers released), so a long run holds only the requests in flight."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t_next = 0.0
    while t_next < duration_s:
        arrivals.append(t_next)
        t_next += rng.exponential(1.0 / rate_hz)
    slots = [None] * len(arrivals)  # (driver time at the call, Handle) once submitted
    done = []                       # (scheduled arrival offset, driver time at the call, info)
    errors = []
    lock = threading.Lock()
    nxt = [0]
    go = threading.Event()
    t0 = 0.0

    def submitter():
        go.wait()
        while not errors:
            with lock:
                j = nxt[0]
                nxt[0] += 1
            if j >= len(arrivals):
                return
            t_arr = t0 + arrivals[j]
            wait = t_arr - time.perf_counter()
            if wait > 0.002:
                time.sleep(wait - 0.0015)  # a sleeping thread wakes up to ~1 ms late
            while time.perf_counter() < t_arr:
                time.sleep(0)  # the last stretch in yields, which release the interpreter lock
            di, req = pool[j % len(pool)]
            t_call = time.perf_counter()
            try:
                slots[j] = (t_call, server.submit(datasets[di], req))
            except Exception as e:  # noqa: BLE001 - re-raised in the calling thread
                errors.append(e)
                return

    threads = [threading.Thread(target=submitter, daemon=True) for _ in range(max(1, int(submit_threads)))]
    for t in threads:
        t.start()
    st0 = server.stats()
    t0 = time.perf_counter()  # the schedule's origin: set before any submitter reads it
    go.set()
    # harvest in arrival order: a request is released once it and every earlier one finished
    j, last_log = 0, t0
    while j < len(arrivals):
        if errors:
            break
        sl = slots[j]
        if sl is not None and sl[1].state() in (_lib.REQ_DONE, _lib.REQ_FAILED):
            done.append((arrivals[j], sl[0], sl[1].info()))
            slots[j] = None
            j += 1
            continue
        if log and time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            log(f"rate {rate_hz}/s: {nxt[0]} submitted, {j} finished, t={last_log - t0:.0f}s, {server.stats()}")
        if sl is not None and all(not t.is_alive() for t in threads):
            sl[1].wait()  # every submission made: block on the oldest
        else:
            time.sleep(0.001)
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    t_end = time.perf_counter() - t0
    st1 = server.stats()
    # driver clock -> server clock
    offset = min((inf["submit_s"] - t_call for _, t_call, inf in done), default=0.0)
    lat, lat_submit, lag, pp, prec = [], [], [], 0, {"f64": 0, "split": 0}
    first_submit, last_done = float("inf"), float("-inf")
    for t_arr, t_call, inf in done:
        if t_arr >= warmup_s:
            lat.append(inf["done_s"] - (t0 + t_arr + offset))
            lat_submit.append(inf["done_s"] - inf["submit_s"])
            lag.append(t_call - (t0 + t_arr))
        pp += inf["pp"]
        first_submit = min(first_submit, inf["submit_s"])
        last_done = max(last_done, inf["done_s"])
        prec["split" if inf["precision"] == _lib.PREC_SPLIT else "f64"] += 1
    span = max(1e-9, last_done - first_submit) if done else 1e-9
    ms = lambda v, q: percentile(v, q) * 1e3 if v else None  # noqa: E731
    rounds = st1.get("rounds", 0) - st0.get("rounds", 0)
    busy = st1.get("gpu_busy_s", 0.0) - st0.get("gpu_busy_s", 0.0)
    return {
        "rate_req_per_s": rate_hz, "duration_s": duration_s, "requests": len(done),
        "latency_p50_ms": ms(lat, 50), "latency_p99_ms": ms(lat, 99),
        "latency_max_ms": max(lat) * 1e3 if lat else None,
        "submit_latency_p50_ms": ms(lat_submit, 50), "submit_latency_p99_ms": ms(lat_submit, 99),
        "arrival_lag_p99_ms": ms(lag, 99),
        "pp_total": int(pp), "pp_per_s": pp / span, "wall_s": t_end, "precision_counts": prec,
        "rounds": int(rounds), "rounds_per_request": rounds / max(1, len(done)),
        "gpu_busy_frac": busy / span,
        # the scheduling thread's time, as fractions of the span: admissions, building rounds, rounds
        **{f"{k}_frac": (st1.get(k, 0.0) - st0.get(k, 0.0)) / span for k in ("admit_s", "build_s", "round_wall_s")},
        "server": st1,
    }


def size_sweep(server: Server, ds, scene, sizes=((8, 8), (128, 64), (256, 128), (600, 144), (1024, 250)),
               precisions=("split", "f64"), reps=4, x0=1000, y0=0):
    """Latency of isolated requests by size: each (width, height) patch of ``scene`` at (x0, y0) is
    submitted ``reps`` times, one at a time; the last repetition's phases are reported (ms):
    submission to admission, admission to its round, the round (work list, kernels, result copies),
    and the whole latency, with every repetition's latency and round."""
    rows = []
    nx, ny = scene.shape
    for (w, h) in sizes:
        w, h = min(w, nx - x0), min(h, ny - y0)
        sub = scene.sub(x0, x0 + w, y0, y0 + h)
        f = lambda a: np.asarray(a).flatten(order="F")  # noqa: E731
        for prec in precisions:
            req = Request(f(sub.x), f(sub.y), f(sub.z), f(sub.mask), f(sub.index), f(sub.sint), dimx=w,
                          precision=prec)
            infos = []
            for _ in range(reps):
                hd = server.submit(ds, req)
                hd.wait()
                infos.append(hd.info())
                time.sleep(0.002)
            i = infos[-1]
            ms = lambda a, b, i=i: (i[b] - i[a]) * 1e3  # noqa: E731
            rows.append({"w": w, "h": h, "precision": prec, "pp": i["pp"], "segments": i["segments"],
                         "submit_to_admit_ms": ms("submit_s", "admit_s"), "admit_to_round_ms": ms("admit_s", "round_s"),
                         "round_ms": ms("round_s", "gpu_done_s"), "latency_ms": ms("submit_s", "done_s"),
                         "latency_ms_each": [(q["done_s"] - q["submit_s"]) * 1e3 for q in infos],
                         "round_ms_each": [(q["gpu_done_s"] - q["round_s"]) * 1e3 for q in infos]})
    return rows


def make_pool(scenes, nreq_per_dataset=32, seed=0):
    """Sub-scene requests (width 128-1024, height 32-250, Sint 500-4000, split:FP64 = 1:1)."""
    from .batch import patch_requests
    pool = []
    for di, sc in enumerate(scenes):
        for r in patch_requests(sc, nreq_per_dataset, seed=seed + di):
            pool.append((di, r))
    rng = np.random.default_rng(seed)
    rng.shuffle(pool)
    return pool


def main(argv=None):
    ap = argparse.ArgumentParser(description="Poisson serving benchmark")
    ap.add_argument("--rates", default="5,20,50")
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--datasets", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--quantum-ms", type=float, default=20.0)
    ap.add_argument("--max-segments", type=int, default=0, help="pulse segments per tile of small requests "
                    "(0: library default; 1: off)")
    ap.add_argument("--pool-per-dataset", type=int, default=32, help="distinct requests per dataset in the pool")
    ap.add_argument("--submit-threads", type=int, default=4, help="driver threads submitting the arrivals")
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--sweep", action="store_true", help="latency of isolated requests by size instead")
    ap.add_argument("--sweep-reps", type=int, default=4, help="repetitions of each request of the size sweep")
    a = ap.parse_args(argv)
    from . import synth
    from .core import Dataset
    log = lambda m: print(f"[serve] {m}", file=sys.stderr, flush=True)  # noqa: E731
    base = synth.full_scene(seed=1, **({"nx": 512, "ny": 250, "naz": 5121, "ntarget": 200} if a.small else {}))
    log(f"scene {base.name} ready")
    if a.sweep:
        ds = Dataset(base.raw, base.tx, base.ty, base.tz, base.near, base.dra, base.lam)
        srv = Server(hbm_budget_bytes=int(a.budget_gb * 1e9), quantum_ms=a.quantum_ms, max_segments=a.max_segments)
        rows = size_sweep(srv, ds, base, x0=min(1000, base.shape[0] // 4), reps=a.sweep_reps)
        srv.close()
        ds.close()
        for r in rows:
            log(json.dumps(r))
        print(json.dumps({"sweep": rows}))
        return
    datasets, scenes = [], []
    for k in range(a.datasets):
        raw = base.raw if k == 0 else np.asfortranarray(base.raw * np.complex64(np.exp(1j * 0.7 * k)))
        datasets.append(Dataset(raw, base.tx, base.ty, base.tz, base.near, base.dra, base.lam))
        scenes.append(base)
        log(f"dataset {k} resident")
    pool = make_pool(scenes, nreq_per_dataset=a.pool_per_dataset)
    log(f"pool of {len(pool)} distinct requests")
    srv = Server(hbm_budget_bytes=int(a.budget_gb * 1e9), quantum_ms=a.quantum_ms, max_segments=a.max_segments)
    results = []
    for rate in [float(r) for r in a.rates.split(",") if r]:
        res = poisson_run(srv, datasets, pool, rate, a.duration, log=log, submit_threads=a.submit_threads)
        log(json.dumps(res))
        results.append(res)
    srv.close()
    print(json.dumps({"serving": results}))


if __name__ == "__main__":
    main()
