#!/usr/bin/env python3
"""Throughput of the per-pixel back-projection on one MI355X.

Workload: the 4000 x 250-pixel scene over 40001 pulses (RawData 4762 x 40001 complex64, ~1.5 GB,
2000 point targets + 20 dB noise, Sint ~ U[2500,3500], mask ~ Bernoulli(0.95)), synthetic and
seeded.  One step = one focusing pass over every pixel with RawData, trajectory and pixel arrays
already resident in HBM.  Unit of work: pixel-pulse pairs (pp) evaluated, counted on the host from
the integer inputs.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "sar-backprojection_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

from sarbp.metrics import split_error_report  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BYTES_PER_PP = 8.0     # one complex64 RawData sample gathered per pixel-pulse pair
# VALU issue: a wave64 instruction occupies one quad-cycle (4 clocks) of its SIMD (FP32 sin/cos two,
# FP64 rsq four), and two adjacent independent FP32 add/mul/fma or 32-bit add/logic instructions of
# one wave can issue in the same quad-cycle (sarbp_probe_rates, profiles/valu_issue_rates__gfx950-probe.md).
# MI355X has 1024 SIMDs (256 CUs x 4) in 8 XCDs; its engine clock is at most 2.4 GHz and runs lower
# under these kernels, so the quad-cycles a launch has are taken from the clock counted in the same
# counter pass (GRBM_GUI_ACTIVE, clocks summed over the 8 XCDs).
SIMDS = 1024
XCDS = 8
PROFILE_JSON = "profiles/traffic_full-4000x250x40001.json"
KERNEL_OF = {"split": "k_bp_split_tile<0>", "f64": "k_bp_f64_tile<false>"}


def default_maps(args):
    return args.split_map == "16,0" and args.f64_map == "1,8"


def _load_profile(pp):
    """Per-launch counters of the default kernels on this workload, from the committed rocprofv3
    summary (None when absent or for another workload)."""
    fn = os.path.join(ROOT, PROFILE_JSON)
    if not os.path.exists(fn):
        return None
    with open(fn) as f:
        tj = json.load(f)
    return tj if tj.get("pp_per_launch") == pp else None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_precision(scene, precision, steps, warmup):
    for _ in range(warmup):
        scene.focus(precision)
    walls, gpus = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        g = scene.focus(precision)
        walls.append(time.perf_counter() - t0)
        gpus.append(g)
    return float(np.mean(walls)), float(np.mean(gpus)) / 1e3


def cpu_baseline(sc, strip_cols, serial_cols, gpu_strip):
    """Oracle (C restatement) on a bounded strip of the same workload, all host cores, compared
    against the GPU result of the same pixels; plus the serial (1-thread) rate on a narrower strip,
    the reference's own loop structure (sar_focuser_puntual.py:259)."""
    from oracle import oracle_lib as O
    x0 = sc.shape[0] // 2 - strip_cols // 2
    st = sc.sub(x0, x0 + strip_cols)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cnt = lambda s: O.count_pp(*[np.asarray(a).flatten(order="F") for a in (s.mask, s.index, s.sint)], s.Naz)  # noqa
    pp = cnt(st)
    t0 = time.perf_counter()
    ref, rcnt = O.focus(*st.args(), nthreads=cores)
    dt = time.perf_counter() - t0
    s1 = sc.sub(x0, x0 + serial_cols)
    pp1 = cnt(s1)
    t1 = time.perf_counter()
    O.focus(*s1.args(), nthreads=1)
    dt1 = time.perf_counter() - t1
    base = {"value": pp / dt, "unit": "pixel·pulses/s", "cores": int(cores), "kind": "port",
            "host_cpus": {"nproc": _nproc(), "os_cpu_count": os.cpu_count(), "affinity": int(cores),
                          "omp_num_threads": os.environ.get("OMP_NUM_THREADS")},
            "sample": f"{strip_cols}x{sc.shape[1]} px strip (x {x0}..{x0 + strip_cols - 1}) of the same "
                      f"workload, {pp} pp, {dt:.2f} s on {cores} threads, C restatement of process_point "
                      f"(OpenMP over pixels)",
            "serial_value": pp1 / dt1,
            "serial_sample": f"{serial_cols}x{sc.shape[1]} px strip, {pp1} pp, {dt1:.2f} s, 1 thread"}
    got = gpu_strip(x0, x0 + strip_cols)
    if got is None:
        return base, {}
    nz = rcnt.ravel() > 0
    rel = np.abs(got.ravel()[nz] - ref.ravel()[nz]) / np.abs(ref.ravel()[nz])
    zeros_ok = bool(np.all(got.ravel()[~nz] == 0))
    return base, {"strip_f64_max_rel_err": float(rel.max()) if nz.any() else 0.0,
                  "strip_zero_pixels_exact": zeros_ok, "strip_pixels_checked": int(nz.sum())}


def _nproc():
    import subprocess
    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip())
    except Exception:
        return None


def small_scene_leg(cores):
    """The reference's CPU-runnable case (500 x 50 px, 5001 pulses, Sint 400): GPU FP64 and split
    rates and parity against the oracle, and the oracle itself on all cores and on one thread."""
    from oracle import oracle_lib as O
    from sarbp import synth
    from sarbp.core import Dataset
    sc = synth.small_scene(seed=0)
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    scene = ds.scene(*sc.flat(), dimx=sc.dimx)
    pp = scene.count_pp()
    out = {"workload": "500x50 px x 5001 pulses, Nra 600, Sint 400", "pp": int(pp)}
    res = {}
    for prec in ("f64", "split"):
        _, ks = time_precision(scene, prec, 5, 1)
        out[f"gpu_{prec}_pp_per_s"] = pp / ks
        res[prec] = scene.result(with_counts=True)
    scene.close()
    ds.close()
    t0 = time.perf_counter()
    ref, rcnt = O.focus(*sc.args(), nthreads=cores)
    out["cpu_pp_per_s"] = pp / (time.perf_counter() - t0)
    ref, rcnt = ref.ravel(order="F"), rcnt.ravel(order="F")  # the scene's flat (F-order) pixel order
    out["cpu_cores"] = int(cores)
    t0 = time.perf_counter()
    O.focus(*sc.args(), nthreads=1)
    out["cpu_serial_pp_per_s"] = pp / (time.perf_counter() - t0)
    got, cnt = res["f64"]
    nz = rcnt > 0
    out["f64_counts_exact"] = bool(np.array_equal(cnt, rcnt))
    out["f64_max_rel_err"] = float((np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])).max())
    out["split_vs_oracle"] = split_error_report(res["split"][0], ref)
    return out


def target_peaks(sc, img_split, img_f64):
    """Point-target responses of the split image against the FP64 image (sarbp.metrics.target_peaks)."""
    from sarbp import metrics
    return metrics.target_peaks(img_split.reshape(sc.shape, order="F"), img_f64.reshape(sc.shape, order="F"),
                                sc.targets[:, :2], float(sc.x[0, 0]), float(sc.y[0, 0]),
                                float(sc.x[1, 0] - sc.x[0, 0]), float(sc.y[0, 1] - sc.y[0, 0]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=30,
                    help="untimed passes first (the GPU's clocks settle over the first ~0.1 s of load)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--strip", type=int, default=4000, help="pixel columns of the all-cores CPU baseline strip "
                    "(4000: the whole scene, ~3.5 s on 16 cores)")
    ap.add_argument("--serial-strip", type=int, default=512, help="pixel columns of the 1-thread CPU strip (~7 s)")
    ap.add_argument("--precisions", default="f64,split")
    ap.add_argument("--small", action="store_true", help="scaled-down scene (quick check, not the metric)")
    ap.add_argument("--f64-map", default="1,8", help="FP64 work mapping kernel,tile_x")
    ap.add_argument("--split-map", default="16,0", help="split work mapping tile_x,accum")
    ap.add_argument("--sweep", action="store_true", help="time every work mapping first and use the fastest")
    ap.add_argument("--batch-requests", type=int, default=64, help="sub-scene requests of the batch leg (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="Poisson serving leg duration (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate of the serving leg")
    ap.add_argument("--serve-load-rate", type=float, default=1200.0,
                    help="arrival rate of the sustained-load serving leg (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (pixel sharding across GPUs "
              "is described, not built)", file=sys.stderr)
        sys.exit(2)

    from sarbp import synth
    from sarbp.core import Dataset

    t0 = time.perf_counter()
    if args.small:
        sc = synth.full_scene(seed=1, nx=512, ny=250, naz=5121, ntarget=200)
    else:
        sc = synth.full_scene(seed=1)
    log(f"scene {sc.name} generated in {time.perf_counter() - t0:.1f} s")
    ds = Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    xf, yf, zf, mf, inf, sf = sc.flat()
    scene = ds.scene(xf, yf, zf, mf, inf, sf, dimx=sc.dimx)
    pp = scene.count_pp()
    log(f"dataset resident ({ds.device_bytes / 1e9:.2f} GB), {scene.n} pixels, {pp} pp")
    k, tx = (int(v) for v in args.f64_map.split(","))
    scene.set_tiling("f64", k, tx, 0)
    tx, acc = (int(v) for v in args.split_map.split(","))
    scene.set_tiling("split", 1, tx, acc)
    sweep = {}
    if args.sweep:
        best = {}
        for prec, maps in (("f64", [(0, 16, 0)] + [(1, t, 0) for t in (64, 32, 16, 8, 4)]),
                           ("split", [(1, t, a) for t in (64, 32, 16, 8) for a in (0, 1)])):
            for m in maps:
                scene.set_tiling(prec, *m)
                _, ks = time_precision(scene, prec, 2, 1)
                sweep[f"{prec}:{m[0]},{m[1]},{m[2]}"] = pp / ks
                log(f"sweep {prec} kernel={m[0]} tile_x={m[1]} accum={m[2]}: {pp / ks:.4e} pp/s (kernel time)")
                if prec not in best or ks < best[prec][0]:
                    best[prec] = (ks, m)
        for prec, (_, m) in best.items():
            scene.set_tiling(prec, *m)
            log(f"using {prec} mapping {m}")

    res = {}
    outs = {}
    for prec in [p for p in args.precisions.split(",") if p]:
        try:
            wall, gpu_s = time_precision(scene, prec, args.steps, args.warmup)
        except ValueError as e:
            log(f"{prec}: unavailable ({e})")
            continue
        res[prec] = {"wall_s": wall, "kernel_s": gpu_s, "pp_per_s": pp / wall}
        outs[prec] = scene.result()
        log(f"{prec}: {wall * 1e3:.2f} ms/step (kernel {gpu_s * 1e3:.2f} ms), {pp / wall:.4e} pp/s")

    head = "split" if "split" in res else "f64"
    extra = {"pp_per_step": pp, "pixels": int(scene.n), "f64_map": args.f64_map, "split_map": args.split_map}
    if sweep:
        extra["sweep_pp_per_s"] = sweep
    prof = _load_profile(pp) if not args.small and not args.sweep else None
    for p, r in res.items():
        extra[f"{p}_pp_per_s"] = r["pp_per_s"]
        extra[f"{p}_kernel_ms"] = r["kernel_s"] * 1e3
        # the same gather-bandwidth roofline for each path
        extra[f"{p}_roofline_frac"] = BYTES_PER_PP * pp / r["kernel_s"] / 1e9 / HBM_PEAK_GBS
        # the VALU-busy fraction of the profiled launch: quad-cycles in which its wavefronts issued
        # VALU work (rocprofv3 SQ_ACTIVE_INST_VALU, which counts each instruction's quad-cycles, less
        # SQ_ACTIVE_INST_VALU2, the quad-cycles that issued two) over the quad-cycles the 1024 SIMDs
        # had in that launch (its GRBM_GUI_ACTIVE clocks per XCD / 4)
        kp = (prof or {}).get("kernels", {}).get(KERNEL_OF[p])
        if kp and default_maps(args):
            c = kp["counters"]
            if "SQ_ACTIVE_INST_VALU2" in c and c.get("GRBM_GUI_ACTIVE"):
                busy = c["SQ_ACTIVE_INST_VALU"] - c["SQ_ACTIVE_INST_VALU2"]
                extra[f"{p}_valu_busy_frac"] = busy / (SIMDS * c["GRBM_GUI_ACTIVE"] / XCDS / 4)
                if kp.get("counter_dispatch_ms"):
                    extra[f"{p}_clock_ghz"] = c["GRBM_GUI_ACTIVE"] / XCDS / (kp["counter_dispatch_ms"] * 1e-3) / 1e9
            extra[f"{p}_valu_lane_instr_per_pp"] = 64.0 * c["SQ_INSTS_VALU"] / pp
    if "split" in res and "f64" in res:
        extra["split_over_f64"] = res["split"]["pp_per_s"] / res["f64"]["pp_per_s"]
        extra["split_vs_f64_errors"] = split_error_report(outs["split"], outs["f64"])
        if sc.targets is not None:
            extra["split_vs_f64_target_peaks"] = target_peaks(sc, outs["split"], outs["f64"])

    # batch leg: uneven sub-scene requests in one persistent launch, mapping chosen by the autotuner
    if args.batch_requests > 0:
        from sarbp.batch import Batch, patch_requests
        reqs = patch_requests(sc, args.batch_requests, seed=0)
        b = Batch(ds, reqs)
        bpp = b.count_pp()
        tp, ch, tms = b.autotune(reps=1)
        times = [b.run() for _ in range(max(1, args.steps))]
        bms = float(np.mean(times))
        extra["batch"] = {"requests": len(reqs), "pixels": int(sum(r.n for r in reqs)), "pp": int(bpp),
                          "ms_per_launch": bms, "pp_per_s": bpp / (bms * 1e-3), "tile_px": tp, "pulse_chunk": ch,
                          "precisions": "split:f64 = 1:1, widths 128-1024, heights 32-250, Sint 500-4000"}
        log(f"batch of {len(reqs)}: {bms:.2f} ms, {bpp / (bms * 1e-3):.4e} pp/s (tile_px {tp}, chunk {ch})")
        b.close()
    # serving leg: Poisson arrivals of the same kind of requests, latency = completion - arrival
    if args.serve_seconds > 0:
        from sarbp.serve import Server, make_pool, poisson_run
        srv = Server()
        pool = make_pool([sc], nreq_per_dataset=32, seed=1)
        sv = poisson_run(srv, [ds], pool, args.serve_rate, args.serve_seconds, warmup_s=1.0)
        srv.close()
        extra["serving"] = {k: sv[k] for k in ("rate_req_per_s", "duration_s", "requests", "latency_p50_ms",
                                              "latency_p99_ms", "latency_max_ms", "pp_per_s", "precision_counts")}
        extra["serving"]["latency"] = ("scheduled Poisson arrival -> result in the caller's buffer (driver lag, "
                                       "submission and queueing included)")
        extra["latency_p50_ms"] = sv["latency_p50_ms"]
        extra["latency_p99_ms"] = sv["latency_p99_ms"]
        log(f"serving {args.serve_rate}/s for {args.serve_seconds}s: p50 {sv['latency_p50_ms']:.2f} ms, "
            f"p99 {sv['latency_p99_ms']:.2f} ms, {sv['pp_per_s']:.3e} pp/s")
        if args.serve_load_rate > 0:  # sustained load: many requests share each round
            srv = Server()
            pool = make_pool([sc], nreq_per_dataset=256, seed=2)
            sl = poisson_run(srv, [ds], pool, args.serve_load_rate, args.serve_seconds, warmup_s=1.0)
            srv.close()
            extra["serving_load"] = {k: sl[k] for k in ("rate_req_per_s", "duration_s", "requests", "latency_p50_ms",
                                                       "latency_p99_ms", "pp_per_s", "rounds", "rounds_per_request",
                                                       "gpu_busy_frac")}
            log(f"serving {args.serve_load_rate}/s: p50 {sl['latency_p50_ms']:.2f} ms, p99 {sl['latency_p99_ms']:.2f} ms, "
                f"{sl['pp_per_s']:.3e} pp/s, {sl['rounds_per_request']:.3f} rounds/request")

    cpu = None
    if not args.no_cpu:
        full_f64 = outs.get("f64")
        shape = sc.shape

        def gpu_strip(a, b):
            if full_f64 is None:
                return None
            return full_f64.reshape(shape, order="F")[a:b]

        cpu, parity = cpu_baseline(sc, args.strip, args.serial_strip, gpu_strip)
        extra.update(parity)
        extra["small_scene"] = small_scene_leg(cpu["cores"])
        log(f"small scene: {extra['small_scene']}")
        log(f"cpu baseline {cpu['value']:.4e} pp/s on {cpu['cores']} cores; strip parity {parity}")

    r = res[head]
    kernel_s = r["kernel_s"]
    achieved = BYTES_PER_PP * pp / kernel_s / 1e9
    traffic, traffic_src = None, None
    kname = KERNEL_OF[head]
    if prof and default_maps(args) and kname in prof.get("kernels", {}):
        kp = prof["kernels"][kname]
        traffic = int(kp["traffic_bytes"])
        traffic_src = ("%s: rocprofv3 --pmc FETCH_SIZE x%.3f + WRITE_SIZE x%.3f per launch of %s (scales from a "
                       "1 GiB copy probe)" % (PROFILE_JSON, prof["fetch_scale"], prof["write_scale"], kname))
        # the profiled session's own figures for the same kernel: the trace's average launch and
        # the roofline fraction it implies, beside this run's HIP-event figure
        extra["profile"] = {"source": PROFILE_JSON, "kernel": kname, "trace_avg_ms": kp["avg_ns"] / 1e6,
                            "trace_roofline_frac": BYTES_PER_PP * pp / (kp["avg_ns"] * 1e-9) / 1e9 / HBM_PEAK_GBS,
                            "session_event_ms": kp.get("bench_event_ms")}
    line = {
        "metric": "pixel·pulses/s split-FP32 vs FP64 at ≤1e-3 rad phase err; p50/p99 request latency",
        "value": r["pp_per_s"],
        "unit": "pixel·pulses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 split" if head == "split" else "f64",
        "data": "synthetic (seeded point targets + 20 dB noise)",
        "config": {"workload": f"{sc.shape[0]}x{sc.shape[1]} px x {sc.Naz} pulses, Nra {sc.Nra}, "
                               f"Sint U[2500,3500], mask 0.95",
                   "precision": head, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if traffic_src:
        extra["traffic_source"] = traffic_src
        extra["traffic_over_algorithmic"] = traffic / (BYTES_PER_PP * pp)
    line.update(extra)
    scene.close()
    ds.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
