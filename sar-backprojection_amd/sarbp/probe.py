"""Counter calibration: time a plain device copy of a known size (libsarbp_probe.so, a measurement
tool built apart from the library: make -C sar-backprojection_amd/csrc probe).

    python -m sarbp.probe --mib 1024 --reps 5

Under ``rocprofv3 --pmc FETCH_SIZE`` (or WRITE_SIZE) the copy kernel reads (writes) exactly
``mib`` MiB per dispatch, which gives the counters' scale for the focusing kernels' traffic."""
from __future__ import annotations

import argparse
import ctypes
import json
import os

_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsarbp_probe.so")


def copy(bytes_: int, reps: int = 5) -> float:
    if not os.path.exists(_SO):
        raise ImportError(f"{_SO} is not built: make -C sar-backprojection_amd/csrc probe")
    L = ctypes.CDLL(_SO)
    L.sarbp_probe_copy.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]
    L.sarbp_probe_copy.restype = ctypes.c_int
    ms = ctypes.c_float(0)
    rc = L.sarbp_probe_copy(int(bytes_), int(reps), ctypes.byref(ms))
    if rc:
        raise RuntimeError(f"sarbp_probe_copy failed ({rc})")
    return float(ms.value)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    ms = copy(a.mib << 20, a.reps)
    print(json.dumps({"copy_bytes": a.mib << 20, "ms": ms, "GBps_read_plus_write": 2 * (a.mib << 20) / ms / 1e6}))


if __name__ == "__main__":
    main()
