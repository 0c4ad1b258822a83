"""ctypes binding of libsarbp.so (the C ABI in include/sarbp.h).

The library is the only compute path: if it is missing, or no HIP device is visible, calls raise
instead of falling back to anything else.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsarbp.so")

OK = 0
EINVAL = -1
ENOMEM = -2
ELAUNCH = -3
EDEVICE = -4
ENODEV = -5
EBUDGET = -6
ESTATE = -7
EAGAIN = -8

PREC_F64 = 0
PREC_SPLIT = 1

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_d = ctypes.c_double
_pp = ctypes.POINTER(ctypes.c_void_p)

FILL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p)


class Request(ctypes.Structure):
    """sarbp_request (include/sarbp.h)."""
    _fields_ = [("x", _p), ("y", _p), ("z", _p), ("mask", _p), ("index", _p), ("sint", _p),
                ("n", _i64), ("dimx", _i64), ("precision", _i32), ("tile_x", _i32)]


_RP = ctypes.POINTER(Request)


class ServerConfig(ctypes.Structure):
    """sarbp_server_config (include/sarbp.h); zero fields take the library defaults."""
    _fields_ = [("hbm_budget_bytes", _i64), ("quantum_ms", _d), ("max_active", _i32), ("pulse_chunk", _i32),
                ("phase_err_limit_rad", _d), ("max_segments", _i32)]


class RequestInfo(ctypes.Structure):
    """sarbp_request_info (include/sarbp.h)."""
    _fields_ = [("state", _i32), ("precision", _i32), ("phase_err_est_rad", _d), ("submit_s", _d),
                ("admit_s", _d), ("done_s", _d), ("pp", _i64), ("rounds", _i32), ("preemptions", _i32),
                ("segments", _i32), ("round_s", _d), ("gpu_done_s", _d)]


class ServerStats(ctypes.Structure):
    """sarbp_server_stats (include/sarbp.h)."""
    _fields_ = [("submitted", _i64), ("completed", _i64), ("failed", _i64), ("rounds", _i64), ("pp_done", _i64),
                ("gpu_busy_s", _d), ("device_bytes_in_use", _i64), ("active", _i32), ("queued", _i32),
                ("admit_s", _d), ("build_s", _d), ("round_wall_s", _d), ("records", _i64)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_PREEMPTED, REQ_FAILED = 0, 1, 2, 3, 4
_i32p = ctypes.POINTER(_i32)
_fp = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every symbol include/sarbp.h declares.
SIGNATURES = {
    "sarbp_abi_version": (ctypes.c_int, []),
    "sarbp_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "sarbp_last_error": (ctypes.c_char_p, []),
    "sarbp_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "sarbp_dataset_create": (ctypes.c_int, [_p, _i64, _i64, _p, _p, _p, _d, _d, _d, _pp]),
    "sarbp_dataset_create_streamed": (ctypes.c_int, [FILL_FN, _p, _i64, _i64, _p, _p, _p, _d, _d, _d, _pp]),
    "sarbp_dataset_destroy": (ctypes.c_int, [_p]),
    "sarbp_dataset_bytes": (ctypes.c_int, [_p, ctypes.POINTER(_i64)]),
    "sarbp_scene_create": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, _pp]),
    "sarbp_scene_set_tiling": (ctypes.c_int, [_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "sarbp_scene_set_bin_min": (ctypes.c_int, [_p, _i32]),
    "sarbp_scene_destroy": (ctypes.c_int, [_p]),
    "sarbp_scene_focus": (ctypes.c_int, [_p, _p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "sarbp_scene_download": (ctypes.c_int, [_p, _p, _p]),
    "sarbp_focus": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, ctypes.c_int, _p, _p]),
    "sarbp_count_pp": (_i64, [_p, _p, _p, _i64, _i64]),
    "sarbp_eval_pairs": (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p]),
    "sarbp_batch_create": (ctypes.c_int, [_p, _RP, _i32, _pp]),
    "sarbp_batch_run": (ctypes.c_int, [_p, _i32, _i32, _fp]),
    "sarbp_batch_autotune": (ctypes.c_int, [_p, _i32, _i32p, _i32p, _fp]),
    "sarbp_batch_download": (ctypes.c_int, [_p, _i32, _p, _p]),
    "sarbp_batch_destroy": (ctypes.c_int, [_p]),
    "sarbp_batch_focus_noalloc": (ctypes.c_int, [_p, _RP, _i32, ctypes.POINTER(_p), ctypes.POINTER(_p), _i32, _i32]),
    "sarbp_split_error_estimate": (ctypes.c_int, [_p, _RP, ctypes.POINTER(_d)]),
    "sarbp_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), _pp]),
    "sarbp_server_submit": (ctypes.c_int, [_p, _p, _RP, _p, _p, ctypes.POINTER(_i64)]),
    "sarbp_server_poll": (ctypes.c_int, [_p, _i64, _i32p]),
    "sarbp_server_wait": (ctypes.c_int, [_p, _i64, _d]),
    "sarbp_server_preempt": (ctypes.c_int, [_p, _i64]),
    "sarbp_server_resume": (ctypes.c_int, [_p, _i64]),
    "sarbp_server_info": (ctypes.c_int, [_p, _i64, ctypes.POINTER(RequestInfo)]),
    "sarbp_server_get_stats": (ctypes.c_int, [_p, ctypes.POINTER(ServerStats)]),
    "sarbp_server_release": (ctypes.c_int, [_p, _i64]),
    "sarbp_server_destroy": (ctypes.c_int, [_p]),
}

_lib = None

# Objects holding library handles.  They are closed at interpreter exit, before the HIP runtime
# tears itself down; finalizers that run later do nothing.
_live = weakref.WeakSet()
_shutting_down = False


def track(obj) -> None:
    _live.add(obj)


def alive() -> bool:
    return not _shutting_down


@atexit.register
def _close_all():
    global _shutting_down
    for o in sorted(list(_live), key=lambda o: getattr(o, "_close_order", 0)):
        try:
            o.close()
        except Exception:
            pass
    _shutting_down = True


class SarbpError(RuntimeError):
    """A libsarbp call failed; ``status`` is the negative status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def load():
    """Load libsarbp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libsarbp.so not built at {LIB_PATH}; run __graft_entry__.build() "
                          "(make -C sar-backprojection_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status == OK:
        return
    L = load()
    msg = L.sarbp_last_error().decode(errors="replace") or L.sarbp_strerror(status).decode()
    text = f"{what}: {msg}" if what else msg
    if status == EINVAL:
        raise ValueError(text)
    if status == ENOMEM:
        raise MemoryError(text)
    raise SarbpError(status, text)


def device_count() -> int:
    n = ctypes.c_int(0)
    load().sarbp_device_count(ctypes.byref(n))
    return int(n.value)


def require_device() -> None:
    if device_count() <= 0:
        raise SarbpError(ENODEV, "no HIP device visible: libsarbp needs an MI355X (gfx950) GPU")
