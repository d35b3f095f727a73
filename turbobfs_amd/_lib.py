"""ctypes binding of libturbobfs.so (the C-ABI declared in include/*.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the shared object is missing, importing the compute entry points
raises, so a run without the HIP library fails loudly instead of silently
computing something else.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libturbobfs.so")

# status codes (include/tbfs.h)
TBFS_OK = 0
DROPIN_FAST = 0
DROPIN_REFERENCE = 1
TBFS_ERR_INVALID = 1
TBFS_ERR_HIP = 2
TBFS_ERR_NOMEM = 3
TBFS_ERR_UNSUPPORTED = 4
TBFS_SIGMA_INEXACT = 5

TBFS_GRAPH_SYMMETRIC = 1
TBFS_GRAPH_DEVICE = 2

MODE_AUTO = 0
MODE_PULL_SC = 1
MODE_PULL_WA = 2
MODE_PUSH = 3
MODE_COOC = 4

PULL_SC = 0
PULL_WA = 1
PULL_LB = 2

LOOP_HOST = 0
LOOP_PERSISTENT = 1
LOOP_CHAINED = 2
LOOP_GRAPH = 3

_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)


class Opts(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("loop", C.c_int32),
        ("sigma", C.c_int32),
        ("bbreak", C.c_int32),
        ("alpha", C.c_float),
        ("beta", C.c_float),
        ("block", C.c_int32),
        ("grid", C.c_int32),
        ("pull", C.c_int32),
        ("fb_lds", C.c_int32),
        ("batch_alpha", C.c_float),
        ("batch_count_bits", C.c_int32),
        ("batch_order", C.c_int32),
        ("stage_ratio", C.c_float),
        ("lb_grid", C.c_int32),
        ("finish", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("levels", C.c_int32),
        ("push_levels", C.c_int32),
        ("pull_levels", C.c_int32),
        ("pad", C.c_int32),
        ("reached", C.c_int64),
        ("edges", C.c_int64),
        ("ovf64", C.c_int64),
        ("ms", C.c_double),
        ("launches", C.c_int32),
        ("pad2", C.c_int32),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("mem_budget", C.c_int64),
        ("slots", C.c_int32),
        ("max_batch", C.c_int32),
        ("open_depth", C.c_int32),
        ("short_slots", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


class QueryDone(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("graph", C.c_int32),
        ("root", C.c_int32),
        ("status", C.c_int32),
        ("batch", C.c_int32),
        ("submit_us", C.c_double),
        ("start_us", C.c_double),
        ("done_us", C.c_double),
        ("level", C.c_int32),
        ("pad", C.c_int32),
    ]


DEGREE_IN = 0
DEGREE_OUT = 1


class DegreeStats(C.Structure):
    _fields_ = [
        ("max", C.c_int32),
        ("argmax", C.c_int32),
        ("mean", C.c_double),
        ("sd", C.c_double),
        ("sum", C.c_int64),
        ("bins", C.c_int64 * 33),
        ("nbins", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


# name -> (restype, argtypes); every symbol the public headers declare
SIGNATURES = {
    # drop-in surface (bfstdcsc.h, bfsbucsc.h, bfstdbucsc.h, bfstdcooc.h)
    "bfs_gpu_td_csc_sc": (C.c_int, [_i32p, _i32p, _i32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bfs_gpu_td_csc_wa": (C.c_int, [_i32p, _i32p, _i32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bfs_gpu_bu_csc_sc": (C.c_int, [_i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bfs_gpu_bu_csc_wa": (C.c_int, [_i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bfs_gpu_tdbu_csc_sc": (C.c_int, [_i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bfs_gpu_tdbu_csc_wa": (C.c_int, [_i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bfs_gpu_td_cooc_sc": (C.c_int, [_i32p, _i32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int]),
    # extended API (tbfs.h)
    "tbfs_opts_default": (None, [C.POINTER(Opts)]),
    "tbfs_status_str": (C.c_char_p, [C.c_int]),
    "tbfs_graph_create": (C.c_int, [_i32p, _i32p, C.c_int32, C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "tbfs_graph_create_cooc": (C.c_int, [_i32p, _i32p, C.c_int32, C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "tbfs_graph_destroy": (C.c_int, [C.c_void_p]),
    "tbfs_graph_info": (C.c_int, [C.c_void_p, _i32p, _i32p, C.POINTER(C.c_uint32)]),
    "tbfs_graph_degree_stats": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(DegreeStats)]),
    "tbfs_batch_tuned": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Opts)]),
    "tbfs_batch_tune_clear": (C.c_int, []),
    "tbfs_ctx_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "tbfs_ctx_destroy": (C.c_int, [C.c_void_p]),
    "tbfs_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "tbfs_bfs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(Opts), C.POINTER(Stats)]),
    "tbfs_bfs_many": (C.c_int, [C.c_void_p, C.c_void_p, _i32p, C.c_int32, C.POINTER(Opts), C.POINTER(Stats)]),
    "tbfs_bfs_multi": (C.c_int, [C.c_void_p, C.c_void_p, _i32p, C.c_int32, C.POINTER(Opts), C.POINTER(Stats)]),
    "tbfs_ctx_read": (C.c_int, [C.c_void_p, C.c_int32, _i32p, _u64p, _u8p]),
    "tbfs_ctx_device_ptrs": (C.c_int, [C.c_void_p, C.POINTER(_i32p), C.POINTER(_u64p), C.POINTER(_u8p)]),
    "tbfs_ctx_mark": (C.c_int, [C.c_void_p, C.c_int32]),
    "tbfs_ctx_elapsed": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "tbfs_calibrate_copy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "tbfs_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "tbfs_device_info": (C.c_int, [C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "tbfs_batch_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "tbfs_batch_destroy": (C.c_int, [C.c_void_p]),
    "tbfs_batch_stream": (C.c_void_p, [C.c_void_p]),
    "tbfs_batch_run": (C.c_int, [C.c_void_p, C.c_void_p, _i32p, C.c_int32, C.POINTER(Opts), C.POINTER(Stats)]),
    "tbfs_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, _i32p, C.c_int32, C.POINTER(Opts),
                                      C.POINTER(Opts), C.POINTER(C.c_double)]),
    "tbfs_batch_read": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _i32p, _u64p, _u8p]),
    "tbfs_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    "tbfs_server_budget_for": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(C.c_int64)]),
    "tbfs_server_destroy": (C.c_int, [C.c_void_p]),
    "tbfs_server_submit": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _i32p, _u64p, _u8p, C.POINTER(C.c_uint64)]),
    "tbfs_server_poll": (C.c_int, [C.c_void_p, C.POINTER(QueryDone), C.c_int32, C.POINTER(C.c_int32), C.c_int64]),
    "tbfs_server_drain": (C.c_int, [C.c_void_p]),
    "tbfs_dropin_release": (C.c_int, []),
    "tbfs_server_hold": (C.c_int, [C.c_void_p, C.c_int32]),
    "tbfs_server_info": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "tbfs_dropin_mapping": (C.c_int, [C.c_int]),
    "tbfs_autotune": (C.c_int, [C.c_void_p, C.c_void_p, _i32p, C.c_int32, C.POINTER(Opts),
                                C.POINTER(Opts), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tbfs_tuned": (C.c_int, [C.c_void_p, C.POINTER(Opts)]),
    "tbfs_tune_save": (C.c_int, [C.c_char_p]),
    "tbfs_tune_load": (C.c_int, [C.c_char_p]),
    "tbfs_tune_clear": (C.c_int, []),
    "tbfs_tune_count": (C.c_int, []),
    "tbfs_bfs_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.POINTER(Opts), C.POINTER(Stats)]),
    "tbfs_device_alloc": (C.c_void_p, [C.c_int64]),
    "tbfs_device_free": (C.c_int, [C.c_void_p]),
    "tbfs_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "tbfs_device_fill": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    "tbfs_host_alloc": (C.c_void_p, [C.c_int64]),
    "tbfs_host_free": (C.c_int, [C.c_void_p]),
    # generators (tbfs_gen.h)
    "tbfs_gen_er": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, C.POINTER(C.c_void_p)]),
    "tbfs_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int32, C.POINTER(C.c_void_p)]),
    "tbfs_gen_grid": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.c_int32, C.POINTER(C.c_void_p)]),
    "tbfs_gen_ws": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.POINTER(C.c_void_p)]),
    "tbfs_csc_host_info": (C.c_int, [C.c_void_p, _i32p, _i32p]),
    "tbfs_csc_host_copy": (C.c_int, [C.c_void_p, _i32p, _i32p]),
    "tbfs_csc_host_free": (None, [C.c_void_p]),
}

_lib = None


def load() -> C.CDLL:
    """Load libturbobfs.so (once) and attach the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(or make -C turbobfs_amd/csrc); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def device_count() -> int:
    n = C.c_int32(0)
    load().tbfs_device_count(C.byref(n))
    return int(n.value)


def device_info(dev: int = 0) -> dict:
    name = C.create_string_buffer(128)
    cus = C.c_int32()
    hbm = C.c_int64()
    coop = C.c_int32()
    check(load().tbfs_device_info(dev, name, 128, C.byref(cus), C.byref(hbm), C.byref(coop)),
          "tbfs_device_info")
    return {"name": name.value.decode(), "cus": cus.value, "hbm_bytes": hbm.value,
            "cooperative": bool(coop.value)}


def status_str(rc: int) -> str:
    return load().tbfs_status_str(rc).decode()


class TbfsError(RuntimeError):
    def __init__(self, what: str, rc: int):
        super().__init__(f"{what}: {status_str(rc)} ({rc})")
        self.rc = rc


def check(rc: int, what: str) -> None:
    if rc != TBFS_OK:
        raise TbfsError(what, rc)
