"""turbobfs_amd -- MI355X-native (gfx950) GPU BFS behind TurboBFS's C host API.

The compute lives in libturbobfs.so (hand-written HIP kernels + a C-ABI,
include/*.h).  This package is the Python side: a ctypes binding that mirrors
the reference's entry points, the resident-graph API, and the synthetic
workload generators.
"""
from . import _lib, graphs  # noqa: F401
from ._lib import LIB_PATH, TbfsError, device_count, device_info, load  # noqa: F401
from .bfs import (  # noqa: F401
    Batch,
    Context,
    Graph,
    Result,
    Server,
    bfs_gpu_bu_csc_sc,
    bfs_gpu_bu_csc_wa,
    bfs_gpu_td_cooc_sc,
    bfs_gpu_td_csc_sc,
    bfs_gpu_td_csc_wa,
    bfs_gpu_tdbu_csc_sc,
    bfs_gpu_tdbu_csc_wa,
    batch_tune_clear,
    batch_tuned,
    default_opts,
    dropin_mapping,
    dropin_release,
    host_array,
    server_budget_for,
    tune_clear,
    tune_count,
    tune_load,
    tune_save,
    tuned,
)

__all__ = [
    "Batch", "Context", "Graph", "Server", "host_array", "Result", "default_opts", "graphs", "load", "LIB_PATH",
    "batch_tuned", "batch_tune_clear", "server_budget_for",
    "tuned", "tune_save", "tune_load", "tune_clear", "tune_count",
    "device_count", "device_info",
    "TbfsError",
    "bfs_gpu_td_csc_sc", "bfs_gpu_td_csc_wa", "bfs_gpu_bu_csc_sc",
    "bfs_gpu_bu_csc_wa", "bfs_gpu_tdbu_csc_sc", "bfs_gpu_tdbu_csc_wa",
    "bfs_gpu_td_cooc_sc", "dropin_mapping", "dropin_release",
]
