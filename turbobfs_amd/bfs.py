"""Python mirror of the library's two surfaces.

1. The reference's GPU entry points, same names and argument meaning
   (TurboBFS_CSC_TD/bfstdcsc.h:54-66, TurboBFS_CSC_BU/bfsbucsc.h:55-67,
   TurboBFS_CSC_TDBU/bfstdbucsc.h:56-68, TurboBFS_COOC_TD/bfstdcooc.h:48-49):
   caller-provided numpy output arrays are filled in place and the C status
   is returned.  They call straight into libturbobfs.so.
2. The resident-graph API of include/tbfs.h: Graph (CSC/COOC in HBM) and
   Context (stream + per-BFS buffers).
"""
from __future__ import annotations

import ctypes as C
import sys
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Opts, Stats, check

_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)


def _ptr(a: np.ndarray, ptype):
    return a.ctypes.data_as(ptype)


def _req(a, dtype, n=None, name="array"):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous numpy {np.dtype(dtype).name} array")
    if n is not None and a.shape[0] < n:
        raise ValueError(f"{name} has {a.shape[0]} entries, needs {n}")
    return a


def _release(obj) -> None:
    """Garbage-collection path of the handle classes.  During interpreter
    shutdown the HIP runtime may already be tearing down (its own exit
    handlers, or a profiler's), so no native call is made then: the process
    is ending and the driver reclaims the device memory.  Close handles
    explicitly (close() or a with-block) to free them earlier."""
    if sys.is_finalizing():
        return
    try:
        obj.close()
    except Exception:
        pass


class _Handle:
    """Context-manager support shared by the handle classes."""

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


# ---------------------------------------------------------------- drop-in ---
def bfs_gpu_td_csc_sc(IC_h, CP_h, S_h, sigma_h, r, nz, n, repetition) -> int:
    lib = _lib.load()
    return lib.bfs_gpu_td_csc_sc(
        _ptr(_req(IC_h, np.int32, nz, "IC_h"), _i32p), _ptr(_req(CP_h, np.int32, n + 1, "CP_h"), _i32p),
        _ptr(_req(S_h, np.int32, n, "S_h"), _i32p), _ptr(_req(sigma_h, np.float32, n, "sigma_h"), _f32p),
        r, nz, n, repetition)


def bfs_gpu_td_csc_wa(IC_h, CP_h, S_h, sigma_h, r, nz, n, repetition) -> int:
    lib = _lib.load()
    return lib.bfs_gpu_td_csc_wa(
        _ptr(_req(IC_h, np.int32, nz, "IC_h"), _i32p), _ptr(_req(CP_h, np.int32, n + 1, "CP_h"), _i32p),
        _ptr(_req(S_h, np.int32, n, "S_h"), _i32p), _ptr(_req(sigma_h, np.float32, n, "sigma_h"), _f32p),
        r, nz, n, repetition)


def bfs_gpu_bu_csc_sc(I_h, CP_h, S_h, r, nz, n, repetition, bbreak) -> int:
    lib = _lib.load()
    return lib.bfs_gpu_bu_csc_sc(
        _ptr(_req(I_h, np.int32, nz, "I_h"), _i32p), _ptr(_req(CP_h, np.int32, n + 1, "CP_h"), _i32p),
        _ptr(_req(S_h, np.int32, n, "S_h"), _i32p), r, nz, n, repetition, bbreak)


def bfs_gpu_bu_csc_wa(I_h, CP_h, S_h, r, nz, n, repetition, bbreak) -> int:
    lib = _lib.load()
    return lib.bfs_gpu_bu_csc_wa(
        _ptr(_req(I_h, np.int32, nz, "I_h"), _i32p), _ptr(_req(CP_h, np.int32, n + 1, "CP_h"), _i32p),
        _ptr(_req(S_h, np.int32, n, "S_h"), _i32p), r, nz, n, repetition, bbreak)


def bfs_gpu_tdbu_csc_sc(I_h, CP_h, S_h, sigma_h, r, nz, n, repetition, bbreak) -> int:
    lib = _lib.load()
    return lib.bfs_gpu_tdbu_csc_sc(
        _ptr(_req(I_h, np.int32, nz, "I_h"), _i32p), _ptr(_req(CP_h, np.int32, n + 1, "CP_h"), _i32p),
        _ptr(_req(S_h, np.int32, n, "S_h"), _i32p), _ptr(_req(sigma_h, np.int32, n, "sigma_h"), _i32p),
        r, nz, n, repetition, bbreak)


def bfs_gpu_tdbu_csc_wa(IC_h, CP_h, S_h, sigma_h, r, nz, n, repetition, bbreak) -> int:
    lib = _lib.load()
    return lib.bfs_gpu_tdbu_csc_wa(
        _ptr(_req(IC_h, np.int32, nz, "IC_h"), _i32p), _ptr(_req(CP_h, np.int32, n + 1, "CP_h"), _i32p),
        _ptr(_req(S_h, np.int32, n, "S_h"), _i32p), _ptr(_req(sigma_h, np.int32, n, "sigma_h"), _i32p),
        r, nz, n, repetition, bbreak)


def bfs_gpu_td_cooc_sc(I_h, J_h, sigma_h, r, nz, n, repetition) -> int:
    lib = _lib.load()
    return lib.bfs_gpu_td_cooc_sc(
        _ptr(_req(I_h, np.int32, nz, "I_h"), _i32p), _ptr(_req(J_h, np.int32, nz, "J_h"), _i32p),
        _ptr(_req(sigma_h, np.float32, n, "sigma_h"), _f32p), r, nz, n, repetition)


def dropin_release() -> None:
    """Free the contexts and device buffers the seven entry points keep
    between calls (later calls allocate again)."""
    check(_lib.load().tbfs_dropin_release(), "tbfs_dropin_release")


def dropin_mapping(mapping: int = -1) -> int:
    """Select the kernels behind the seven entry points above (process-wide):
    _lib.DROPIN_FAST (default) or _lib.DROPIN_REFERENCE (the reference's
    one-lane / one-wavefront mappings); -1 only queries.  Returns the
    previous mapping."""
    rc = _lib.load().tbfs_dropin_mapping(int(mapping))
    if rc not in (_lib.DROPIN_FAST, _lib.DROPIN_REFERENCE):
        check(rc, "tbfs_dropin_mapping")
    return rc


# ------------------------------------------------------- resident graphs ---
def default_opts(**kw) -> Opts:
    o = Opts()
    _lib.load().tbfs_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(f"tbfs_opts has no field {k}")
        setattr(o, k, v)
    return o


class Graph(_Handle):
    """A CSC (or COOC) graph resident in HBM."""

    def __init__(self, CP=None, IC=None, *, I=None, J=None, n=None,
                 symmetric: bool = False):
        lib = _lib.load()
        self._h = C.c_void_p()
        flags = _lib.TBFS_GRAPH_SYMMETRIC if symmetric else 0
        if CP is not None:
            CP = np.ascontiguousarray(CP, dtype=np.int32)
            IC = np.ascontiguousarray(IC, dtype=np.int32)
            self.n = int(CP.shape[0] - 1)
            self.nz = int(IC.shape[0])
            check(lib.tbfs_graph_create(_ptr(CP, _i32p), _ptr(IC, _i32p), self.n,
                                        self.nz, flags, C.byref(self._h)),
                  "tbfs_graph_create")
        else:
            I = np.ascontiguousarray(I, dtype=np.int32)
            J = np.ascontiguousarray(J, dtype=np.int32)
            self.n = int(n)
            self.nz = int(I.shape[0])
            check(lib.tbfs_graph_create_cooc(_ptr(I, _i32p), _ptr(J, _i32p), self.n,
                                             self.nz, flags, C.byref(self._h)),
                  "tbfs_graph_create_cooc")

    @property
    def handle(self):
        return self._h

    def degree_stats(self, which: str = "in") -> dict:
        """Degree statistics computed on the device (max, first vertex with
        it, mean, sample sd, degree sum, power-of-two bins: bins[0] degree 0,
        bins[k] degrees in [2^(k-1), 2^k)), of the in-lists or out-lists."""
        if which not in ("in", "out"):
            raise ValueError(f"which must be 'in' or 'out', not {which!r}")
        ds = _lib.DegreeStats()
        w = _lib.DEGREE_IN if which == "in" else _lib.DEGREE_OUT
        check(_lib.load().tbfs_graph_degree_stats(self._h, w, C.byref(ds)),
              "tbfs_graph_degree_stats")
        return {"max": ds.max, "argmax": ds.argmax, "mean": ds.mean, "sd": ds.sd,
                "sum": ds.sum, "bins": [int(ds.bins[k]) for k in range(ds.nbins)]}

    def close(self):
        if self._h:
            _lib.load().tbfs_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        _release(self)


@dataclass
class Result:
    S: np.ndarray        # int32, -1 = unreached
    sigma: np.ndarray    # uint64, count mod 2^64
    ovf: np.ndarray      # uint8, count >= 2^64
    stats: dict


class Context(_Handle):
    """Stream + per-BFS device buffers for graphs of up to n_cap vertices."""

    def __init__(self, n_cap: int):
        self._h = C.c_void_p()
        check(_lib.load().tbfs_ctx_create(int(n_cap), C.byref(self._h)), "tbfs_ctx_create")
        self.n_cap = int(n_cap)

    @property
    def stream(self) -> int:
        return int(_lib.load().tbfs_ctx_stream(self._h) or 0)

    def run(self, g: Graph, root: int, opts: Opts | None = None) -> dict:
        st = Stats()
        o = opts if opts is not None else default_opts()
        check(_lib.load().tbfs_bfs(self._h, g.handle, int(root), C.byref(o), C.byref(st)),
              "tbfs_bfs")
        return {f: getattr(st, f) for f, _ in Stats._fields_ if f not in ("pad", "pad2")}

    def run_many(self, g: Graph, roots, opts: Opts | None = None) -> list[dict]:
        """BFS from every root in turn, back to back on the device; returns
        the per-BFS stats (results of the last BFS stay on the device)."""
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        k = int(roots.shape[0])
        st = (Stats * k)()
        o = opts if opts is not None else default_opts()
        check(_lib.load().tbfs_bfs_many(self._h, g.handle, _ptr(roots, _i32p), k, C.byref(o),
                                        st), "tbfs_bfs_many")
        return [{f: getattr(x, f) for f, _ in Stats._fields_ if f not in ("pad", "pad2")} for x in st]

    def run_multi(self, g: Graph, roots, opts: Opts | None = None) -> dict:
        """One BFS from all roots at once (depth 0 = every root); results stay
        on the device (read with read())."""
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        st = Stats()
        o = opts if opts is not None else default_opts()
        check(_lib.load().tbfs_bfs_multi(self._h, g.handle, _ptr(roots, _i32p),
                                         int(roots.shape[0]), C.byref(o), C.byref(st)),
              "tbfs_bfs_multi")
        return {f: getattr(st, f) for f, _ in Stats._fields_ if f not in ("pad", "pad2")}

    def mark(self, slot: int) -> None:
        """Record timing marker `slot` on this context's stream."""
        check(_lib.load().tbfs_ctx_mark(self._h, int(slot)), "tbfs_ctx_mark")

    def elapsed(self, a: int, b: int) -> float:
        """Device milliseconds between markers a and b (waits for b)."""
        ms = C.c_double()
        check(_lib.load().tbfs_ctx_elapsed(self._h, int(a), int(b), C.byref(ms)),
              "tbfs_ctx_elapsed")
        return float(ms.value)

    def calibrate_copy(self, nbytes: int, reps: int = 5) -> float:
        """Average ms of a plain device copy of nbytes (reads and writes nbytes)."""
        ms = C.c_double()
        check(_lib.load().tbfs_calibrate_copy(self._h, int(nbytes), int(reps), C.byref(ms)),
              "tbfs_calibrate_copy")
        return float(ms.value)

    def autotune(self, g: Graph, roots, base: Opts | None = None):
        """Tune the single-source switch thresholds and launch shapes on g
        with these roots (tbfs_autotune); returns (best opts, ms per BFS of
        the winner, ms per BFS of the defaults).  The winner is recorded in
        the process's table for graphs of g's class."""
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        best = Opts()
        ms = C.c_double()
        dms = C.c_double()
        check(_lib.load().tbfs_autotune(self._h, g.handle, _ptr(roots, _i32p),
                                        int(roots.shape[0]),
                                        C.byref(base) if base is not None else None,
                                        C.byref(best), C.byref(ms), C.byref(dms)), "tbfs_autotune")
        return best, float(ms.value), float(dms.value)

    def read(self, n: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        S = np.empty(n, dtype=np.int32)
        sigma = np.empty(n, dtype=np.uint64)
        ovf = np.empty(n, dtype=np.uint8)
        check(_lib.load().tbfs_ctx_read(self._h, n, _ptr(S, _i32p), _ptr(sigma, _u64p),
                                        _ptr(ovf, _u8p)), "tbfs_ctx_read")
        return S, sigma, ovf

    def bfs(self, g: Graph, root: int, opts: Opts | None = None) -> Result:
        stats = self.run(g, root, opts)
        S, sigma, ovf = self.read(g.n)
        stats["ovf64"] = int(ovf[S >= 0].sum())
        return Result(S, sigma, ovf, stats)

    def close(self):
        if self._h:
            _lib.load().tbfs_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        _release(self)


class Batch(_Handle):
    """Many sources per BFS: up to b_cap (<= 1024) roots on one graph at once,
    frontiers bit-packed 64 sources per word.  Each source's result equals
    Context.bfs from that root."""

    def __init__(self, n_cap: int, b_cap: int):
        self._h = C.c_void_p()
        check(_lib.load().tbfs_batch_create(int(n_cap), int(b_cap), C.byref(self._h)),
              "tbfs_batch_create")
        self.n_cap = int(n_cap)
        self.b_cap = int(b_cap)
        self._n = 0
        self._count = 0

    def run(self, g: Graph, roots, opts: Opts | None = None) -> dict:
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        st = Stats()
        o = opts if opts is not None else default_opts()
        check(_lib.load().tbfs_batch_run(self._h, g.handle, _ptr(roots, _i32p),
                                         int(roots.shape[0]), C.byref(o), C.byref(st)),
              "tbfs_batch_run")
        self._n = g.n
        self._g = g
        self._count = int(roots.shape[0])
        out = {f: getattr(st, f) for f, _ in Stats._fields_ if f not in ("pad", "pad2")}
        out["count_bits"] = int(st.pad)
        return out

    def autotune(self, g: Graph, roots, base: Opts | None = None):
        """Time the launch-grid / push-threshold candidates on this graph and
        these roots; returns (best opts, best ms)."""
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        best = Opts()
        ms = C.c_double()
        check(_lib.load().tbfs_batch_autotune(self._h, g.handle, _ptr(roots, _i32p),
                                              int(roots.shape[0]),
                                              C.byref(base) if base is not None else None,
                                              C.byref(best), C.byref(ms)), "tbfs_batch_autotune")
        self._n = g.n
        self._g = g
        self._count = int(roots.shape[0])
        return best, float(ms.value)

    def read(self, first: int = 0, count: int | None = None):
        """(S, sigma, ovf) of sources first .. first+count-1, shape (count, n)."""
        if count is None:
            count = self._count - first
        n = self._n
        S = np.empty((count, n), dtype=np.int32)
        sigma = np.empty((count, n), dtype=np.uint64)
        ovf = np.empty((count, n), dtype=np.uint8)
        check(_lib.load().tbfs_batch_read(self._h, int(first), int(count), int(n),
                                          _ptr(S, _i32p), _ptr(sigma, _u64p), _ptr(ovf, _u8p)),
              "tbfs_batch_read")
        return S, sigma, ovf

    def close(self):
        if self._h:
            _lib.load().tbfs_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        _release(self)


def batch_tuned(g: Graph, count: int) -> Opts | None:
    """The autotuner table's grid / batch_alpha for a batch of `count`
    sources on g (None when no entry applies)."""
    o = default_opts()
    rc = _lib.load().tbfs_batch_tuned(g.handle, int(count), C.byref(o))
    if rc == _lib.TBFS_ERR_UNSUPPORTED:
        return None
    check(rc, "tbfs_batch_tuned")
    return o


def batch_tune_clear() -> None:
    check(_lib.load().tbfs_batch_tune_clear(), "tbfs_batch_tune_clear")


def tuned(g: Graph) -> Opts | None:
    """The single-source values the autotuner's table holds for g's class
    (alpha, beta, stage_ratio, lb_grid on top of the defaults), or None."""
    o = default_opts()
    rc = _lib.load().tbfs_tuned(g.handle, C.byref(o))
    if rc == _lib.TBFS_ERR_UNSUPPORTED:
        return None
    check(rc, "tbfs_tuned")
    return o


def tune_save(path: str) -> None:
    check(_lib.load().tbfs_tune_save(str(path).encode()), "tbfs_tune_save")


def tune_load(path: str) -> None:
    check(_lib.load().tbfs_tune_load(str(path).encode()), "tbfs_tune_load")


def tune_clear() -> None:
    check(_lib.load().tbfs_tune_clear(), "tbfs_tune_clear")


def tune_count() -> int:
    return int(_lib.load().tbfs_tune_count())


def host_array(count: int, dtype) -> np.ndarray:
    """A numpy array over pinned host memory (tbfs_host_alloc); freed when the
    array is garbage-collected."""
    dt = np.dtype(dtype)
    nbytes = max(1, int(count) * dt.itemsize)
    ptr = _lib.load().tbfs_host_alloc(nbytes)
    if not ptr:
        raise MemoryError("tbfs_host_alloc failed")
    buf = (C.c_uint8 * nbytes).from_address(ptr)
    arr = np.frombuffer(buf, dtype=dt, count=int(count))
    fin = weakref.finalize(buf, _host_free, ptr)
    fin.atexit = False  # at exit the process's pinned memory goes with it
    return arr


def _host_free(ptr: int) -> None:
    if not sys.is_finalizing():
        _lib.load().tbfs_host_free(C.c_void_p(ptr))


def server_budget_for(graphs, slots: int, words: int) -> int:
    """The Server mem_budget that gives `slots` slots room for `words`
    64-source words of the largest of `graphs` each."""
    arr = (C.c_void_p * len(graphs))(*[g.handle for g in graphs])
    b = C.c_int64()
    check(_lib.load().tbfs_server_budget_for(arr, len(graphs), int(slots), int(words),
                                             C.byref(b)), "tbfs_server_budget_for")
    return int(b.value)


class Server(_Handle):
    """Resident graphs + native worker threads, one per batch slot, that run
    queries as they arrive: single-source runs, batches, or (deep graphs,
    more levels than open_depth) continuous batches that later queries join
    at level boundaries.  short_slots slots stay free of long jobs."""

    def __init__(self, graphs, mem_budget: int = 0, slots: int = 0, max_batch: int = 0,
                 open_depth: int = 0, short_slots: int = 0):
        self._graphs = list(graphs)  # keep the graphs alive
        arr = (C.c_void_p * len(self._graphs))(*[g.handle for g in self._graphs])
        o = _lib.ServerOpts()
        o.mem_budget = int(mem_budget)
        o.slots = int(slots)
        o.max_batch = int(max_batch)
        o.open_depth = int(open_depth)
        o.short_slots = int(short_slots)
        self._h = C.c_void_p()
        check(_lib.load().tbfs_server_create(arr, len(self._graphs), C.byref(o), C.byref(self._h)),
              "tbfs_server_create")
        self._bufs = {}

    def submit(self, graph: int, root: int, S=None, sigma=None, ovf=None) -> int:
        """Queue a query; S / sigma / ovf (C-contiguous numpy int32 / uint64 /
        uint8 of at least n entries, n = that graph's vertex count, or None)
        receive the result.  The server keeps a reference to them until poll()
        reports the query done, since the device writes them asynchronously."""
        # (a graph index out of range is the library's to reject)
        n = self._graphs[int(graph)].n if 0 <= int(graph) < len(self._graphs) else 0
        if S is not None:
            _req(S, np.int32, n, "S")
        if sigma is not None:
            _req(sigma, np.uint64, n, "sigma")
        if ovf is not None:
            _req(ovf, np.uint8, n, "ovf")
        t = C.c_uint64()
        check(_lib.load().tbfs_server_submit(
            self._h, int(graph), int(root),
            _ptr(S, _i32p) if S is not None else None,
            _ptr(sigma, _u64p) if sigma is not None else None,
            _ptr(ovf, _u8p) if ovf is not None else None, C.byref(t)), "tbfs_server_submit")
        if S is not None or sigma is not None or ovf is not None:
            self._bufs[int(t.value)] = (S, sigma, ovf)
        return int(t.value)

    def hold(self, on: bool = True) -> None:
        """Hold (or release) dispatch: queries submitted while held are taken
        together once released (up to each graph's batch cap)."""
        check(_lib.load().tbfs_server_hold(self._h, 1 if on else 0), "tbfs_server_hold")

    def poll(self, max_items: int = 1024, wait_us: int = 0) -> list[dict]:
        buf = (_lib.QueryDone * max_items)()
        k = C.c_int32()
        check(_lib.load().tbfs_server_poll(self._h, buf, int(max_items), C.byref(k), int(wait_us)),
              "tbfs_server_poll")
        out = [{f: getattr(buf[i], f) for f, _ in _lib.QueryDone._fields_} for i in range(k.value)]
        for d in out:  # the device no longer writes these buffers
            self._bufs.pop(d["ticket"], None)
        return out

    def drain(self) -> None:
        check(_lib.load().tbfs_server_drain(self._h), "tbfs_server_drain")

    def info(self, graph: int = -1) -> dict:
        cap = C.c_int32()
        nb = C.c_int64()
        nq = C.c_int64()
        now = C.c_double()
        check(_lib.load().tbfs_server_info(self._h, int(graph), C.byref(cap), C.byref(nb),
                                           C.byref(nq), C.byref(now)), "tbfs_server_info")
        return {"batch_cap": cap.value, "batches": nb.value, "batched": nq.value,
                "now_us": now.value}

    def close(self):
        if self._h:
            _lib.load().tbfs_server_destroy(self._h)
            self._h = C.c_void_p()
        self._bufs.clear()  # destroy waited for the device's last copies

    def __del__(self):
        _release(self)
