"""Host-side graph construction: the benchmark workloads and CSC utilities.

Generators call the deterministic C++ generators in libturbobfs.so
(include/tbfs_gen.h); no GPU is touched.  Graphs are returned as a pair of
int32 numpy arrays (CP, IC) in the reference's CSC layout: column v of IC
lists the in-neighbours u of v (TurboBFS_CSC_TD/bfstdcsc.h:30-34).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class CSC:
    CP: np.ndarray  # int32[n+1]
    IC: np.ndarray  # int32[nz]

    @property
    def n(self) -> int:
        return int(self.CP.shape[0] - 1)

    @property
    def nz(self) -> int:
        return int(self.IC.shape[0])

    def indeg(self) -> np.ndarray:
        return np.diff(self.CP).astype(np.int64)


def _i32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _take(handle: C.c_void_p) -> CSC:
    lib = _lib.load()
    n = C.c_int32()
    nz = C.c_int32()
    try:
        lib.tbfs_csc_host_info(handle, C.byref(n), C.byref(nz))
        CP = np.empty(n.value + 1, dtype=np.int32)
        IC = np.empty(max(nz.value, 0), dtype=np.int32)
        lib.tbfs_csc_host_copy(handle, _i32(CP), _i32(IC))
    finally:
        lib.tbfs_csc_host_free(handle)
    return CSC(CP, IC)


def _gen(fn_name: str, *args) -> CSC:
    lib = _lib.load()
    h = C.c_void_p()
    rc = getattr(lib, fn_name)(*args, C.byref(h))
    if rc != 0:
        raise ValueError(f"{fn_name}{args}: generator returned {rc}")
    return _take(h)


def erdos_renyi(n: int, p: float, seed: int) -> CSC:
    """Undirected G(n, p), symmetric CSC."""
    return _gen("tbfs_gen_er", n, p, seed)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
         c: float = 0.19, seed: int = 2, relabel: bool = True) -> CSC:
    """Symmetrized, deduplicated R-MAT (d = 1 - a - b - c)."""
    return _gen("tbfs_gen_rmat", scale, edge_factor, a, b, c, seed,
                1 if relabel else 0)


def grid(w: int, h: int, remove_frac: float = 0.0, seed: int = 3,
         relabel: bool = True) -> CSC:
    """w x h 4-neighbour grid with a fraction of its edges removed."""
    return _gen("tbfs_gen_grid", w, h, remove_frac, seed, 1 if relabel else 0)


def watts_strogatz(n: int, k: int = 10, p: float = 0.1, seed: int = 0) -> CSC:
    return _gen("tbfs_gen_ws", n, k, p, seed)


def block_diagonal(graphs: list[CSC]) -> tuple[CSC, np.ndarray]:
    """Pack graphs as diagonal blocks of one CSC.  Returns (csc, offsets)
    where block i owns vertex ids [offsets[i], offsets[i+1])."""
    ns = np.array([g.n for g in graphs], dtype=np.int64)
    nzs = np.array([g.nz for g in graphs], dtype=np.int64)
    off = np.zeros(len(graphs) + 1, dtype=np.int64)
    off[1:] = np.cumsum(ns)
    eoff = np.zeros(len(graphs) + 1, dtype=np.int64)
    eoff[1:] = np.cumsum(nzs)
    if off[-1] > 2**31 - 2 or eoff[-1] > 2**31 - 1:
        raise ValueError("block-diagonal graph exceeds int32 indexing")
    CP = np.empty(int(off[-1]) + 1, dtype=np.int32)
    IC = np.empty(int(eoff[-1]), dtype=np.int32)
    for i, g in enumerate(graphs):
        CP[off[i]:off[i + 1]] = g.CP[:-1] + eoff[i]
        IC[eoff[i]:eoff[i + 1]] = g.IC + off[i]
    CP[-1] = eoff[-1]
    return CSC(CP, IC), off


def components(g: CSC) -> np.ndarray:
    """Weakly connected component label of every vertex (scipy)."""
    from scipy.sparse import csc_matrix
    from scipy.sparse.csgraph import connected_components

    n = g.n
    A = csc_matrix((np.ones(g.nz, dtype=np.int8), g.IC, g.CP), shape=(n, n))
    _, lab = connected_components(A, directed=True, connection="weak")
    return lab


def largest_component_vertices(g: CSC) -> np.ndarray:
    lab = components(g)
    big = np.bincount(lab).argmax()
    return np.flatnonzero(lab == big).astype(np.int32)


def sample_sources(candidates: np.ndarray, count: int, seed: int) -> np.ndarray:
    """`count` distinct vertices drawn uniformly from candidates (numpy PCG64
    with a fixed seed; sources are also recorded with every result)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = min(count, len(candidates))
    return np.sort(rng.choice(candidates, size=k, replace=False)).astype(np.int32)


def nonzero_degree_vertices(g: CSC) -> np.ndarray:
    return np.flatnonzero(np.diff(g.CP) > 0).astype(np.int32)


def csc_to_cooc(g: CSC) -> tuple[np.ndarray, np.ndarray]:
    """COO sorted by column (the reference's COOC, TurboBFS_COOC_TD/bfstdcooc.h:29-32):
    edge I[k] -> J[k] with J non-decreasing."""
    J = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.CP))
    return g.IC.copy(), J


def transpose(g: CSC) -> CSC:
    """CSC of A^T (out-lists of g as in-lists), rows ascending."""
    J = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.CP))
    order = np.lexsort((J, g.IC))
    IC = J[order].astype(np.int32)
    CP = np.zeros(g.n + 1, dtype=np.int64)
    np.add.at(CP, g.IC.astype(np.int64) + 1, 1)
    return CSC(np.cumsum(CP).astype(np.int32), IC)
