"""Matrix Market reader producing the reference's CSC / COOC layouts.

Follows the reference's input path (paths relative to the reference root):
banner and size line (TurboBFS_CSC_TD/mmiof.c:97-218), 1-based "row col
[value]" entries (TurboBFS_CSC_TD/readMMfilef.c:33-70), and for undirected
graphs (`ug`) the symmetric lower triangle expanded to both directions
(SSC -> CSC, TurboBFS_CSC_TD/sparseformatransf.c:60-104; nz doubled,
TurboBFS_CSC_TD/bfstdcsc_main.c:123).  Entry u -> v (MM row u, column v)
lands in CSC column v.

Differences, by design: entries are stably sorted by column (the reference
requires a column-sorted file, TurboBFS_CSC_TD/bfstdcsc_main.c:159-161,
:268, and builds a wrong CSC otherwise), undirected columns come out with
rows ascending (what the reference produces for sorted files), the value
column is detected from the banner instead of a `w` argument, and the
CP[n+1] write past the end in CP_RP_ug/CP_RP_dg (sparseformatransf.c:96-101,
:114-117) has no counterpart.
"""
from __future__ import annotations

import gzip

import numpy as np

from .graphs import CSC


def _open(path: str):
    return gzip.open(path, "rt") if path.endswith(".gz") else open(path, "r")


class MMError(ValueError):
    pass


def read_banner(path: str) -> dict:
    with _open(path) as f:
        line = f.readline()
    tok = line.strip().split()
    if len(tok) < 5 or tok[0] != "%%MatrixMarket":
        raise MMError(f"{path}: not a Matrix Market file")
    obj, fmt, field, sym = (t.lower() for t in tok[1:5])
    if obj != "matrix" or fmt != "coordinate":
        raise MMError(f"{path}: only sparse coordinate matrices are supported")
    if field == "complex":
        raise MMError(f"{path}: complex matrices are not supported")
    return {"field": field, "symmetry": sym}


def read_entries(path: str):
    """Returns (M, N, rows0, cols0) with 0-based int32 indices in file order."""
    ban = read_banner(path)
    with _open(path) as f:
        size = None
        for line in f:
            if line.startswith("%") or not line.strip():
                continue
            size = line.split()
            break
        if size is None or len(size) < 3:
            raise MMError(f"{path}: missing size line")
        M, N, nz = (int(x) for x in size[:3])
        data = np.loadtxt(f, dtype=np.float64, ndmin=2, max_rows=nz) if nz > 0 \
            else np.zeros((0, 2))
    if data.shape[0] != nz:
        raise MMError(f"{path}: expected {nz} entries, found {data.shape[0]}")
    ncol = 2 if ban["field"] == "pattern" else 3
    if nz and data.shape[1] < ncol:
        raise MMError(f"{path}: entries have {data.shape[1]} columns, expected {ncol}")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    if nz and (rows.min() < 0 or cols.min() < 0 or rows.max() >= M or cols.max() >= N):
        raise MMError(f"{path}: index out of range")
    return M, N, rows.astype(np.int32), cols.astype(np.int32), ban


def read_csc(path: str, ug: bool | None = None) -> CSC:
    """CSC in-lists of the graph in `path`.  ug=None takes the banner's
    symmetry; ug=True treats the entries as the lower triangle of an
    undirected graph (both directions stored)."""
    M, N, rows, cols, ban = read_entries(path)
    if ug is None:
        ug = ban["symmetry"] == "symmetric"
    n = max(M, N)
    if ug:
        src = np.concatenate([rows, cols])
        dst = np.concatenate([cols, rows])
        order = np.lexsort((src, dst))  # by column, rows ascending
    else:
        src, dst = rows, cols
        order = np.argsort(dst, kind="stable")  # by column, file order inside
    IC = src[order].astype(np.int32)
    cnt = np.bincount(dst, minlength=n).astype(np.int64)
    CP = np.zeros(n + 1, dtype=np.int64)
    CP[1:] = np.cumsum(cnt)
    if CP[-1] > 2**31 - 1:
        raise MMError(f"{path}: more than 2^31-1 entries")
    return CSC(CP.astype(np.int32), IC)


def read_cooc(path: str, ug: bool | None = None):
    """(I, J, n): COO sorted by column, edge I[k] -> J[k]
    (TurboBFS_COOC_TD/sparseformatransf.c:62-95)."""
    g = read_csc(path, ug)
    J = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.CP))
    return g.IC.copy(), J, g.n
