"""tbfs_bfs_noalloc, the entry point that allocates nothing, against
tbfs_bfs (the entry point it is a variant of) through the C-ABI: exact S,
sigma and overflow flags and the same counters, on two calls in a row, the
second on a smaller graph (the library's scratch array then holds state from
the larger one); a graph over the scratch array's bound is refused."""
import ctypes as C

import numpy as np
import pytest

import turbobfs_amd as tb
from turbobfs_amd import _lib, graphs

pytestmark = pytest.mark.gpu


def _device_outputs(n):
    lib = _lib.load()
    ptrs = [lib.tbfs_device_alloc(n * w) for w in (4, 8, 1)]
    assert all(ptrs), "tbfs_device_alloc failed"
    for p, w in zip(ptrs, (4, 8, 1)):  # poison: every entry must be written
        assert lib.tbfs_device_fill(C.c_void_p(p), 0x5A, n * w) == 0
    return ptrs


def _to_host(ptrs, n):
    lib = _lib.load()
    S = np.empty(n, np.int32)
    sig = np.empty(n, np.uint64)
    ovf = np.empty(n, np.uint8)
    for a, p in zip((S, sig, ovf), ptrs):
        assert lib.tbfs_device_copy(a.ctypes.data_as(C.c_void_p), C.c_void_p(p), a.nbytes) == 0
    return S, sig, ovf


def _noalloc(gr, root, ptrs, opts=None):
    lib = _lib.load()
    st = _lib.Stats()
    o = opts if opts is not None else tb.default_opts()
    rc = lib.tbfs_bfs_noalloc(gr.handle, int(root), C.c_void_p(ptrs[0]), C.c_void_p(ptrs[1]),
                              C.c_void_p(ptrs[2]), None, C.byref(o), C.byref(st))
    return rc, st


def test_noalloc_matches_tbfs_bfs_two_calls(gpu):
    lib = _lib.load()
    big = graphs.rmat(16, 16, 0.57, 0.19, 0.19, seed=3, relabel=True)
    small = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    deep = graphs.grid(48, 48, 0.0, seed=0, relabel=False)  # counts past 2^64: flags
    cases = [(big, int(graphs.nonzero_degree_vertices(big)[7])), (small, 0), (deep, 0)]
    for g, r in cases:
        print(f"[noalloc] n={g.n} nz={g.nz} root={r}", flush=True)
        gr = tb.Graph(g.CP, g.IC, symmetric=True)
        ctx = tb.Context(g.n)
        ref = ctx.bfs(gr, r)
        ptrs = _device_outputs(g.n)
        try:
            rc, st = _noalloc(gr, r, ptrs)
            assert rc == 0, rc
            S, sig, ovf = _to_host(ptrs, g.n)
        finally:
            for p in ptrs:
                lib.tbfs_device_free(C.c_void_p(p))
        np.testing.assert_array_equal(S, ref.S)
        np.testing.assert_array_equal(sig, ref.sigma)
        np.testing.assert_array_equal(ovf, ref.ovf)
        assert st.levels == ref.stats["levels"]
        assert st.reached == ref.stats["reached"]
        assert st.edges == ref.stats["edges"]
        ctx.close()
        gr.close()


def test_noalloc_rejects_graph_over_bound(gpu):
    """One vertex over the bound (an edgeless graph): refused on the host, by
    the size check that comes before any launch."""
    lib = _lib.load()
    n = (1 << 24) + 1
    g = graphs.CSC(np.zeros(n + 1, np.int32), np.zeros(0, np.int32))
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ptrs = [lib.tbfs_device_alloc(64) for _ in range(3)]
    try:
        rc, _ = _noalloc(gr, 0, ptrs)
        assert rc == _lib.TBFS_ERR_INVALID
    finally:
        for p in ptrs:
            lib.tbfs_device_free(C.c_void_p(p))
    gr.close()
