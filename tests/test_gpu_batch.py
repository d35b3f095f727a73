"""Batched sources (tbfs_batch_*): many roots per BFS, 64 per frontier word.

Every source's S and sigma (uint64 + overflow flag) must equal the oracle's
single-source BFS from that root, bit for bit; the batch counters must equal
the sums over sources.  Cases: ragged batch sizes (one word, partial words,
the full 1024), repeated roots, directed input, in-lists split over several
slices (hubs), counts past 2^64, consecutive batches of different shapes.
"""
import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import graphs, mtx
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _csc_from_edges(n, src, dst, symmetric=True):
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    if symmetric:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    key = np.unique(dst * n + src)
    col, row = key // n, key % n
    CP = np.zeros(n + 1, np.int32)
    np.add.at(CP, col + 1, 1)
    CP = np.cumsum(CP).astype(np.int32)
    return graphs.CSC(CP, row.astype(np.int32))


def _expected(g, roots):
    S = np.empty((len(roots), g.n), np.int32)
    sig = np.empty((len(roots), g.n), np.uint64)
    ovf = np.empty((len(roots), g.n), np.uint8)
    deg = np.diff(g.CP).astype(np.int64)
    edges = 0
    levels = 0
    for i, r in enumerate(roots):
        S[i], sig[i], ovf[i], lv = oracle.textbook(g.CP, g.IC, int(r))
        edges += int(deg[S[i] >= 0].sum())
        levels = max(levels, lv)
    return S, sig, ovf, edges, levels


def _check_batch(g, roots, batch=None, symmetric=True, first=0, count=None, alpha=0.0,
                 order=0):
    roots = np.asarray(roots, np.int32)
    gr = tb.Graph(g.CP, g.IC, symmetric=symmetric)
    b = batch if batch is not None else tb.Batch(g.n, max(64, len(roots)))
    print(f"[batch] n={g.n} nz={g.nz} sources={len(roots)} alpha={alpha} order={order}",
          flush=True)
    st = b.run(gr, roots, tb.default_opts(batch_alpha=alpha, batch_order=order))
    S, sig, ovf = b.read(first, count)
    k = S.shape[0]
    eS, esig, eovf, edges, levels = _expected(g, roots[first:first + k])
    for i in range(k):
        assert np.array_equal(S[i], eS[i]), f"S of source {first + i} (root {roots[first + i]})"
        assert np.array_equal(sig[i], esig[i]), f"sigma of source {first + i}"
        assert np.array_equal(ovf[i], eovf[i]), f"ovf of source {first + i}"
    if first == 0 and k == len(roots):
        assert st["reached"] == int((eS >= 0).sum())
        assert st["edges"] == edges
        assert st["levels"] == levels
        assert (st["ovf64"] != 0) == bool(eovf.any())
    return st


@pytest.mark.parametrize("count", [1, 37, 64, 100, 256])
def test_batch_er4096_sizes(gpu, count):
    g = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), count, seed=count)
    _check_batch(g, roots)


def test_batch_er4096_full_1024_with_repeats(gpu):
    g = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    rng = np.random.default_rng(7)
    roots = rng.integers(0, g.n, size=1024).astype(np.int32)
    roots[5] = roots[700] = roots[1023] = 0   # repeated roots, across words
    _check_batch(g, roots)


def test_batch_rmat12_and_golden_vectors(gpu):
    """R-MAT scale 12; the first roots are those of the golden vectors that
    the reference's own sequential code wrote (depths, -1 = unreached)."""
    fx = np.load(golden("rmat12_seed2.npz"))
    g = graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    extra = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 200, seed=2)
    roots = np.concatenate([fx["roots"].astype(np.int32), extra])
    _check_batch(g, roots)
    b = tb.Batch(g.n, len(roots))
    b.run(tb.Graph(g.CP, g.IC, symmetric=True), roots)
    S, sig, _ = b.read(0, len(fx["roots"]))
    for i in range(len(fx["roots"])):
        assert np.array_equal(S[i], fx["S_bu"][i])
        assert np.array_equal(sig[i].astype(np.float64), fx["sigma_td"][i].astype(np.float64))


def test_batch_rmat14_hubs_split(gpu):
    # R-MAT hubs have in-lists much longer than one pull slice
    g = graphs.rmat(14, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    assert np.diff(g.CP).max() > 1024
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 130, seed=4)
    _check_batch(g, roots)


def test_batch_star_and_path(gpu):
    # a 5000-leaf star (its hub's in-list spans 20 slices) joined to a path
    leaves = 5000
    n = leaves + 1 + 40
    src = [0] * leaves + list(range(leaves + 1, n - 1)) + [leaves]
    dst = list(range(1, leaves + 1)) + list(range(leaves + 2, n)) + [leaves + 1]
    g = _csc_from_edges(n, src, dst)
    roots = [0, 1, leaves, n - 1, 17, 0] + list(range(2000, 2060))
    _check_batch(g, roots)


def test_batch_grid_counts_past_2_64(gpu):
    # 32-bit counts overflow first: the batch runs again with 64-bit counts
    g = graphs.grid(48, 48, 0.0, seed=0, relabel=False)
    roots = [0, 47, 48 * 47, 48 * 48 - 1, 1000, 5] * 20
    st = _check_batch(g, roots)
    assert st["ovf64"] != 0
    assert st["count_bits"] == 64


def test_batch_count_widths(gpu):
    """Forced 64-bit counts give the same results as the default; forced
    32-bit counts give every count modulo 2^32 with the flag raised exactly
    where the count reaches 2^32."""
    g = graphs.rmat(13, 16, 0.57, 0.19, 0.19, seed=6, relabel=True)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 80, seed=6)
    st = _check_batch(g, roots)
    assert st["count_bits"] == 32
    b = tb.Batch(g.n, 80)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    st64 = b.run(gr, roots, tb.default_opts(batch_count_bits=64))
    assert st64["count_bits"] == 64
    S64, sig64, ovf64 = b.read()
    b.run(gr, roots, tb.default_opts(batch_count_bits=32))
    S32, sig32, ovf32 = b.read()
    assert np.array_equal(S64, S32) and np.array_equal(sig64, sig32)
    eS, esig, eovf, _, _ = _expected(g, roots)
    assert np.array_equal(S64, eS) and np.array_equal(sig64, esig)

    grid = graphs.grid(40, 40, 0.0, seed=0, relabel=False)
    r = np.array([0, 39, 40 * 40 - 1], np.int32)
    bg = tb.Batch(grid.n, 64)
    st = bg.run(tb.Graph(grid.CP, grid.IC, symmetric=True), r,
                tb.default_opts(batch_count_bits=32))
    assert st["count_bits"] == 32 and st["ovf64"] != 0
    S, sig, ovf = bg.read()
    eS, esig, eovf, _, _ = _expected(grid, r)
    assert np.array_equal(S, eS)
    assert np.array_equal(sig, esig & np.uint64(0xFFFFFFFF))
    assert np.array_equal(ovf, (eovf.astype(bool) | (esig >= np.uint64(2**32))).astype(np.uint8))


def test_batch_perturbed_grid(gpu):
    g = graphs.grid(96, 64, 0.1, seed=3, relabel=True)
    roots = graphs.sample_sources(graphs.largest_component_vertices(g), 70, seed=3)
    _check_batch(g, roots)


def test_batch_directed_tina(gpu):
    g = mtx.read_csc(golden("Tina_AskCal.mtx"))
    roots = np.arange(g.n, dtype=np.int32)
    _check_batch(g, roots, symmetric=False)


@pytest.mark.parametrize("alpha", [1e-9, 4.0, 1e12])
@pytest.mark.parametrize("case", ["rmat14", "grid_overflow", "directed_tina", "star"])
def test_batch_direction_invariance(gpu, case, alpha):
    """Every level top-down (alpha ~ 0), the default switch, every level
    bottom-up (alpha huge): identical results."""
    sym = True
    if case == "rmat14":
        g = graphs.rmat(14, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
        roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 100, seed=9)
    elif case == "grid_overflow":
        g = graphs.grid(48, 48, 0.0, seed=0, relabel=False)
        roots = np.array([0, 47, 48 * 48 - 1, 1000] * 17, np.int32)
    elif case == "directed_tina":
        g = mtx.read_csc(golden("Tina_AskCal.mtx"))
        roots = np.arange(g.n, dtype=np.int32)
        sym = False
    else:
        leaves = 3000
        n = leaves + 1 + 40
        src = [0] * leaves + list(range(leaves + 1, n - 1)) + [leaves]
        dst = list(range(1, leaves + 1)) + list(range(leaves + 2, n)) + [leaves + 1]
        g = _csc_from_edges(n, src, dst)
        roots = np.array([0, 5, leaves + 3, n - 1] * 20, np.int32)
    st = _check_batch(g, roots, symmetric=sym, alpha=alpha)
    if alpha < 1e-6:
        assert st["pull_levels"] == 0 and st["push_levels"] == st["levels"]
    if alpha > 1e6:  # a frontier without out-edges (if any) still goes top-down
        assert st["pull_levels"] >= st["levels"] - 1


@pytest.mark.parametrize("order", [-1, 1])
@pytest.mark.parametrize("case", ["rmat13", "grid_overflow", "directed_tina", "star"])
def test_batch_degree_order(gpu, case, order):
    """Batches on the degree-ordered copy of the graph (order 1; by default
    only for skewed degrees) and on the graph itself (order -1) give every
    source exactly the oracle's result in the caller's numbering."""
    sym = True
    if case == "rmat13":
        g = graphs.rmat(13, 16, 0.57, 0.19, 0.19, seed=6, relabel=True)
        roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 130, seed=6)
    elif case == "grid_overflow":
        g = graphs.grid(48, 48, 0.0, seed=0, relabel=False)
        roots = np.array([0, 47, 48 * 48 - 1, 1000, 5], np.int32)
    elif case == "directed_tina":
        g = mtx.read_csc(golden("Tina_AskCal.mtx"))
        roots = np.arange(g.n, dtype=np.int32)
        sym = False
    else:
        n = 3000
        g = _csc_from_edges(n, np.zeros(n - 1, np.int64), np.arange(1, n))
        roots = np.array([0, 1, 2999, 17], np.int32)
    _check_batch(g, roots, symmetric=sym, order=order)


def test_batch_degree_order_default_and_invalid(gpu):
    g = graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 64, seed=1)
    _check_batch(g, roots, order=0)  # skewed: runs on the copy
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    b = tb.Batch(g.n, 64)
    with pytest.raises(tb.TbfsError):
        b.run(gr, roots, tb.default_opts(batch_order=2))


def test_batch_autotune_keeps_results(gpu):
    g = graphs.rmat(13, 16, 0.57, 0.19, 0.19, seed=8, relabel=True)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 128, seed=8)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    b = tb.Batch(g.n, 128)
    best, ms = b.autotune(gr, roots)
    assert ms > 0 and best.batch_alpha in (4.0, 32.0, 128.0)
    S0, sig0, ovf0 = b.read()          # the last candidate's results
    b.run(gr, roots, best)
    S1, sig1, ovf1 = b.read()
    assert np.array_equal(S0, S1) and np.array_equal(sig0, sig1) and np.array_equal(ovf0, ovf1)
    eS, esig, eovf, _, _ = _expected(g, roots)
    assert np.array_equal(S1, eS) and np.array_equal(sig1, esig)


def test_batch_reuse_and_partial_read(gpu):
    g1 = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    g2 = graphs.grid(40, 30, 0.05, seed=1, relabel=True)
    b = tb.Batch(4096, 512)
    r1 = graphs.sample_sources(graphs.nonzero_degree_vertices(g1), 500, seed=11)
    _check_batch(g1, r1, batch=b)
    r2 = graphs.sample_sources(graphs.largest_component_vertices(g2), 65, seed=12)
    _check_batch(g2, r2, batch=b)
    _check_batch(g1, r1[:130], batch=b, first=64, count=66)


def test_batch_equals_single_source_engine(gpu):
    g = graphs.rmat(13, 16, 0.57, 0.19, 0.19, seed=5, relabel=True)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 96, seed=5)
    b = tb.Batch(g.n, 96)
    b.run(gr, roots)
    S, sig, ovf = b.read()
    ctx = tb.Context(g.n)
    for i in range(0, 96, 7):
        res = ctx.bfs(gr, int(roots[i]))
        assert np.array_equal(S[i], res.S)
        assert np.array_equal(sig[i], res.sigma)
        assert np.array_equal(ovf[i], res.ovf)


def test_batch_invalid_inputs(gpu):
    g = graphs.erdos_renyi(256, 0.05, seed=3)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    b = tb.Batch(256, 64)
    with pytest.raises(tb.TbfsError):
        b.run(gr, np.zeros(65, np.int32))       # more than b_cap
    with pytest.raises(tb.TbfsError):
        b.run(gr, np.array([0, 256], np.int32))  # root out of range
    with pytest.raises(tb.TbfsError):
        b.run(gr, np.zeros(0, np.int32))         # empty batch
    big = graphs.erdos_renyi(300, 0.05, seed=3)
    with pytest.raises(tb.TbfsError):
        b.run(tb.Graph(big.CP, big.IC, symmetric=True), np.zeros(3, np.int32))  # n > n_cap
    with pytest.raises(tb.TbfsError):
        tb.Batch(256, 1025)
    st = b.run(gr, np.array([3, 3, 4], np.int32))
    assert st["levels"] > 0
    with pytest.raises(tb.TbfsError):
        b.read(2, 2)                             # past the last batch's sources


# ---------------------------------------------------------------------------
# many roots in one BFS (tbfs_bfs_multi): block-diagonal packing of graphs


def _with_super_source(g, roots):
    """g plus a vertex n with an edge n -> r for every distinct root r: a BFS
    from n has S' = S_multi + 1 and the same path counts."""
    n = g.n
    rs = np.unique(np.asarray(roots, np.int64))
    deg = np.diff(g.CP).astype(np.int64)
    add = np.zeros(n + 1, np.int64)
    add[rs] = 1
    newdeg = np.concatenate([deg, [0]]) + add
    CP = np.zeros(n + 2, np.int64)
    CP[1:] = np.cumsum(newdeg)
    IC = np.empty(int(CP[-1]), np.int32)
    for v in range(n):
        a, b = int(g.CP[v]), int(g.CP[v + 1])
        IC[CP[v]:CP[v] + (b - a)] = g.IC[a:b]
        if add[v]:
            IC[CP[v + 1] - 1] = n
    return graphs.CSC(CP.astype(np.int32), IC)


def test_multi_root_block_diagonal_ws(gpu):
    rng = np.random.default_rng(8)
    parts = [graphs.watts_strogatz(int(rng.integers(100, 3000)), 10, 0.1, seed=i)
             for i in range(24)]
    g, off = graphs.block_diagonal(parts)
    roots = np.array([off[i] + (i * 37) % parts[i].n for i in range(len(parts))], np.int32)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    st = ctx.run_multi(gr, roots)
    S, sig, ovf = ctx.read(g.n)
    levels = 0
    for i, p in enumerate(parts):
        eS, esig, eovf, lv = oracle.textbook(p.CP, p.IC, int(roots[i] - off[i]))
        a, b = int(off[i]), int(off[i + 1])
        assert np.array_equal(S[a:b], eS), f"block {i}"
        assert np.array_equal(sig[a:b], esig), f"block {i}"
        assert np.array_equal(ovf[a:b], eovf)
        levels = max(levels, lv)
    assert st["levels"] == levels
    assert st["reached"] == int((S >= 0).sum())


@pytest.mark.parametrize("case", ["rmat13", "grid_overflow", "directed_tina"])
def test_multi_root_equals_super_source(gpu, case):
    if case == "rmat13":
        g = graphs.rmat(13, 16, 0.57, 0.19, 0.19, seed=3, relabel=True)
        roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 50, seed=1)
        roots = np.concatenate([roots, roots[:5]])          # repeats
        sym = True
    elif case == "grid_overflow":
        g = graphs.grid(48, 48, 0.0, seed=0, relabel=False)
        roots = np.array([0, 48 * 48 - 1], np.int32)
        sym = True
    else:
        g = mtx.read_csc(golden("Tina_AskCal.mtx"))
        roots = np.array([10, 3], np.int32)
        sym = False
    gr = tb.Graph(g.CP, g.IC, symmetric=sym)
    ctx = tb.Context(g.n)
    ctx.run_multi(gr, roots)
    S, sig, ovf = ctx.read(g.n)
    h = _with_super_source(g, roots)
    eS, esig, eovf, _ = oracle.textbook(h.CP, h.IC, g.n)
    eS = eS[:g.n]
    assert np.array_equal(S, np.where(eS >= 0, eS - 1, -1))
    assert np.array_equal(sig, esig[:g.n])
    assert np.array_equal(ovf, eovf[:g.n])


def test_multi_root_after_graph_loop_then_close(gpu):
    """Regression: a multi-root run with more roots than the context's root
    buffer holds replaces that buffer, which a captured level loop reads its
    root from.  The capture must be dropped (and captured again by the next
    graph-loop BFS), nothing else freed twice, and closing the context must
    leave no error behind for the next context."""
    g = graphs.rmat(10, 16, 0.57, 0.19, 0.19, seed=3, relabel=True)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    graph_loop = tb.default_opts(loop=tb._lib.LOOP_GRAPH)

    def same(res, r):
        S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, r)
        assert np.array_equal(res.S, S) and np.array_equal(res.sigma, sig)
        assert np.array_equal(res.ovf, ovf)

    ctx = tb.Context(g.n)
    same(ctx.bfs(gr, 5, graph_loop), 5)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 40, seed=2)
    ctx.run_multi(gr, roots)
    S, sig, ovf = ctx.read(g.n)
    h = _with_super_source(g, roots)
    eS, esig, _, _ = oracle.textbook(h.CP, h.IC, g.n)
    assert np.array_equal(S, np.where(eS[:g.n] >= 0, eS[:g.n] - 1, -1))
    assert np.array_equal(sig, esig[:g.n])
    same(ctx.bfs(gr, 7, graph_loop), 7)
    ctx.close()
    ctx2 = tb.Context(g.n)
    same(ctx2.bfs(gr, 9), 9)
    same(ctx2.bfs(gr, 11, graph_loop), 11)
    ctx2.close()


def test_multi_root_invalid(gpu):
    g = graphs.erdos_renyi(256, 0.05, seed=3)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(256)
    with pytest.raises(tb.TbfsError):
        ctx.run_multi(gr, np.array([0, 300], np.int32))
    with pytest.raises(tb.TbfsError):
        ctx.run_multi(gr, np.zeros(0, np.int32))
    with pytest.raises(tb.TbfsError):
        ctx.run_multi(gr, np.array([1], np.int32), tb.default_opts(mode=1))  # PULL_SC
