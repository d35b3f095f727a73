"""Parity at the full sizes of the measured workloads.

The same graphs, sources and seeds as the benchmark: R-MAT scale 20
(n = 1,048,576, 31.4M CSC entries) single-source under every level loop and
the reference's kernel mappings; the 2048 x 2048 perturbed grid (3797 and
2309 levels from its first two sources) under the chained and graph-captured
loops; R-MAT scale 18 with 1024
sources in one batch; 512 Watts-Strogatz graphs as one block-diagonal CSC
(16.2M vertices, 162M entries) in one multi-root BFS; and a serving mix over
resident graphs of scales 18 and 20 plus a 1024 x 1024 grid.  Every checked
result is compared bit for bit with the oracle's O(n + m) BFS
(S, sigma mod 2^64, overflow flags); where checking every source would take
minutes on the host, the batch is also checked source by source against the
single-source engine (a different set of kernels) and against the batch's
own edge and vertex totals.
"""
import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import _lib as L, graphs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat20():
    g = graphs.rmat(20, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    src = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 64, seed=2)
    return g, src


def _same(res, exp, what):
    S, sig, ovf, lv = exp
    np.testing.assert_array_equal(res.S, S, err_msg=f"{what}: S")
    np.testing.assert_array_equal(res.sigma, sig, err_msg=f"{what}: sigma")
    np.testing.assert_array_equal(res.ovf, ovf, err_msg=f"{what}: ovf")
    assert res.stats["levels"] == lv, what


def test_rmat20_single_source_every_loop(gpu, rmat20):
    g, src = rmat20
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    deg = np.diff(g.CP).astype(np.int64)
    loops = [("chained", dict(loop=L.LOOP_CHAINED)),
             ("chained_no_lds", dict(loop=L.LOOP_CHAINED, fb_lds=-1)),
             ("graph", dict(loop=L.LOOP_GRAPH)),
             ("persistent", dict(loop=L.LOOP_PERSISTENT)),
             ("host", dict(loop=L.LOOP_HOST))]
    for r in src[:3]:
        exp = oracle.textbook(g.CP, g.IC, int(r))
        for name, kw in loops:
            print(f"[rmat20] root={r} loop={name}", flush=True)
            res = ctx.bfs(gr, int(r), tb.default_opts(mode=L.MODE_AUTO, pull=L.PULL_LB, **kw))
            _same(res, exp, f"{name} root {r}")
            assert res.stats["edges"] == int(deg[exp[0] >= 0].sum())
    ctx.close()


def test_rmat20_sixteen_sources_many(gpu, rmat20):
    """16 of the 64 benchmark sources, run back to back as the benchmark
    times them (tbfs_bfs_many, chained loop, launch shapes planned from the
    previous call), each checked against the oracle; the last source's
    arrays, and every source's level / vertex / edge counts."""
    g, src = rmat20
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    deg = np.diff(g.CP).astype(np.int64)
    roots = np.asarray(src[:16], np.int32)
    ctx.run_many(gr, roots[:4])  # history for the launch plan
    stats = ctx.run_many(gr, roots)
    for r, st in zip(roots, stats):
        S, sig, ovf, lv = oracle.textbook(g.CP, g.IC, int(r))
        assert st["levels"] == lv, int(r)
        assert st["reached"] == int((S >= 0).sum()), int(r)
        assert st["edges"] == int(deg[S >= 0].sum()), int(r)
    S, sig, ovf = ctx.read(g.n)
    exp = oracle.textbook(g.CP, g.IC, int(roots[-1]))
    np.testing.assert_array_equal(S, exp[0])
    np.testing.assert_array_equal(sig, exp[1])
    np.testing.assert_array_equal(ovf, exp[2])
    # and each of the 16 alone, arrays checked
    for r in roots:
        print(f"[rmat20] root={r} single", flush=True)
        _same(ctx.bfs(gr, int(r)), oracle.textbook(g.CP, g.IC, int(r)), f"root {r}")
    ctx.close()


def test_rmat20_reference_mappings(gpu, rmat20):
    """The reference's kernel mappings (one lane per column, one wavefront per
    column, edge-parallel COOC, push) and the depth-only bottom-up sweep."""
    g, src = rmat20
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    r = int(src[0])
    exp = oracle.textbook(g.CP, g.IC, r)
    modes = [("td_csc_sc", dict(mode=L.MODE_PULL_SC, loop=L.LOOP_HOST)),
             ("td_csc_wa", dict(mode=L.MODE_PULL_WA, loop=L.LOOP_HOST)),
             ("push", dict(mode=L.MODE_PUSH, pull=L.PULL_LB, loop=L.LOOP_CHAINED)),
             ("td_cooc", dict(mode=L.MODE_COOC, loop=L.LOOP_HOST))]
    for name, kw in modes:
        print(f"[rmat20] root={r} mode={name}", flush=True)
        _same(ctx.bfs(gr, r, tb.default_opts(**kw)), exp, name)
    print(f"[rmat20] root={r} depth only", flush=True)
    res = ctx.bfs(gr, r, tb.default_opts(mode=L.MODE_AUTO, pull=L.PULL_LB,
                                         loop=L.LOOP_CHAINED, alpha=1e9, sigma=0))
    np.testing.assert_array_equal(res.S, exp[0])
    assert res.stats["levels"] == exp[3]
    # the bottom-up entry points' mappings with bbreak (stop a column at its
    # first parent): depths only
    for name, mode in (("bu_csc_sc", L.MODE_PULL_SC), ("bu_csc_wa", L.MODE_PULL_WA)):
        for rr in (r, int(src[1])):
            e = exp if rr == r else oracle.textbook(g.CP, g.IC, rr)
            print(f"[rmat20] root={rr} mode={name} bbreak=1", flush=True)
            res = ctx.bfs(gr, rr, tb.default_opts(mode=mode, loop=L.LOOP_HOST, sigma=0,
                                                  bbreak=1))
            np.testing.assert_array_equal(res.S, e[0], err_msg=name)
            assert res.stats["levels"] == e[3], name
    ctx.close()


def test_rmat20_dropin_entry_points(gpu, rmat20):
    """The seven entry points at full size (host arrays in and out, the
    default kernels): S, and sigma in the reference's type, against the
    oracle, with the inexact status exactly where a count leaves it."""
    from tests.test_gpu_parity import _check_dropin
    g, src = rmat20
    print(f"[rmat20] drop-in root={int(src[0])}", flush=True)
    _check_dropin(g, int(src[0]), mappings=(L.DROPIN_FAST,))


def test_grid2048_high_diameter(gpu):
    g = graphs.grid(2048, 2048, 0.1, seed=3, relabel=True)
    src = graphs.sample_sources(graphs.largest_component_vertices(g), 16, seed=3)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    for r in src[:2]:
        exp = oracle.textbook(g.CP, g.IC, int(r))
        assert exp[3] > 2000  # thousands of levels, most of them tiny
        loops = [("chained", dict(loop=L.LOOP_CHAINED)), ("graph", dict(loop=L.LOOP_GRAPH))]
        if r == src[0]:  # the benchmark's other two loops, on one source
            loops += [("persistent_1024x64", dict(loop=L.LOOP_PERSISTENT, block=1024, grid=64)),
                      ("host", dict(loop=L.LOOP_HOST))]
        for name, kw in loops:
            print(f"[grid2048] root={r} loop={name} levels={exp[3]}", flush=True)
            res = ctx.bfs(gr, int(r), tb.default_opts(mode=L.MODE_AUTO, pull=L.PULL_LB, **kw))
            _same(res, exp, f"{name} root {r}")
    ctx.close()


def test_batch_rmat18_1024_sources(gpu):
    g = graphs.rmat(18, 16, 0.57, 0.19, 0.19, seed=4, relabel=True)
    src = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 1024, seed=4)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    b = tb.Batch(g.n, len(src))
    print(f"[batch rmat18] n={g.n} nz={g.nz} sources={len(src)}", flush=True)
    st = b.run(gr, src)
    ctx = tb.Context(g.n)
    deg = np.diff(g.CP).astype(np.int64)
    reached = edges = levels = 0
    for first in range(0, len(src), 64):
        S, sig, ovf = b.read(first, 64)
        for i in range(64):
            r = int(src[first + i])
            if (first + i) % 32 == 5:  # 32 sources against the oracle
                eS, esig, eovf, _ = oracle.textbook(g.CP, g.IC, r)
                np.testing.assert_array_equal(S[i], eS, err_msg=f"S of source {first + i}")
                np.testing.assert_array_equal(sig[i], esig, err_msg=f"sigma of {first + i}")
                np.testing.assert_array_equal(ovf[i], eovf, err_msg=f"ovf of {first + i}")
            one = ctx.bfs(gr, r)  # every source against the single-source engine
            np.testing.assert_array_equal(S[i], one.S, err_msg=f"S of source {first + i}")
            np.testing.assert_array_equal(sig[i], one.sigma, err_msg=f"sigma of {first + i}")
            np.testing.assert_array_equal(ovf[i], one.ovf, err_msg=f"ovf of {first + i}")
            reached += int((S[i] >= 0).sum())
            edges += int(deg[S[i] >= 0].sum())
            levels = max(levels, one.stats["levels"])
    assert st["reached"] == reached
    assert st["edges"] == edges
    assert st["levels"] == levels
    ctx.close()
    b.close()


def test_multigraph_512_watts_strogatz(gpu):
    rng = np.random.default_rng(5)
    parts = [graphs.watts_strogatz(int(rng.integers(1000, 64001)), 10, 0.1, seed=i)
             for i in range(512)]
    bd, off = graphs.block_diagonal(parts)
    roots = np.array([off[i] + int(rng.integers(0, parts[i].n)) for i in range(len(parts))],
                     np.int32)
    gb = tb.Graph(bd.CP, bd.IC, symmetric=True)
    ctx = tb.Context(bd.n)
    print(f"[multigraph] n={bd.n} nz={bd.nz} graphs={len(parts)}", flush=True)
    st = ctx.run_multi(gb, roots)
    S, sig, ovf = ctx.read(bd.n)
    levels = 0
    for i, p in enumerate(parts):
        eS, esig, eovf, lv = oracle.textbook(p.CP, p.IC, int(roots[i] - off[i]))
        a, z = int(off[i]), int(off[i]) + p.n
        np.testing.assert_array_equal(S[a:z], eS, err_msg=f"S of block {i}")
        np.testing.assert_array_equal(sig[a:z], esig, err_msg=f"sigma of block {i}")
        np.testing.assert_array_equal(ovf[a:z], eovf, err_msg=f"ovf of block {i}")
        levels = max(levels, lv)
    assert st["levels"] == levels
    assert st["reached"] == int((S >= 0).sum())
    ctx.close()


def test_serving_mix_at_scale(gpu):
    cs = [graphs.rmat(18, 16, 0.57, 0.19, 0.19, seed=18, relabel=True),
          graphs.rmat(20, 16, 0.57, 0.19, 0.19, seed=20, relabel=True),
          graphs.grid(1024, 1024, 0.1, seed=3, relabel=True)]
    gs = [tb.Graph(g.CP, g.IC, symmetric=True) for g in cs]
    srv = tb.Server(gs)
    rng = np.random.default_rng(5)
    qs = [(0, int(rng.integers(0, cs[0].n))) for _ in range(16)]
    qs += [(1, int(rng.integers(0, cs[1].n))) for _ in range(6)]
    qs += [(2, int(rng.integers(0, cs[2].n))) for _ in range(16)]
    order = rng.permutation(len(qs))
    store = {}
    print(f"[serve] queries={len(qs)} on rmat18 / rmat20 / grid1024", flush=True)
    for j in order:
        gi, r = qs[j]
        n = cs[gi].n
        bufs = (tb.host_array(n, np.int32), tb.host_array(n, np.uint64),
                tb.host_array(n, np.uint8))
        store[srv.submit(gi, r, *bufs)] = (gi, r) + bufs
    done = []
    while len(done) < len(qs):
        done += srv.poll(4096, wait_us=100000)
    assert max(d["batch"] for d in done) > 1  # queries on one graph shared a batch
    for d in done:
        assert d["status"] == 0, d
        gi, r, S, sig, ovf = store[d["ticket"]]
        eS, esig, eovf, _ = oracle.textbook(cs[gi].CP, cs[gi].IC, r)
        np.testing.assert_array_equal(S, eS, err_msg=f"S graph {gi} root {r}")
        np.testing.assert_array_equal(sig, esig, err_msg=f"sigma graph {gi} root {r}")
        np.testing.assert_array_equal(ovf, eovf, err_msg=f"ovf graph {gi} root {r}")
    srv.close()


def _serving_graphs():
    """The serving benchmark's eight resident graphs: R-MAT scales 16..22
    (ef16, a=.57 b=.19 c=.19, symmetrized, relabelled, seed = scale) and a
    1024 x 1024 grid with 10% of its edges removed (seed 3)."""
    cs = [graphs.rmat(sc, 16, 0.57, 0.19, 0.19, seed=sc, relabel=True) for sc in range(16, 23)]
    cs.append(graphs.grid(1024, 1024, 0.1, seed=3, relabel=True))
    return cs


@pytest.mark.timeout(600)
def test_serving_benchmark_graph_set(gpu):
    """The serving benchmark's graph set under the default server: every graph
    is queried (seed 5) with results checked bit for bit against the oracle,
    and the graphs whose batch cap is below 1024 sources (R-MAT 21 and 22:
    512 and 256 with the default budget) get bursts larger than the cap, so
    batches at the cap run too; a checked query is the same in a full batch
    as alone."""
    cs = _serving_graphs()
    gs = [tb.Graph(g.CP, g.IC, symmetric=True) for g in cs]
    srv = tb.Server(gs)
    caps = [srv.info(i)["batch_cap"] for i in range(len(gs))]
    print(f"[serve8] caps={caps}", flush=True)
    rng = np.random.default_rng(5)
    store = {}
    capped = {}
    # the whole set is queued before any of it runs: the graphs with more
    # queries than their cap then run batches at the cap
    srv.hold(True)
    for gi, g in enumerate(cs):
        n = g.n
        extra = 2 * caps[gi] if caps[gi] < 1024 else 0  # unchecked: fill the cap
        roots = [int(rng.integers(0, n)) for _ in range(4)]
        for r in roots[:2]:
            bufs = (tb.host_array(n, np.int32), tb.host_array(n, np.uint64),
                    tb.host_array(n, np.uint8))
            store[srv.submit(gi, r, *bufs)] = (gi, r) + bufs
        for _ in range(extra):
            capped[srv.submit(gi, int(rng.integers(0, n)))] = gi
        for r in roots[2:]:
            bufs = (tb.host_array(n, np.int32), tb.host_array(n, np.uint64),
                    tb.host_array(n, np.uint8))
            store[srv.submit(gi, r, *bufs)] = (gi, r) + bufs
    srv.hold(False)
    total = len(store) + len(capped)
    done = []
    while len(done) < total:
        done += srv.poll(8192, wait_us=100000)
    assert all(d["status"] == 0 for d in done)
    for gi in range(len(cs)):
        if caps[gi] < 1024:
            sizes = [d["batch"] for d in done if d["graph"] == gi]
            assert max(sizes) == caps[gi], (gi, caps[gi], max(sizes))
    srv.close()
    for d in done:
        if d["ticket"] not in store:
            continue
        gi, r, S, sig, ovf = store[d["ticket"]]
        print(f"[serve8] check graph={gi} root={r} batch={d['batch']}", flush=True)
        eS, esig, eovf, _ = oracle.textbook(cs[gi].CP, cs[gi].IC, r)
        np.testing.assert_array_equal(S, eS, err_msg=f"S graph {gi} root {r}")
        np.testing.assert_array_equal(sig, esig, err_msg=f"sigma graph {gi} root {r}")
        np.testing.assert_array_equal(ovf, eovf, err_msg=f"ovf graph {gi} root {r}")
    for x in gs:
        x.close()
