"""Seeded random graphs against the oracle, through every device path.

Sixty graphs drawn from a fixed seed: Erdos-Renyi, R-MAT (skewed, with hubs
past the slice threshold), perturbed grids (long thin tails), Watts-Strogatz,
and directed graphs with self-loops, duplicate-free random arcs and isolated
vertices.  Each goes through the chained loop (default options), the
host-driven loop, the persistent loop, the graph-captured loop, a batch of
random roots (repeats included) and a multi-root BFS whose roots sit in
different components; every S, sigma (mod 2^64) and overflow flag must equal
the oracle's, and components without a root stay unreached.  Every third
graph also goes through the seven reference entry points, and twelve of them
are served together to a burst of random queries.
"""
import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import _lib as L, graphs

pytestmark = pytest.mark.gpu


def _directed(n, m, rng):
    src = rng.integers(0, n, size=m)
    dst = rng.integers(0, n, size=m)
    key = np.unique(dst.astype(np.int64) * n + src)  # column = head, row = tail
    col, row = key // n, key % n
    CP = np.zeros(n + 1, np.int64)
    CP[1:] = np.cumsum(np.bincount(col, minlength=n))
    return graphs.CSC(CP.astype(np.int32), row.astype(np.int32))


def _draw(i):
    rng = np.random.default_rng(20261203 + i)
    kind = i % 5
    if kind == 0:
        n = int(rng.integers(1, 3000))
        return f"er{n}", graphs.erdos_renyi(n, float(rng.uniform(0.2, 12.0)) / max(n - 1, 1),
                                            seed=i), True
    if kind == 1:
        sc = int(rng.integers(5, 12))
        ef = int(rng.integers(1, 24))
        return f"rmat{sc}x{ef}", graphs.rmat(sc, ef, 0.57, 0.19, 0.19, seed=i,
                                             relabel=bool(rng.integers(0, 2))), True
    if kind == 2:
        w, h = int(rng.integers(2, 90)), int(rng.integers(2, 90))
        return f"grid{w}x{h}", graphs.grid(w, h, float(rng.uniform(0.0, 0.3)), seed=i,
                                           relabel=bool(rng.integers(0, 2))), True
    if kind == 3:
        n = int(rng.integers(12, 2500))
        return f"ws{n}", graphs.watts_strogatz(n, 2 * int(rng.integers(1, 5)),
                                               float(rng.uniform(0, 0.5)), seed=i), True
    n = int(rng.integers(2, 2500))
    return f"dir{n}", _directed(n, int(rng.integers(n // 2, 8 * n)), rng), False


@pytest.mark.parametrize("case", range(60))
def test_random_graph_every_path(gpu, case):
    name, g, symmetric = _draw(case)
    rng = np.random.default_rng(case)
    gr = tb.Graph(g.CP, g.IC, symmetric=symmetric)
    roots = rng.integers(0, g.n, size=3).astype(np.int32)
    exp = {int(r): oracle.textbook(g.CP, g.IC, int(r)) for r in roots}
    ctx = tb.Context(g.n)
    print(f"[random] {name} n={g.n} nz={g.nz} roots={roots.tolist()}", flush=True)
    for loop in (L.LOOP_CHAINED, L.LOOP_HOST, L.LOOP_PERSISTENT, L.LOOP_GRAPH):
        for r in roots:
            res = ctx.bfs(gr, int(r), tb.default_opts(loop=loop))
            S, sig, ovf, lv = exp[int(r)]
            np.testing.assert_array_equal(res.S, S, err_msg=f"{name} loop {loop} root {r}")
            np.testing.assert_array_equal(res.sigma, sig, err_msg=f"{name} loop {loop} root {r}")
            np.testing.assert_array_equal(res.ovf, ovf, err_msg=f"{name} loop {loop} root {r}")
            assert res.stats["levels"] == lv
    # a batch of random roots, repeats included
    k = int(rng.integers(1, 131))
    broots = rng.integers(0, g.n, size=k).astype(np.int32)
    b = tb.Batch(g.n, k)
    b.run(gr, broots)
    bS, bsig, bovf = b.read()
    for i, r in enumerate(broots):
        S, sig, ovf, _ = exp[int(r)] if int(r) in exp else oracle.textbook(g.CP, g.IC, int(r))
        np.testing.assert_array_equal(bS[i], S, err_msg=f"{name} batch source {i} root {r}")
        np.testing.assert_array_equal(bsig[i], sig, err_msg=f"{name} batch source {i}")
        np.testing.assert_array_equal(bovf[i], ovf, err_msg=f"{name} batch source {i}")
    b.close()
    # multi-root: one root in each of up to 8 components; S / sigma of each
    # component equal that root's single-source result there
    if symmetric:
        comp = graphs.components(g)
        ids = np.unique(comp)[:8]
        mroots = np.array([int(np.flatnonzero(comp == c)[0]) for c in ids], np.int32)
        ctx.run_multi(gr, mroots)
        mS, msig, movf = ctx.read(g.n)
        for r, c in zip(mroots, ids):
            S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, int(r))
            sel = comp == c
            np.testing.assert_array_equal(mS[sel], S[sel], err_msg=f"{name} multi root {r}")
            np.testing.assert_array_equal(msig[sel], sig[sel], err_msg=f"{name} multi root {r}")
            np.testing.assert_array_equal(movf[sel], ovf[sel], err_msg=f"{name} multi root {r}")
        rest = ~np.isin(comp, ids)
        assert (mS[rest] == -1).all() and (msig[rest] == 0).all() and (movf[rest] == 0).all()
    ctx.close()


@pytest.mark.parametrize("case", range(0, 60, 3))
def test_random_graph_dropin(gpu, case):
    """The seven reference entry points on every third graph, from a random
    root: the reference's output types, exact where representable and the
    inexact status where not."""
    from tests.test_gpu_parity import _check_dropin
    name, g, symmetric = _draw(case)
    r = int(np.random.default_rng(case + 1000).integers(0, g.n))
    print(f"[random drop-in] {name} n={g.n} nz={g.nz} root={r}", flush=True)
    _check_dropin(g, r)


def test_random_graphs_served(gpu):
    """Twelve of the graphs resident in one server, 150 queries on random
    (graph, root) pairs submitted at once: every result equals the oracle's."""
    drawn = [_draw(i) for i in range(12)]
    cs = [g for _, g, _ in drawn]
    srv = tb.Server([tb.Graph(g.CP, g.IC, symmetric=sym) for _, g, sym in drawn])
    rng = np.random.default_rng(77)
    store = {}
    for _ in range(150):
        gi = int(rng.integers(0, len(cs)))
        r = int(rng.integers(0, cs[gi].n))
        n = cs[gi].n
        bufs = (tb.host_array(n, np.int32), tb.host_array(n, np.uint64), tb.host_array(n, np.uint8))
        store[srv.submit(gi, r, *bufs)] = (gi, r) + bufs
    done = []
    while len(done) < len(store):
        done += srv.poll(4096, wait_us=100000)
    for d in done:
        assert d["status"] == 0, d
        gi, r, S, sig, ovf = store[d["ticket"]]
        eS, esig, eovf, _ = oracle.textbook(cs[gi].CP, cs[gi].IC, r)
        np.testing.assert_array_equal(S, eS, err_msg=f"graph {gi} root {r}")
        np.testing.assert_array_equal(sig, esig, err_msg=f"graph {gi} root {r}")
        np.testing.assert_array_equal(ovf, eovf, err_msg=f"graph {gi} root {r}")
    srv.close()

