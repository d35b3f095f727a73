"""The chained loop's first level run by its init kernel.

For one root in AUTO or PUSH mode the init kernel pushes from the root itself
when the root has at most 256 out-list entries and the direction rule picks a
push (otherwise the first level launch does it).  These graphs put the cases
side by side, each checked bit for bit against the oracle (S, sigma mod 2^64,
overflow flags, level count): roots with 256 and 257 out-list entries,
repeated edges and a self-loop at the root (counts add per edge, as the
reference's sum over in-list entries does), a directed root no edge enters,
an isolated root, and a root whose first level the rule makes a pull.  Runs go
through tbfs_bfs, tbfs_bfs_many (back to back on one context) and the
graph-captured loop (which seeds its root from device memory and never fuses).
"""
import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import _lib as L, graphs

pytestmark = pytest.mark.gpu


def _csc(n, src, dst):
    """CSC with every listed edge src -> dst kept (repeats and self-loops
    included), in-lists in ascending row order."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.lexsort((src, dst))
    col, row = dst[order], src[order]
    CP = np.zeros(n + 1, np.int64)
    np.add.at(CP, col + 1, 1)
    return graphs.CSC(np.cumsum(CP).astype(np.int32), row.astype(np.int32))


def _graph(seed=11):
    """Symmetric multigraph on 4000 vertices: a random background (average
    degree ~6) plus vertex 0 joined to 256 others, vertex 1 to 257, vertex 2
    joined to vertex 5 three times and to itself twice, vertex 3 isolated."""
    rng = np.random.default_rng(seed)
    n = 4000
    m = 12000
    a = rng.integers(4, n, m)
    b = rng.integers(4, n, m)
    keep = a != b
    src = list(a[keep])
    dst = list(b[keep])
    hub0 = rng.choice(np.arange(4, n), 256, replace=False)
    hub1 = rng.choice(np.arange(4, n), 257, replace=False)
    src += [0] * 256 + [1] * 257 + [2, 2, 2] + [2, 2]
    dst += list(hub0) + list(hub1) + [5, 5, 5] + [2, 2]
    src += [2] * 4
    dst += list(rng.integers(4, n, 4))
    s = np.asarray(src, np.int64)
    d = np.asarray(dst, np.int64)
    # both directions of every edge (a self-loop once per listing)
    loops = s == d
    src2 = np.concatenate([s, d[~loops]])
    dst2 = np.concatenate([d, s[~loops]])
    return _csc(n, src2, dst2)


def _directed(seed=12):
    """Directed graph: vertex 0 has 200 out-edges and no in-edge; a random
    background of 8000 edges; a few repeated edges out of vertex 0."""
    rng = np.random.default_rng(seed)
    n = 3000
    a = rng.integers(1, n, 8000)
    b = rng.integers(1, n, 8000)
    keep = a != b
    outs = rng.choice(np.arange(1, n), 200, replace=False)
    src = np.concatenate([a[keep], np.zeros(200, np.int64), [0, 0]])
    dst = np.concatenate([b[keep], outs, outs[:2]])
    return _csc(n, src, dst)


def _same(res, exp, what):
    S, sig, ovf, lv = exp
    np.testing.assert_array_equal(res.S, S, err_msg=f"{what}: S")
    np.testing.assert_array_equal(res.sigma, sig, err_msg=f"{what}: sigma")
    np.testing.assert_array_equal(res.ovf, ovf, err_msg=f"{what}: ovf")
    assert res.stats["levels"] == lv, what


@pytest.mark.parametrize("kind", ["symmetric", "directed"])
def test_init_runs_depth0(gpu, kind):
    g = _graph() if kind == "symmetric" else _directed()
    symmetric = kind == "symmetric"
    gr = tb.Graph(g.CP, g.IC, symmetric=symmetric)
    ctx = tb.Context(g.n)
    outdeg = np.bincount(g.IC, minlength=g.n)
    roots = [0, 1, 2, 3, 7] if symmetric else [0, 5, 9]
    if symmetric:
        assert outdeg[0] == 256 and outdeg[1] == 257 and outdeg[3] == 0
    cases = [("chained", dict(loop=L.LOOP_CHAINED)),
             ("push", dict(mode=L.MODE_PUSH, loop=L.LOOP_CHAINED)),
             ("graph", dict(loop=L.LOOP_GRAPH)),
             ("pull first", dict(loop=L.LOOP_CHAINED, alpha=1e9))]
    for r in roots:
        exp = oracle.textbook(g.CP, g.IC, r)
        for name, kw in cases:
            print(f"[init depth 0] {kind} root={r} ({outdeg[r]} out) {name}", flush=True)
            o = tb.default_opts(**({"mode": L.MODE_AUTO, "pull": L.PULL_LB} | kw))
            _same(ctx.bfs(gr, r, o), exp, f"{kind} root {r} {name}")
    # back to back, the results of each checked through its counts and the
    # last one's arrays
    rr = np.asarray(roots * 3, np.int32)
    stats = ctx.run_many(gr, rr)
    deg_in = np.diff(g.CP).astype(np.int64)
    for r, st in zip(rr, stats):
        S, sig, ovf, lv = oracle.textbook(g.CP, g.IC, int(r))
        assert st["levels"] == lv and st["reached"] == int((S >= 0).sum())
        assert st["edges"] == int(deg_in[S >= 0].sum())
    S, sig, ovf = ctx.read(g.n)
    exp = oracle.textbook(g.CP, g.IC, int(rr[-1]))
    np.testing.assert_array_equal(S, exp[0])
    np.testing.assert_array_equal(sig, exp[1])
    ctx.close()
