"""The single-source autotuner (tbfs_autotune): every candidate gives the
oracle's results, the winner is recorded for the graph's class and taken by
runs with default options, and the table survives a JSON round trip."""
import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import _lib as L, graphs

pytestmark = pytest.mark.gpu


def _same(res, exp, what):
    S, sig, ovf, lv = exp
    np.testing.assert_array_equal(res.S, S, err_msg=f"{what}: S")
    np.testing.assert_array_equal(res.sigma, sig, err_msg=f"{what}: sigma")
    np.testing.assert_array_equal(res.ovf, ovf, err_msg=f"{what}: ovf")
    assert res.stats["levels"] == lv, what


EXTREMES = [dict(alpha=6.0, beta=96.0, stage_ratio=0.25, lb_grid=256),
            dict(alpha=30.0, beta=8.0, stage_ratio=4.0, lb_grid=2048),
            dict(alpha=10.0, beta=16.0, stage_ratio=0.5, lb_grid=64)]


def test_autotune_single_source(gpu, tmp_path):
    tb.tune_clear()
    g = graphs.rmat(16, 16, 0.57, 0.19, 0.19, seed=9, relabel=True)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 8, seed=9)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    best, ms, dms = ctx.autotune(gr, roots)
    print(f"[tune] best alpha={best.alpha} beta={best.beta} stage={best.stage_ratio} "
          f"lb_grid={best.lb_grid} ms={ms:.4f} default={dms:.4f}", flush=True)
    assert 0 < ms <= dms
    assert tb.tune_count() == 1
    t = tb.tuned(gr)
    assert t is not None
    assert (t.alpha, t.beta, t.stage_ratio, t.lb_grid) == \
        (best.alpha, best.beta, best.stage_ratio, best.lb_grid)
    exp = {int(r): oracle.textbook(g.CP, g.IC, int(r)) for r in roots[:4]}
    for r in roots[:4]:  # default options: the tuned values apply
        _same(ctx.bfs(gr, int(r)), exp[int(r)], f"tuned root {r}")
    for kw in EXTREMES:  # candidates far from the defaults
        for r in roots[:2]:
            _same(ctx.bfs(gr, int(r), tb.default_opts(**kw)), exp[int(r)], f"{kw} root {r}")
    path = tmp_path / "tune.json"
    tb.tune_save(str(path))
    tb.tune_clear()
    assert tb.tuned(gr) is None
    tb.tune_load(str(path))
    t2 = tb.tuned(gr)
    assert (t2.alpha, t2.beta, t2.stage_ratio, t2.lb_grid) == \
        (best.alpha, best.beta, best.stage_ratio, best.lb_grid)
    tb.tune_clear()
    ctx.close()
    gr.close()


def test_tunables_high_diameter(gpu):
    """The same candidates on a perturbed grid (hundreds of levels, counts
    past 2^64): results are the oracle's under each."""
    tb.tune_clear()
    g = graphs.grid(96, 96, 0.1, seed=4, relabel=True)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    r = int(graphs.largest_component_vertices(g)[0])
    exp = oracle.textbook(g.CP, g.IC, r)
    for kw in EXTREMES:
        for loop in (L.LOOP_CHAINED, L.LOOP_GRAPH):
            _same(ctx.bfs(gr, r, tb.default_opts(loop=loop, **kw)), exp, f"{kw} loop {loop}")
    ctx.close()
    gr.close()
