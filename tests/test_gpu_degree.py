"""Degree statistics on the device (tbfs_graph_degree_stats) and the batch
autotuner's table keyed by them (tbfs_batch_tuned / tbfs_batch_tune_clear).

The statistics restate the reference's degree.c (TurboBFS_CSC_TD/degree.c:
degree_max keeps the first vertex with the largest degree, degree_sta
divides the squared deviations by n - 1, degree_dist_ac groups degrees into
0, 1, [2, 4), [4, 8), ...); they are checked against numpy on the same
arrays and against closed forms (star, regular graph, one vertex).
"""
import math

import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import graphs, mtx
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _expect(deg):
    deg = np.asarray(deg, np.int64)
    n = deg.shape[0]
    bins = np.zeros(33, np.int64)
    for d, c in zip(*np.unique(deg, return_counts=True)):
        bins[0 if d == 0 else int(d).bit_length()] += c
    nb = int(np.nonzero(bins)[0].max()) + 1
    return {"max": int(deg.max()), "argmax": int(np.argmax(deg)), "sum": int(deg.sum()),
            "mean": float(deg.sum()) / n, "sd": float(np.std(deg, ddof=1)) if n > 1 else 0.0,
            "bins": [int(x) for x in bins[:nb]]}


def _check(got, exp):
    assert got["max"] == exp["max"]
    assert got["argmax"] == exp["argmax"]
    assert got["sum"] == exp["sum"]
    assert got["bins"] == exp["bins"]
    assert math.isclose(got["mean"], exp["mean"], rel_tol=1e-12, abs_tol=1e-12)
    assert math.isclose(got["sd"], exp["sd"], rel_tol=1e-9, abs_tol=1e-12)


@pytest.mark.parametrize("name", ["rmat14", "grid", "er4096"])
def test_degree_stats_symmetric(gpu, name):
    g = {"rmat14": lambda: graphs.rmat(14, 16, 0.57, 0.19, 0.19, seed=3, relabel=True),
         "grid": lambda: graphs.grid(64, 48, 0.1, seed=3, relabel=True),
         "er4096": lambda: graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)}[name]()
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    exp = _expect(np.diff(g.CP))
    _check(gr.degree_stats("in"), exp)
    _check(gr.degree_stats("out"), exp)


def test_degree_stats_directed_in_and_out(gpu):
    g = mtx.read_csc(golden("Tina_AskCal.mtx"))
    gr = tb.Graph(g.CP, g.IC, symmetric=False)
    _check(gr.degree_stats("in"), _expect(np.diff(g.CP)))
    _check(gr.degree_stats("out"), _expect(np.bincount(g.IC, minlength=g.n)))


def test_degree_stats_closed_forms(gpu):
    # star: centre n-1, leaves 1
    n = 1000
    src = np.zeros(n - 1, np.int64)
    dst = np.arange(1, n)
    key = np.unique(np.concatenate([dst * n + src, src * n + dst]))
    col, row = key // n, key % n
    CP = np.zeros(n + 1, np.int64)
    CP[1:] = np.cumsum(np.bincount(col, minlength=n))
    st = tb.Graph(CP.astype(np.int32), row.astype(np.int32), symmetric=True).degree_stats()
    m = 2.0 * (n - 1) / n
    sd = math.sqrt(((n - 1 - m) ** 2 + (n - 1) * (1 - m) ** 2) / (n - 1))
    assert (st["max"], st["argmax"], st["sum"]) == (n - 1, 0, 2 * (n - 1))
    assert math.isclose(st["mean"], m, rel_tol=1e-12) and math.isclose(st["sd"], sd, rel_tol=1e-9)
    assert st["bins"][1] == n - 1 and sum(st["bins"]) == n
    assert st["bins"][(n - 1).bit_length()] == 1
    # cycle: every degree 2, sd 0, first vertex holds the maximum
    n2 = 64
    s2 = np.arange(n2)
    d2 = (s2 + 1) % n2
    key = np.unique(np.concatenate([d2 * n2 + s2, s2 * n2 + d2]))
    CP2 = np.zeros(n2 + 1, np.int64)
    CP2[1:] = np.cumsum(np.bincount(key // n2, minlength=n2))
    st2 = tb.Graph(CP2.astype(np.int32), (key % n2).astype(np.int32), symmetric=True).degree_stats()
    assert (st2["max"], st2["argmax"], st2["mean"], st2["sd"]) == (2, 0, 2.0, 0.0)
    assert st2["bins"] == [0, 0, n2]
    # one isolated vertex: sd defined as 0
    st3 = tb.Graph(np.zeros(2, np.int32), np.zeros(0, np.int32), symmetric=True).degree_stats()
    assert (st3["max"], st3["argmax"], st3["sum"], st3["sd"], st3["bins"]) == (0, 0, 0, 0.0, [1])


def test_degree_stats_invalid(gpu):
    import ctypes as C
    g = graphs.erdos_renyi(256, 0.05, seed=2)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    with pytest.raises(ValueError):
        gr.degree_stats("both")
    ds = tb._lib.DegreeStats()
    lib = tb.load()
    assert lib.tbfs_graph_degree_stats(gr.handle, 7, C.byref(ds)) == tb._lib.TBFS_ERR_INVALID
    assert lib.tbfs_graph_degree_stats(gr.handle, 0, None) == tb._lib.TBFS_ERR_INVALID
    assert lib.tbfs_graph_degree_stats(None, 0, C.byref(ds)) == tb._lib.TBFS_ERR_INVALID


def test_tuning_table_keyed_by_degrees(gpu):
    tb.batch_tune_clear()
    g = graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=5, relabel=True)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 128, seed=5)
    assert tb.batch_tuned(gr, 128) is None
    b = tb.Batch(g.n, 128)
    b.run(gr, roots)
    S0, sig0, ovf0 = b.read()
    best, ms = b.autotune(gr, roots)
    t = tb.batch_tuned(gr, 128)
    assert t is not None and (t.grid, t.batch_alpha) == (best.grid, best.batch_alpha)
    assert tb.batch_tuned(gr, 64) is None  # one word instead of two: another key
    # default options now take the tuned values; every result stays exact
    b.run(gr, roots)
    S1, sig1, ovf1 = b.read()
    np.testing.assert_array_equal(S0, S1)
    np.testing.assert_array_equal(sig0, sig1)
    np.testing.assert_array_equal(ovf0, ovf1)
    for i in (0, 77, 127):
        eS, esig, eovf, _ = oracle.textbook(g.CP, g.IC, int(roots[i]))
        np.testing.assert_array_equal(S1[i], eS)
        np.testing.assert_array_equal(sig1[i], esig)
    tb.batch_tune_clear()
    assert tb.batch_tuned(gr, 128) is None
    b.close()
