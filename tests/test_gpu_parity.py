"""GPU parity: every entry point through the C-ABI against the oracle.

S must match bit-exactly; sigma must match exactly (uint64 value and overflow
flag through the extended API; at the drop-in boundary, the reference's
float/int output type with the inexact flag raised exactly when a count
leaves that type's exact range).
"""
import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import graphs, mtx
from tests.conftest import golden

pytestmark = pytest.mark.gpu

F32_EXACT = 1 << 24
I32_MAX = 2**31 - 1


def _er4096():
    return graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)


def _dropin_all(g, r, directed_ok=True):
    """Run the seven drop-in entry points; return dict name -> outputs."""
    n, nz = g.n, g.nz
    out = {}
    for name in ("bfs_gpu_td_csc_sc", "bfs_gpu_td_csc_wa"):
        S = np.full(n, 7, np.int32)
        sig = np.full(n, -1, np.float32)
        rc = getattr(tb, name)(g.IC, g.CP, S, sig, r, nz, n, 1)
        out[name] = (rc, S, sig)
    for name in ("bfs_gpu_bu_csc_sc", "bfs_gpu_bu_csc_wa"):
        for bb in (0, 1):
            S = np.full(n, 7, np.int32)
            rc = getattr(tb, name)(g.IC, g.CP, S, r, nz, n, 1, bb)
            out[f"{name}[bbreak={bb}]"] = (rc, S, None)
    for name in ("bfs_gpu_tdbu_csc_sc", "bfs_gpu_tdbu_csc_wa"):
        S = np.full(n, 7, np.int32)
        sig = np.full(n, 7, np.int32)
        rc = getattr(tb, name)(g.IC, g.CP, S, sig, r, nz, n, 1, 1)
        out[name] = (rc, S, sig)
    I, J = graphs.csc_to_cooc(g)
    sig = np.full(n, -1, np.float32)
    rc = tb.bfs_gpu_td_cooc_sc(I, J, sig, r, nz, n, 1)
    out["bfs_gpu_td_cooc_sc"] = (rc, None, sig)
    return out


MAPPINGS = (tb._lib.DROPIN_FAST, tb._lib.DROPIN_REFERENCE)


def _dropin_with(mapping, g, r):
    """_dropin_all under one kernel mapping of the entry points."""
    prev = tb.dropin_mapping(mapping)
    try:
        return _dropin_all(g, r)
    finally:
        tb.dropin_mapping(prev)


def _check_dropin(g, r, mappings=MAPPINGS):
    """Every entry point under each kernel mapping against the oracle: the
    reference's output conventions and types, TBFS_SIGMA_INEXACT (5) exactly
    when oracle.count_exceed finds a count past the type's exact range."""
    S0, sig0, ovf0, _ = oracle.seq_td_csc(g.CP, g.IC, r, s_unreached=0)
    S1, sig1, ovf1, _ = oracle.seq_td_csc(g.CP, g.IC, r, s_unreached=-1)
    exp_f32 = np.where(ovf1 == 1, np.float32(np.inf), sig1.astype(np.float32))
    f32_inexact = oracle.count_exceed(sig1, ovf1, F32_EXACT) > 0
    i32_inexact = oracle.count_exceed(sig1, ovf1, I32_MAX) > 0
    exp_i32 = (sig1 & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
    for m in mappings:
        _check_outputs(_dropin_with(m, g, r), S0, S1, exp_f32, exp_i32, f32_inexact,
                       i32_inexact, m)


def _check_outputs(res, S0, S1, exp_f32, exp_i32, f32_inexact, i32_inexact, m):
    for name, (rc, S, sig) in res.items():
        name = f"{name} (mapping {m})"
        if name.startswith("bfs_gpu_td_csc"):
            assert rc == (5 if f32_inexact else 0), name
            np.testing.assert_array_equal(S, S0, err_msg=name)
            np.testing.assert_array_equal(sig, exp_f32, err_msg=name)
        elif name.startswith("bfs_gpu_bu_csc"):
            assert rc == 0, name
            np.testing.assert_array_equal(S, S1, err_msg=name)
        elif name.startswith("bfs_gpu_tdbu_csc"):
            assert rc == (5 if i32_inexact else 0), name
            np.testing.assert_array_equal(S, S1, err_msg=name)
            np.testing.assert_array_equal(sig, exp_i32, err_msg=name)
        else:  # cooc
            assert rc == (5 if f32_inexact else 0), name
            np.testing.assert_array_equal(sig, exp_f32, err_msg=name)


def test_er4096_dropin(gpu):
    """Config: undirected G(4096, 8/4095) seed 1, source 0, all seven entry points."""
    _check_dropin(_er4096(), 0)


@pytest.mark.parametrize("fname,r", [
    ("mycielskian4.mtx", 0), ("Tina_AskCal.mtx", 10), ("cage4.mtx", 0),
    ("Ragusa18.mtx", 0), ("luxembourg_osm.mtx.gz", 0),
])
def test_reference_fixtures_dropin(gpu, fname, r):
    g = mtx.read_csc(golden(fname))
    _check_dropin(g, r)


L = tb._lib
HOST = L.LOOP_HOST
PERS = L.LOOP_PERSISTENT
MODES = [
    ("auto_sc", dict(mode=L.MODE_AUTO, pull=L.PULL_SC, loop=HOST)),
    ("auto_wa", dict(mode=L.MODE_AUTO, pull=L.PULL_WA, loop=HOST)),
    ("auto_lb", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=HOST)),
    ("pull_sc", dict(mode=L.MODE_PULL_SC, loop=HOST)),
    ("pull_wa", dict(mode=L.MODE_PULL_WA, loop=HOST)),
    ("push", dict(mode=L.MODE_PUSH, pull=L.PULL_SC, loop=HOST)),
    ("push_lb", dict(mode=L.MODE_PUSH, pull=L.PULL_LB, loop=HOST)),
    ("cooc", dict(mode=L.MODE_COOC, loop=HOST)),
    ("pers_auto", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=PERS)),
    ("pers_push", dict(mode=L.MODE_PUSH, pull=L.PULL_LB, loop=PERS)),
    ("pers_pull", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=PERS, alpha=1e9)),
    ("pers_auto_g64", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=PERS, grid=64)),
    ("pers_auto_b1024", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=PERS, block=1024)),
    ("chain_auto", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED)),
    ("chain_auto_no_lds", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED,
                               fb_lds=-1)),
    ("chain_push", dict(mode=L.MODE_PUSH, pull=L.PULL_LB, loop=L.LOOP_CHAINED)),
    ("chain_pull_depth_only", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED,
                                   alpha=1e9, sigma=0)),
    ("chain_auto_no_finisher", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED,
                                    finish=-1)),
    ("chain_auto_finisher_early", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED,
                                       finish=65536)),
    ("chain_depth_only_finisher", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED,
                                       sigma=0, finish=65536)),
    ("graph_auto", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_GRAPH)),
    ("graph_auto_finisher_early", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_GRAPH,
                                       finish=65536)),
    ("graph_push", dict(mode=L.MODE_PUSH, pull=L.PULL_LB, loop=L.LOOP_GRAPH)),
]


def _check_modes(g, roots, symmetric, modes=MODES):
    gr = tb.Graph(g.CP, g.IC, symmetric=symmetric)
    ctx = tb.Context(g.n)
    for r in roots:
        S, sig, ovf, lv = oracle.textbook(g.CP, g.IC, int(r))
        for name, kw in modes:
            res = ctx.bfs(gr, int(r), tb.default_opts(**kw))
            np.testing.assert_array_equal(res.S, S, err_msg=f"{name} r={r} S")
            if kw.get("sigma", 1) == 0:
                assert res.stats["levels"] == lv, name
                continue
            np.testing.assert_array_equal(res.sigma, sig, err_msg=f"{name} r={r} sigma")
            np.testing.assert_array_equal(res.ovf, ovf, err_msg=f"{name} r={r} ovf")
            assert res.stats["levels"] == lv, name
            assert res.stats["reached"] == int((S >= 0).sum()), name
            assert res.stats["edges"] == int(np.diff(g.CP)[S >= 0].sum()), name


def test_rmat14_all_modes(gpu):
    g = graphs.rmat(14, 16, seed=2)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 4, seed=2)
    _check_modes(g, roots, symmetric=True)
    _check_modes(g, roots[:1], symmetric=False)


def test_grid_overflow_all_modes(gpu):
    """Unrelabelled 48x48 grid from a corner: sigma = C(x+y, x) passes 2^64
    (x+y >= 68), so the overflow flag and the wrapped value are exercised."""
    g = graphs.grid(48, 48, 0.0, seed=0, relabel=False)
    S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, 0)
    assert ovf.any() and not ovf.all()
    _check_modes(g, [0, 48 * 48 - 1], symmetric=True)


def test_perturbed_grid_modes(gpu):
    g = graphs.grid(96, 64, 0.1, seed=3, relabel=True)
    _check_modes(g, [0, 17, 5000], symmetric=True)


def test_hub_columns_split_across_slices(gpu):
    """Star + clique mix: columns with in-degree > 256 take the static
    pull-slice path and out-lists > 256 the sliced push path."""
    hub_leaves = 5000
    n = hub_leaves + 1 + 40
    src, dst = [], []
    for v in range(1, hub_leaves + 1):          # star around vertex 0
        src += [0, v]
        dst += [v, 0]
    base = hub_leaves + 1
    for a in range(40):                          # clique joined to two leaves
        for b in range(40):
            if a != b:
                src.append(base + a)
                dst.append(base + b)
    for a in range(40):
        for leaf in (7, 11):
            src += [base + a, leaf]
            dst += [leaf, base + a]
    src = np.array(src, np.int32)
    dst = np.array(dst, np.int32)
    order = np.lexsort((src, dst))
    IC = src[order]
    CP = np.zeros(n + 1, np.int64)
    CP[1:] = np.cumsum(np.bincount(dst, minlength=n))
    g = graphs.CSC(CP.astype(np.int32), IC)
    _check_modes(g, [0, 3, base + 5], symmetric=True)
    _check_modes(g, [base], symmetric=False)


def _sym_csc(n, src, dst):
    src = np.asarray(src, np.int32)
    dst = np.asarray(dst, np.int32)
    s2 = np.concatenate([src, dst])
    d2 = np.concatenate([dst, src])
    key = np.unique(d2.astype(np.int64) * n + s2)
    dst_u, src_u = (key // n).astype(np.int32), (key % n).astype(np.int32)
    CP = np.zeros(n + 1, np.int64)
    CP[1:] = np.cumsum(np.bincount(dst_u, minlength=n))
    return graphs.CSC(CP.astype(np.int32), src_u)


def test_slice_boundaries(gpu):
    """Lists of 255..1025 entries around the 256-entry slice size: the root
    reaches 3 vertices, each joined to all of layer A (1100 vertices, count
    3); hub i joins the first d_i - 5 vertices of A (count 3 (d_i - 5)) and
    5 leaves of its own, so its list holds d_i entries.  Pull levels take the
    hubs through 1..5 slices, push levels take the lists in slices."""
    degs = [255, 256, 257, 300, 511, 512, 513, 768, 1025]
    na = 1100
    n_pre, a0 = 3, 4
    h0 = a0 + na
    l0 = h0 + len(degs)
    nleaf = 5
    n = l0 + len(degs) * nleaf
    src, dst = [], []
    for p in range(1, 1 + n_pre):
        src.append(0)
        dst.append(p)
        src += [p] * na
        dst += list(range(a0, a0 + na))
    for i, d in enumerate(degs):
        src += [h0 + i] * (d - nleaf)
        dst += list(range(a0, a0 + d - nleaf))
        for k in range(nleaf):
            src.append(h0 + i)
            dst.append(l0 + i * nleaf + k)
    g = _sym_csc(n, src, dst)
    assert list(np.diff(g.CP)[h0:l0]) == degs
    pull_modes = MODES + [
        ("chain_pull", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED, alpha=1e9)),
        ("chain_pull_no_lds", dict(mode=L.MODE_AUTO, pull=L.PULL_LB, loop=L.LOOP_CHAINED,
                                   alpha=1e9, fb_lds=-1)),
    ]
    _check_modes(g, [0, a0 + 3, h0 + 8, l0 + 2], symmetric=True, modes=pull_modes)


@pytest.mark.parametrize("loop", [L.LOOP_PERSISTENT, L.LOOP_CHAINED, L.LOOP_HOST, L.LOOP_GRAPH])
def test_run_many_matches_single(gpu, loop):
    g = graphs.rmat(12, 16, seed=7)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 6, seed=1)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    stats = ctx.run_many(gr, roots, tb.default_opts(loop=loop))
    S, sig, ovf = ctx.read(g.n)
    S2, sig2, ovf2, lv = oracle.textbook(g.CP, g.IC, int(roots[-1]))
    np.testing.assert_array_equal(S, S2)
    np.testing.assert_array_equal(sig, sig2)
    for r, st in zip(roots, stats):
        Sr, _, _, lvr = oracle.textbook(g.CP, g.IC, int(r))
        assert st["levels"] == lvr and st["reached"] == int((Sr >= 0).sum())
        assert st["edges"] == int(np.diff(g.CP)[Sr >= 0].sum())
        assert st["ms"] > 0


def test_directed_reference_fixture_modes(gpu):
    g = mtx.read_csc(golden("Tina_AskCal.mtx"))
    _check_modes(g, range(g.n), symmetric=False)


def test_edge_cases(gpu):
    # single vertex, no edges
    g = graphs.CSC(np.array([0, 0], np.int32), np.zeros(0, np.int32))
    _check_dropin(g, 0)
    # isolated root in a larger graph, plus a duplicate entry and a self-loop
    CP = np.array([0, 0, 3, 5, 6], np.int32)
    IC = np.array([2, 2, 1, 1, 3, 2], np.int32)  # col1: 2,2,1 ; col2: 1,3 ; col3: 2
    g = graphs.CSC(CP, IC)
    for r in range(4):
        _check_dropin(g, r)
        _check_modes(g, [r], symmetric=False)


def test_invalid_inputs_rejected(gpu):
    g = _er4096()
    S = np.zeros(g.n, np.int32)
    sig = np.zeros(g.n, np.float32)
    assert tb.bfs_gpu_td_csc_sc(g.IC, g.CP, S, sig, g.n, g.nz, g.n, 1) == 1
    assert tb.bfs_gpu_td_csc_sc(g.IC, g.CP, S, sig, -1, g.nz, g.n, 1) == 1
    assert tb.bfs_gpu_td_csc_sc(g.IC, g.CP, S, sig, 0, g.nz, g.n, 0) == 1
    bad = g.CP.copy()
    bad[5] = bad[6] + 1  # not monotone
    assert tb.bfs_gpu_td_csc_sc(g.IC, bad, S, sig, 0, g.nz, g.n, 1) == 1
    badic = g.IC.copy()
    badic[3] = g.n
    assert tb.bfs_gpu_bu_csc_sc(badic, g.CP, S, 0, g.nz, g.n, 1, 1) == 1


def test_launch_budget_follows_the_graph(gpu):
    """Chained loop: a BFS uses one level launch per depth plus the one that
    finds the frontier empty; the next call's budget is the most launches
    the previous call needed plus one.  A deep graph (a 300 x 4 grid, 303
    levels) after a shallow one (R-MAT 12) runs past the shrunken budget and
    is finished by the host-checked re-run; a shallow graph after the deep
    one runs with the deep budget's spare launches.  Every result stays
    exact."""
    deep = graphs.grid(300, 4, 0.0, seed=0, relabel=False)
    shallow = graphs.rmat(12, 16, seed=7)
    gd = tb.Graph(deep.CP, deep.IC, symmetric=True)
    gs = tb.Graph(shallow.CP, shallow.IC, symmetric=True)
    ctx = tb.Context(max(deep.n, shallow.n))
    every_level = tb.default_opts(finish=-1)  # one launch per level throughout
    for g, gr, roots in ((shallow, gs, [0, 1, 2]), (deep, gd, [0, 5]), (shallow, gs, [3, 4]),
                         (deep, gd, [7, 0])):
        st = ctx.run_many(gr, roots, every_level)
        for r, x in zip(roots, st):
            S, sig, ovf, lv = oracle.textbook(g.CP, g.IC, int(r))
            assert x["levels"] == lv and x["launches"] == lv + 1, (r, x)
        S, sig, ovf, lv = oracle.textbook(g.CP, g.IC, int(roots[-1]))
        got = ctx.read(g.n)  # the last source's results stay in the context
        np.testing.assert_array_equal(got[0], S)
        np.testing.assert_array_equal(got[1], sig)
        np.testing.assert_array_equal(got[2], ovf)
    ctx.close()


def test_repetition_same_result(gpu):
    g = _er4096()
    S1 = np.zeros(g.n, np.int32)
    s1 = np.zeros(g.n, np.float32)
    S3 = np.zeros(g.n, np.int32)
    s3 = np.zeros(g.n, np.float32)
    assert tb.bfs_gpu_td_csc_wa(g.IC, g.CP, S1, s1, 5, g.nz, g.n, 1) == 0
    assert tb.bfs_gpu_td_csc_wa(g.IC, g.CP, S3, s3, 5, g.nz, g.n, 3) == 0
    np.testing.assert_array_equal(S1, S3)
    np.testing.assert_array_equal(s1, s3)


@pytest.mark.parametrize("fixture", ["er4096_seed1", "rmat12_seed2", "luxembourg_osm_r0"])
def test_dropin_equals_reference_golden_vectors(gpu, fixture):
    """Golden vectors written by the reference's own sequential C code
    (tests/golden/make_fixtures.py): every drop-in entry point reproduces them
    exactly (all counts here are below 2^24, so float is exact)."""
    fx = np.load(golden(fixture + ".npz"))
    if fixture == "er4096_seed1":
        g = graphs.CSC(fx["CP"], fx["IC"])
    elif fixture == "rmat12_seed2":
        g = graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    else:
        g = mtx.read_csc(golden("luxembourg_osm.mtx.gz"))
    for i, r in enumerate(fx["roots"]):
        for m in MAPPINGS:
            _check_golden(_dropin_with(m, g, int(r)), fx, i, m)


def _check_golden(res, fx, i, m):
    for name, (rc, S, sig) in res.items():
        name = f"{name} (mapping {m})"
        assert rc == 0, name
        if name.startswith("bfs_gpu_td_csc"):
            np.testing.assert_array_equal(S, fx["S_td"][i], err_msg=name)
            np.testing.assert_array_equal(sig, fx["sigma_td"][i], err_msg=name)
        elif name.startswith("bfs_gpu_bu_csc"):
            np.testing.assert_array_equal(S, fx["S_bu"][i], err_msg=name)
        elif name.startswith("bfs_gpu_tdbu_csc"):
            np.testing.assert_array_equal(S, fx["S_bu"][i], err_msg=name)
            np.testing.assert_array_equal(sig, fx["sigma_tdbu"][i], err_msg=name)
        else:
            np.testing.assert_array_equal(sig, fx["sigma_cooc"][i], err_msg=name)


def test_finisher_takes_the_tail(gpu):
    """Once at most `finish` in-list entries are unexplored, one workgroup
    runs the remaining levels: with finish = 65536 a BFS on R-MAT 14 (nz
    ~ 460k) ends within fewer launches than one per level (the default, off);
    results and level / vertex / edge counts are the same."""
    g = graphs.rmat(14, 16, seed=2)
    gr = tb.Graph(g.CP, g.IC, symmetric=True)
    ctx = tb.Context(g.n)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 6, seed=5)
    for r in roots:
        S, sig, ovf, lv = oracle.textbook(g.CP, g.IC, int(r))
        runs = {}
        for name, fin in (("off", -1), ("default", 0), ("late", 1024), ("early", 65536)):
            res = ctx.bfs(gr, int(r), tb.default_opts(finish=fin))
            np.testing.assert_array_equal(res.S, S, err_msg=f"{name} r={r}")
            np.testing.assert_array_equal(res.sigma, sig, err_msg=f"{name} r={r}")
            np.testing.assert_array_equal(res.ovf, ovf, err_msg=f"{name} r={r}")
            assert res.stats["levels"] == lv and res.stats["reached"] == int((S >= 0).sum())
            assert res.stats["edges"] == int(np.diff(g.CP)[S >= 0].sum())
            runs[name] = res.stats["launches"]
        assert runs["off"] == lv + 1 and runs["default"] == lv + 1, runs
        assert runs["early"] < runs["off"], runs
    ctx.close()


def test_dropin_calls_reuse_and_release(gpu):
    """Back-to-back calls of an entry point on graphs of different sizes (the
    reused contexts and buffers must not carry anything from one call to the
    next), then a release of what the calls keep, then calls again."""
    small = _er4096()
    big = graphs.rmat(14, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    for rep in range(2):
        for g, r in ((big, 3), (small, 0), (big, 11), (small, 7)):
            print(f"[dropin reuse] rep={rep} n={g.n} root={r}", flush=True)
            _check_dropin(g, r, mappings=(tb._lib.DROPIN_FAST,))
        tb.dropin_release()
