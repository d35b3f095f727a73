"""Pinning the oracle (CPU only).

The oracle restates the reference's sequential BFS.  It is checked here
against things that do not share its formula: closed forms, brute-force path
enumeration on tiny graphs, an independent queue BFS, the known answers for
the reference's bundled Matrix Market files, and the reference's own C
sequential code compiled from its sources into oracle/_ref (float / int path
counts, so compared where those are exact).
"""
import itertools
import math

import numpy as np
import pytest

import oracle
from turbobfs_amd import graphs, mtx
from tests.conftest import golden

U64 = 1 << 64


def csc_from_edges(n, edges, undirected=True):
    src, dst = [], []
    for u, v in edges:
        src.append(u)
        dst.append(v)
        if undirected:
            src.append(v)
            dst.append(u)
    src = np.array(src, np.int64)
    dst = np.array(dst, np.int64)
    order = np.lexsort((src, dst))
    CP = np.zeros(n + 1, np.int64)
    CP[1:] = np.cumsum(np.bincount(dst, minlength=n))
    return graphs.CSC(CP.astype(np.int32), src[order].astype(np.int32))


def expected_pair(big):
    """Exact integer -> (value mod 2^64, overflow flag)."""
    return big % U64, int(big >= U64)


def all_variants(g, r):
    """(S, sigma, ovf) from the three oracle functions; S with -1 unreached."""
    S1, s1, o1, _ = oracle.seq_td_csc(g.CP, g.IC, r, s_unreached=-1)
    S2, s2, o2, _ = oracle.textbook(g.CP, g.IC, r)
    I, J = graphs.csc_to_cooc(g)
    s3, o3, _ = oracle.seq_td_cooc(I, J, g.n, r)
    np.testing.assert_array_equal(S1, S2)
    np.testing.assert_array_equal(s1, s2)
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(s1, s3)
    np.testing.assert_array_equal(o1, o3)
    return S1, s1, o1


# ------------------------------------------------------------ closed forms ---
def test_path_graph():
    n = 9
    g = csc_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    S, sig, ovf = all_variants(g, 3)
    np.testing.assert_array_equal(S, np.abs(np.arange(n) - 3))
    assert (sig == 1).all() and not ovf.any()


def test_even_cycle_antipode_has_two_paths():
    n = 10
    g = csc_from_edges(n, [(i, (i + 1) % n) for i in range(n)])
    S, sig, _ = all_variants(g, 0)
    np.testing.assert_array_equal(S, [min(i, n - i) for i in range(n)])
    assert sig[5] == 2 and (np.delete(sig, 5) == 1).all()


def test_complete_and_star():
    g = csc_from_edges(6, [(i, j) for i in range(6) for j in range(i + 1, 6)])
    S, sig, _ = all_variants(g, 2)
    assert S[2] == 0 and (np.delete(S, 2) == 1).all() and (sig == 1).all()
    g = csc_from_edges(7, [(0, i) for i in range(1, 7)])
    S, sig, _ = all_variants(g, 3)
    np.testing.assert_array_equal(S, [1, 2, 2, 0, 2, 2, 2])
    assert (sig == 1).all()


def test_complete_bipartite():
    a, b = 4, 6
    g = csc_from_edges(a + b, [(i, a + j) for i in range(a) for j in range(b)])
    S, sig, _ = all_variants(g, 0)
    assert (S[1:a] == 2).all() and (S[a:] == 1).all()
    assert (sig[1:a] == b).all() and (sig[a:] == 1).all()


def test_hypercube_factorials():
    """Q_k: a vertex at Hamming distance j from 0 has j! shortest paths;
    20! < 2^64 < 21! so k = 21 crosses the 64-bit range at j = 21."""
    for k in (6, 21):
        n = 1 << k
        edges = [(v, v | (1 << b)) for v in range(n) for b in range(k) if not v >> b & 1]
        g = csc_from_edges(n, edges)
        S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, 0)
        pop = np.array([bin(v).count("1") for v in range(n)])
        np.testing.assert_array_equal(S, pop)
        for j in range(k + 1):
            val, flag = expected_pair(math.factorial(j))
            sel = pop == j
            assert (sig[sel] == np.uint64(val)).all() and (ovf[sel] == flag).all(), j
        if k == 6:
            all_variants(g, 0)


def test_grid_binomials_cross_64_bits():
    """Unperturbed w x h grid from a corner: sigma(x, y) = C(x+y, x), which
    first exceeds 2^64 at x + y = 68."""
    w = h = 40
    g = graphs.grid(w, h, 0.0, seed=0, relabel=False)
    S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, 0)
    S1, s1, o1, _ = oracle.seq_td_csc(g.CP, g.IC, 0, s_unreached=-1)
    np.testing.assert_array_equal(S, S1)
    np.testing.assert_array_equal(sig, s1)
    np.testing.assert_array_equal(ovf, o1)
    crossed = 0
    for y in range(h):
        for x in range(w):
            v = y * w + x
            assert S[v] == x + y
            val, flag = expected_pair(math.comb(x + y, x))
            assert sig[v] == np.uint64(val) and ovf[v] == flag, (x, y)
            crossed += flag
    assert crossed > 0


# ---------------------------------------------------- brute force (tiny) ---
def brute_force(n, adj_out, r):
    """Enumerate every simple path from r by DFS; the shortest ones give S and
    sigma.  Exponential; only for tiny graphs."""
    best = {r: 0}
    count = {r: 1}

    def dfs(v, depth, seen):
        for w in adj_out[v]:
            if w in seen:
                continue
            nd = depth + 1
            if w not in best or nd < best[w]:
                best[w] = nd
                count[w] = 1
            elif nd == best[w]:
                count[w] += 1
            seen.add(w)
            dfs(w, nd, seen)
            seen.discard(w)

    dfs(r, 0, {r})
    S = np.array([best.get(v, -1) for v in range(n)], np.int32)
    sig = np.array([count.get(v, 0) for v in range(n)], np.uint64)
    return S, sig


@pytest.mark.parametrize("seed", range(12))
def test_brute_force_random_digraphs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 9))
    pairs = [(u, v) for u in range(n) for v in range(n) if u != v]
    m = int(rng.integers(n, min(len(pairs), 3 * n) + 1))
    pick = rng.choice(len(pairs), size=m, replace=False)
    edges = [pairs[i] for i in pick]
    g = csc_from_edges(n, edges, undirected=False)
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
    for r in range(n):
        S_bf, sig_bf = brute_force(n, adj, r)
        S, sig, ovf = all_variants(g, r)
        np.testing.assert_array_equal(S, S_bf)
        np.testing.assert_array_equal(sig, sig_bf)
        assert not ovf.any()


def test_pass_capped_run_is_a_prefix_of_the_full_run():
    """orc_bfs_seq_td_csc_passes (the timing sample of long BFS runs) is the
    same loop stopped after k passes: depths up to k, and the counts of the
    depths below k, equal the full run's; nothing else is reached."""
    g = graphs.grid(40, 40, 0.1, seed=3, relabel=True)
    S, sig, ovf, passes = oracle.seq_td_csc(g.CP, g.IC, 5)
    for k in (1, 5, 17, passes + 3):
        Sp, sp, op, d = oracle.seq_td_csc_passes(g.CP, g.IC, 5, k)
        assert d == min(k, passes)
        np.testing.assert_array_equal(Sp, np.where(S <= k, S, -1))
        np.testing.assert_array_equal(sp, np.where((S >= 0) & (S < k), sig, 0))
    with pytest.raises(ValueError):
        oracle.seq_td_csc_passes(g.CP, g.IC, 5, 0)


def test_td_convention_unreached_zero():
    g = csc_from_edges(5, [(0, 1), (1, 2)])
    S0, _, _, passes = oracle.seq_td_csc(g.CP, g.IC, 0, s_unreached=0)
    np.testing.assert_array_equal(S0, [0, 1, 2, 0, 0])
    assert passes == 3  # the reference's final d: one pass past the last depth


def test_planted_mistake_is_caught():
    """A one-edge change in the oracle's input changes its answer, and the
    independent BFS sees the same change: the comparisons are not vacuous."""
    g = graphs.erdos_renyi(300, 0.02, seed=5)
    S, sig, _ = all_variants(g, 0)
    IC = g.IC.copy()
    # a depth-1 vertex whose only parent is the root: point that in-edge at a
    # vertex of depth >= 3 instead, so v's depth must change
    v = int(np.flatnonzero((S == 1) & (sig == 1))[0])
    k = g.CP[v] + int(np.flatnonzero(g.IC[g.CP[v]:g.CP[v + 1]] == 0)[0])
    IC[k] = int(np.flatnonzero(S >= 3)[0])
    S2, sig2, _, _ = oracle.seq_td_csc(g.CP, IC, 0, s_unreached=-1)
    S3, sig3, _, _ = oracle.textbook(g.CP, IC, 0)
    assert not (np.array_equal(S, S2) and np.array_equal(sig, sig2))
    np.testing.assert_array_equal(S2, S3)
    np.testing.assert_array_equal(sig2, sig3)


# ------------------------------------- the reference's bundled .mtx files ---
KNOWN = {
    ("mycielskian4.mtx", 0): ([0, 1, 2, 1, 2, 2, 1, 2, 1, 2, 2],
                              [1, 1, 2, 1, 2, 2, 1, 1, 1, 1, 2]),
    ("Tina_AskCal.mtx", 10): ([2, 2, 1, 2, -1, 1, -1, 1, -1, -1, 0],
                              [1, 3, 1, 2, 0, 1, 0, 1, 0, 0, 1]),
    ("cage4.mtx", 0): ([0, 1, 2, 1, 1, 2, 2, 1, 2], [1, 1, 2, 1, 1, 3, 3, 1, 2]),
}


@pytest.mark.parametrize("key", list(KNOWN))
def test_known_answers_small_files(key):
    fname, r = key
    g = mtx.read_csc(golden(fname))
    S, sig, _ = all_variants(g, r)
    np.testing.assert_array_equal(S, KNOWN[key][0])
    np.testing.assert_array_equal(sig, KNOWN[key][1])


def test_known_answers_ragusa18_and_luxembourg():
    g = mtx.read_csc(golden("Ragusa18.mtx"))
    S, sig, _ = all_variants(g, 0)
    assert S.max() == 4 and (S >= 0).sum() == 18 and sig.max() == 2
    g = mtx.read_csc(golden("luxembourg_osm.mtx.gz"))
    S, sig, ovf, levels = oracle.textbook(g.CP, g.IC, 0)
    assert levels == 1035 and S.max() == 1034
    assert (S >= 0).all() and g.n == 114599 and sig.max() == 420 and not ovf.any()


# ------------------------------- the reference's own sequential C code ---
needs_ref = pytest.mark.skipif(not oracle.ref_available("td"),
                               reason="reference sources were not built into oracle/_ref")


def _ref_cases():
    yield "er300", graphs.erdos_renyi(300, 0.02, seed=5), [0, 7, 150]
    yield "rmat10", graphs.rmat(10, 8, seed=3), [1, 2, 3]
    yield "grid20", graphs.grid(20, 20, 0.1, seed=3), [0, 399]
    for f, r in (("mycielskian4.mtx", 0), ("Tina_AskCal.mtx", 10), ("cage4.mtx", 0),
                 ("Ragusa18.mtx", 0)):
        yield f, mtx.read_csc(golden(f)), [r]


@needs_ref
@pytest.mark.parametrize("name,g,roots", list(_ref_cases()), ids=lambda x: x if isinstance(x, str) else "")
def test_restatement_matches_reference_code(name, g, roots):
    for r in roots:
        S, sig, ovf = all_variants(g, r)
        assert not ovf.any() and sig.max() < (1 << 24)
        # CSC_TD: float sigma, S stays 0 where unreached
        S_td, sig_td = oracle.ref_seq_td_csc(g.CP, g.IC, r, "td")
        np.testing.assert_array_equal(S_td, np.where(S < 0, 0, S))
        np.testing.assert_array_equal(sig_td.astype(np.uint64), sig)
        # CSC_BU copy: float sigma, S = -1 where unreached
        S_bu, sig_bu = oracle.ref_seq_td_csc(g.CP, g.IC, r, "bu")
        np.testing.assert_array_equal(S_bu, S)
        np.testing.assert_array_equal(sig_bu.astype(np.uint64), sig)
        # TDBU copy: int sigma
        S_i, sig_i = oracle.ref_seq_td_csc(g.CP, g.IC, r, "tdbu")
        np.testing.assert_array_equal(S_i, S)
        np.testing.assert_array_equal(sig_i.astype(np.uint64), sig)
        # COOC: float sigma
        I, J = graphs.csc_to_cooc(g)
        sig_c = oracle.ref_seq_td_cooc(I, J, g.n, r)
        np.testing.assert_array_equal(sig_c.astype(np.uint64), sig)


@needs_ref
def test_reference_int_counts_break_discovery_past_int32():
    """The reference's int copy (TurboBFS_CSC_TDBU/bfstdcsc_seq.c) agrees with
    the restatement up to the first depth whose counts exceed int32; past it
    a wrapped (negative) sum fails the reference's "> 0.9" test and its S goes
    wrong.  The restatement keeps exact counts mod 2^64 with the overflow
    flag, and the library reports TBFS_SIGMA_INEXACT for that range instead
    of reproducing the breakage."""
    g = graphs.grid(24, 24, 0.0, seed=0, relabel=False)
    S, sig, ovf = all_variants(g, 0)
    assert oracle.count_exceed(sig, ovf, 2**31 - 1) > 0
    S_i, sig_i = oracle.ref_seq_td_csc(g.CP, g.IC, 0, "tdbu")
    first_big = int(S[sig > 2**31 - 1].min())
    ok = S < first_big
    np.testing.assert_array_equal(S_i[ok], S[ok])
    np.testing.assert_array_equal(sig_i[ok].astype(np.uint64), sig[ok])
    assert not np.array_equal(S_i, S)  # the reference's own result breaks


def test_count_exceed_boundaries():
    """oracle.count_exceed(sigma, ovf, limit) = number of vertices whose exact
    count is above `limit` (sigma > limit, or the 2^64 overflow flag): the
    test the drop-in entry points' TBFS_SIGMA_INEXACT status follows, at the
    float limit 2^24 and the int32 limit 2^31 - 1."""
    for limit in (2**24, 2**31 - 1):
        sig = np.array([0, 1, limit - 1, limit, limit + 1, 2**63, 2**64 - 1, 5], np.uint64)
        ovf = np.array([0, 0, 0, 0, 0, 0, 0, 1], np.uint8)
        # above the limit: limit + 1, 2^63, 2^64 - 1, and the flagged 5
        assert oracle.count_exceed(sig, ovf, limit) == 4
        assert oracle.count_exceed(sig[:4], ovf[:4], limit) == 0  # == limit is exact
        assert oracle.count_exceed(sig, np.zeros_like(ovf), limit) == 3
    # closed form: a k x k grid from its corner has C(i + j, i) paths to (i, j)
    from math import comb
    k = 24
    g = graphs.grid(k, k, 0.0, seed=0, relabel=False)
    S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, 0)
    for limit in (2**24, 2**31 - 1):
        want = sum(1 for i in range(k) for j in range(k) if comb(i + j, i) > limit)
        assert want > 0
        assert oracle.count_exceed(sig, ovf, limit) == want
