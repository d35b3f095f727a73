"""Host-side graph construction (CPU only): generators, packing, the Matrix
Market reader, and the golden fixtures' inputs."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from turbobfs_amd import graphs, mtx
from tests.conftest import golden
from tests.golden.make_fixtures import csc_sha


def _check_simple_symmetric(g):
    n = g.n
    assert g.CP[0] == 0 and g.CP[-1] == g.nz and (np.diff(g.CP) >= 0).all()
    assert g.IC.min(initial=0) >= 0 and g.IC.max(initial=0) < n
    col = np.repeat(np.arange(n), np.diff(g.CP))
    assert not (col == g.IC).any(), "self-loop"
    key = col.astype(np.int64) * n + g.IC
    assert (np.diff(key) > 0).all(), "rows not strictly ascending in a column"
    t = graphs.transpose(g)
    np.testing.assert_array_equal(t.CP, g.CP)
    np.testing.assert_array_equal(t.IC, g.IC)


def test_er_deterministic_and_simple():
    a = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    b = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    c = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=2)
    assert csc_sha(a) == csc_sha(b) != csc_sha(c)
    _check_simple_symmetric(a)
    # E ~ Binomial(4096*4095/2, 8/4095): mean 16384, sd ~ 128
    assert abs(a.nz / 2 - 16384) < 6 * 128


def test_er_fixture_is_what_the_generator_makes():
    fx = np.load(golden("er4096_seed1.npz"))
    g = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    np.testing.assert_array_equal(g.CP, fx["CP"])
    np.testing.assert_array_equal(g.IC, fx["IC"])
    assert csc_sha(g) == str(fx["sha256"])


def test_rmat_fixture_hash_and_properties():
    fx = np.load(golden("rmat12_seed2.npz"))
    g = graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    assert csc_sha(g) == str(fx["sha256"]) and g.n == int(fx["n"]) and g.nz == int(fx["nz"])
    _check_simple_symmetric(g)
    deg = np.diff(g.CP)
    assert deg.max() > 20 * deg[deg > 0].mean()  # skewed (power-law-like)


def test_rmat_relabel_is_a_permutation():
    a = graphs.rmat(10, 8, seed=4, relabel=False)
    b = graphs.rmat(10, 8, seed=4, relabel=True)
    assert a.nz == b.nz
    assert sorted(np.diff(a.CP)) == sorted(np.diff(b.CP))


def test_grid_edges_removed_exactly():
    w, h = 64, 48
    E = h * (w - 1) + w * (h - 1)
    full = graphs.grid(w, h, 0.0, seed=3, relabel=False)
    assert full.nz == 2 * E
    cut = graphs.grid(w, h, 0.1, seed=3, relabel=True)
    assert cut.nz == 2 * (E - round(0.1 * E))
    _check_simple_symmetric(cut)
    deg = np.diff(full.CP)
    assert deg.max() == 4 and deg.min() == 2


def test_watts_strogatz_degree_sum_and_simple():
    g = graphs.watts_strogatz(2000, 10, 0.1, seed=9)
    _check_simple_symmetric(g)
    assert g.nz == 2000 * 10  # rewiring keeps the edge count
    g0 = graphs.watts_strogatz(2000, 10, 0.0, seed=9)
    assert (np.diff(g0.CP) == 10).all()


def test_block_diagonal_packing():
    parts = [graphs.watts_strogatz(n, 4, 0.2, seed=i) for i, n in enumerate((50, 80, 33))]
    g, off = graphs.block_diagonal(parts)
    assert g.n == 163 and g.nz == sum(p.nz for p in parts)
    for i, p in enumerate(parts):
        for r in (0, p.n - 1):
            S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, int(off[i] + r))
            Sp, sigp, _, _ = oracle.textbook(p.CP, p.IC, r)
            np.testing.assert_array_equal(S[off[i]:off[i + 1]], Sp)
            np.testing.assert_array_equal(sig[off[i]:off[i + 1]], sigp)
            assert (S[:off[i]] == -1).all() and (S[off[i + 1]:] == -1).all()


def test_sources_are_seeded_and_distinct():
    g = graphs.rmat(10, 8, seed=4)
    a = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 16, seed=2)
    b = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 16, seed=2)
    np.testing.assert_array_equal(a, b)
    assert len(set(a.tolist())) == 16 and (np.diff(g.CP)[a] > 0).all()


# ------------------------------------------------------- Matrix Market ---
def test_mtx_reader_banner_and_counts():
    g = mtx.read_csc(golden("mycielskian4.mtx"))
    assert g.n == 11 and g.nz == 40  # 20 symmetric entries, both directions
    g = mtx.read_csc(golden("Tina_AskCal.mtx"))
    assert g.n == 11 and g.nz == 29
    g = mtx.read_csc(golden("cage4.mtx"))  # real general: value column ignored
    assert g.n == 9 and g.nz == 49


def test_mtx_reader_rejects_garbage(tmp_path):
    p = tmp_path / "x.mtx"
    p.write_text("not a matrix\n")
    with pytest.raises(mtx.MMError):
        mtx.read_csc(str(p))
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 2\n4 1\n")
    with pytest.raises(mtx.MMError):
        mtx.read_csc(str(p))


needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(oracle.REF_DIR, "libref_fmt.so")),
                               reason="reference sources were not built into oracle/_ref")


def _reference_csc(path, ug, weighted):
    """Run the reference's own reader + SSC->CSC transform (compiled from its
    sources) on a plain .mtx file."""
    lib = C.CDLL(os.path.join(oracle.REF_DIR, "libref_fmt.so"))
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    f = libc.fopen(path.encode(), b"r")
    assert f
    code = (C.c_char * 4)()
    lib.mm_read_banner.argtypes = [C.c_void_p, C.c_void_p]
    assert lib.mm_read_banner(C.c_void_p(f), code) == 0
    M, N, nz = C.c_int(), C.c_int(), C.c_int()
    lib.mm_read_mtx_crd_size.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    assert lib.mm_read_mtx_crd_size(C.c_void_p(f), C.byref(M), C.byref(N), C.byref(nz)) == 0
    n = N.value
    nzf = nz.value
    i32 = lambda k: np.zeros(k, np.int32)  # noqa: E731
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rows, cols, out_d, in_d = i32(nzf), i32(nzf), i32(n + 2), i32(n + 2)
    if weighted:
        vals = np.zeros(nzf, np.float32)
        lib.readMMfile_w.argtypes = [C.c_void_p] * 6 + [C.c_int]
        lib.readMMfile_w(C.c_void_p(f), p(rows), p(cols), p(vals), p(out_d), p(in_d), nzf)
    else:
        lib.readMMfile_uw.argtypes = [C.c_void_p] * 5 + [C.c_int]
        lib.readMMfile_uw(C.c_void_p(f), p(rows), p(cols), p(out_d), p(in_d), nzf)
    libc.fclose(C.c_void_p(f))
    # arrays get two spare slots: CP_RP_* write index n+1 (loop runs to i <= n)
    if ug:
        deg, CPLT, RPLT, CP, RP = i32(n + 2), i32(n + 2), i32(n + 2), i32(n + 2), i32(n + 2)
        lib.degree_ug.argtypes = [C.c_void_p] * 3 + [C.c_int]
        lib.degree_ug(p(out_d), p(in_d), p(deg), n)
        lib.CP_RP_ug.argtypes = [C.c_void_p] * 7 + [C.c_int]
        lib.CP_RP_ug(p(out_d), p(in_d), p(deg), p(CPLT), p(RPLT), p(CP), p(RP), n)
        IC = i32(2 * nzf)
        lib.SSCCSC_uw.argtypes = [C.c_void_p] * 6 + [C.c_int]
        lib.SSCCSC_uw(p(rows), p(CPLT), p(RPLT), p(RP), p(CP), p(IC), n)
        return graphs.CSC(CP[:n + 1].copy(), IC)
    CP, RP = i32(n + 2), i32(n + 2)
    lib.CP_RP_dg.argtypes = [C.c_void_p] * 4 + [C.c_int]
    lib.CP_RP_dg(p(out_d), p(in_d), p(CP), p(RP), n)
    return graphs.CSC(CP[:n + 1].copy(), rows)


@needs_ref
@pytest.mark.parametrize("fname,ug,w", [("mycielskian4.mtx", 1, 0), ("Tina_AskCal.mtx", 0, 0),
                                         ("cage4.mtx", 0, 1), ("Ragusa18.mtx", 0, 1)])
def test_mtx_reader_matches_reference_reader(fname, ug, w):
    ref = _reference_csc(golden(fname), ug, w)
    ours = mtx.read_csc(golden(fname), ug=bool(ug))
    np.testing.assert_array_equal(ours.CP, ref.CP)
    np.testing.assert_array_equal(ours.IC, ref.IC)


@needs_ref
def test_mtx_reader_matches_reference_reader_luxembourg(tmp_path):
    import gzip
    import shutil

    plain = tmp_path / "luxembourg_osm.mtx"
    with gzip.open(golden("luxembourg_osm.mtx.gz"), "rb") as src, open(plain, "wb") as dst:
        shutil.copyfileobj(src, dst)
    ref = _reference_csc(str(plain), 1, 0)
    ours = mtx.read_csc(golden("luxembourg_osm.mtx.gz"))
    np.testing.assert_array_equal(ours.CP, ref.CP)
    np.testing.assert_array_equal(ours.IC, ref.IC)
    fx = np.load(golden("luxembourg_osm_r0.npz"))
    assert csc_sha(ours) == str(fx["sha256"])
