"""Benchmark: single-source BFS with exact path counts on R-MAT scale 20.

Workload (one step = one BFS from one source, graph resident in HBM):
  R-MAT scale 20 (n = 1,048,576), edge factor 16, a=.57 b=.19 c=.19 d=.05,
  symmetrized, deduplicated, randomly relabelled (seed 2), int32 CSC;
  64 random sources of nonzero degree (seed 2); steps cycle through them.
value = GTEPS = sum over timed BFS of (m_s / 2) / total time, where m_s is
the number of in-list entries of the vertices reached (undirected edges
counted once).  Every timed BFS computes depths S and exact uint64 path
counts sigma.

Prints ONE JSON line.  Also reports the roofline of the dominant kernel, a
CPU baseline (the reference's own sequential bfs_seq_td_csc, compiled from
its sources into oracle/_ref, else this project's C restatement), and
(unless --no-batch) the batched-sources throughput against the per-source
loop (R-MAT scale 18, 1024 sources per batch), a 512-graph block-diagonal
multi-root BFS against a per-graph loop, the per-level cost of each level
loop on a high-diameter 2048x2048 grid, and (unless --no-serve) the query
server's saturated throughput and p50 / p99 latency under Poisson arrivals.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per BFS measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
# this bench (scaled by the counters' reading of the calibration copy)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bfs_rmat20_traffic.json")


def measured_traffic(scale: int, mode: str):
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("scale") != scale or t.get("mode") != mode:
        return None
    return t.get("hbm_bytes_per_bfs")
METRIC = "GTEPS on 1 MI355X + batched sources/s & p99 latency vs per-source loop, exact"


def _quiet_call(fn, *a):
    """Run fn with the process's stdout fd sent to /dev/null (the reference's
    sequential code prints its timings with printf)."""
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        return fn(*a)
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


def cpu_baseline(g, sources, budget_s: float):
    """Time the reference's sequential BFS (or the restatement) on host cores,
    one thread, on a bounded sample of the same sources."""
    import oracle

    kind = "reference" if oracle.ref_available("td") else "port"
    edges = 0
    t_total = 0.0
    done = 0
    for r in sources:
        t0 = time.perf_counter()
        if kind == "reference":
            S, _ = _quiet_call(oracle.ref_seq_td_csc, g.CP, g.IC, int(r), "td")
            reached = (S > 0)
            reached[int(r)] = True
        else:
            S, _, _, _ = oracle.seq_td_csc(g.CP, g.IC, int(r), s_unreached=-1)
            reached = S >= 0
        t_total += time.perf_counter() - t0
        edges += int(np.diff(g.CP)[reached].sum())
        done += 1
        if t_total > budget_s:
            break
    gteps = (edges / 2) / t_total / 1e9
    return {
        "value": gteps, "unit": "GTEPS", "cores": 1, "kind": kind,
        "sample": f"{done} of {len(sources)} sources, bfs_seq_td_csc "
                  f"(O(levels*(n+nz)) sweep), single thread, "
                  f"{t_total:.1f} s; 1 thread of {os.cpu_count()} host cores ({_cpu_model()})",
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "cpu model unknown"


def bench_dropin(tb, g, sources, nsrc: int = 4):
    """The reference's entry point bfs_gpu_tdbu_csc_sc as a caller of the
    reference uses it: host arrays in, host arrays out, the graph uploaded
    and checked on every call.  Host wall time per call (PCIe included), one
    warm-up call first; with repetition = 8 one upload serves 8 BFS."""
    n, nz = g.n, g.nz
    S = np.zeros(n, np.int32)
    sig = np.zeros(n, np.int32)
    tb.bfs_gpu_tdbu_csc_sc(g.IC, g.CP, S, sig, int(sources[0]), nz, n, 1, 1)
    out = {"entry_point": "bfs_gpu_tdbu_csc_sc (int32 sigma; returns 5 where a count "
                          "leaves int32)"}
    deg = np.diff(g.CP).astype(np.int64)
    for rep in (1, 8):
        t = []
        edges = 0
        for r in sources[:nsrc]:
            t0 = time.perf_counter()
            rc = tb.bfs_gpu_tdbu_csc_sc(g.IC, g.CP, S, sig, int(r), nz, n, rep, 1)
            t.append(time.perf_counter() - t0)
            if rc not in (0, 5):
                raise RuntimeError(f"bfs_gpu_tdbu_csc_sc returned {rc}")
            edges += int(deg[S >= 0].sum())
        ms = 1e3 * sum(t) / len(t)
        out[f"repetition_{rep}"] = {
            "ms_per_call": ms, "ms_per_bfs": ms / rep,
            "gteps_pcie_inclusive": edges / 2 * rep / (sum(t)) / 1e9}
    tb.dropin_release()  # the kept contexts and buffers go before the next section
    return out


def _time_variant(tb, ctx, gr, roots, n, o):
    ctx.run_many(gr, roots[:1], o)
    ctx.run_many(gr, roots[:2], o)  # launch plan history for this graph
    ctx.mark(0)
    st = ctx.run_many(gr, roots, o)
    ctx.mark(1)
    ms = ctx.elapsed(0, 1)
    edges = sum(x["edges"] for x in st)
    b_alg = len(roots) * (4 * (n + 1) + 12 * n) + 4 * edges
    return {"gteps": edges / 2 / (ms * 1e-3) / 1e9, "ms_per_bfs": ms / len(roots),
            "alg_gbs": b_alg / (ms * 1e-3) / 1e9, "sources": len(roots)}


def bench_variants(tb, g, gr, ctx, sources, nsrc: int = 16, names=None):
    """GTEPS and algorithmic GB/s (B_alg / time) on the resident R-MAT graph,
    sources in turn (`names`: only these rows).  `entry_points`: the kernels
    the seven reference entry points run by default (each distinct
    configuration once, with the entry points it serves; the COOC entry point
    on a COOC upload of the graph); `reference_mappings`: the reference's own
    thread / warp mappings, which tbfs_dropin_mapping(TBFS_DROPIN_REFERENCE)
    selects.  `measured_gbs`: HBM bytes per BFS from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes of this table, divided by the time."""
    from turbobfs_amd import _lib
    H, CH, LB = _lib.LOOP_HOST, _lib.LOOP_CHAINED, _lib.PULL_LB
    roots = np.asarray(sources[:nsrc], np.int32)
    n = g.n
    traffic = _variant_traffic()
    out = {"entry_points": {}, "reference_mappings": {}}
    fast = [
        ("tdbu_sigma", "bfs_gpu_td_csc_sc, bfs_gpu_td_csc_wa, bfs_gpu_tdbu_csc_sc, "
                       "bfs_gpu_tdbu_csc_wa",
         dict(mode=_lib.MODE_AUTO, pull=LB, loop=CH), False),
        ("tdbu_depth_only", "bfs_gpu_bu_csc_sc, bfs_gpu_bu_csc_wa",
         dict(mode=_lib.MODE_AUTO, pull=LB, loop=CH, sigma=0, bbreak=1), False),
        ("tdbu_sigma_cooc_upload", "bfs_gpu_td_cooc_sc",
         dict(mode=_lib.MODE_AUTO, pull=LB, loop=CH), True),
    ]
    for name, eps, kw, cooc in fast:
        if names and name not in names:
            continue
        if cooc:
            I, J = _cooc_of(g)
            with tb.Graph(I=I, J=J, n=n) as gc:
                row = _time_variant(tb, ctx, gc, roots, n, tb.default_opts(**kw))
        else:
            row = _time_variant(tb, ctx, gr, roots, n, tb.default_opts(**kw))
        row["entry_points"] = eps
        _add_measured(row, traffic.get(name))
        out["entry_points"][name] = row
    ref = [
        ("td_csc_sc", dict(mode=_lib.MODE_PULL_SC, loop=H)),
        ("td_csc_wa", dict(mode=_lib.MODE_PULL_WA, loop=H)),
        ("bu_csc_sc_bbreak", dict(mode=_lib.MODE_PULL_SC, loop=H, sigma=0, bbreak=1)),
        ("bu_csc_wa_bbreak", dict(mode=_lib.MODE_PULL_WA, loop=H, sigma=0, bbreak=1)),
        ("td_cooc", dict(mode=_lib.MODE_COOC, loop=H)),
        ("tdbu_push_pull_host", dict(mode=_lib.MODE_AUTO, pull=LB, loop=H)),
    ]
    for name, kw in ref:
        if names and name not in names:
            continue
        row = _time_variant(tb, ctx, gr, roots, n, tb.default_opts(**kw))
        _add_measured(row, traffic.get(name))
        out["reference_mappings"][name] = row
    return out


def _cooc_of(g):
    """COOC of a CSC graph (edge IC[k] -> its column), ordered by column."""
    J = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.CP))
    return np.ascontiguousarray(g.IC), J


VARIANT_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "variants_rmat20_traffic.json")


def _variant_traffic():
    try:
        with open(VARIANT_TRAFFIC_FILE) as f:
            return json.load(f).get("hbm_bytes_per_bfs", {})
    except (OSError, ValueError):
        return {}


def _add_measured(row, bytes_per_bfs):
    if bytes_per_bfs:
        row["hbm_bytes_per_bfs"] = bytes_per_bfs
        row["measured_gbs"] = bytes_per_bfs / (row["ms_per_bfs"] * 1e-3) / 1e9


def bench_er4096(tb, graphs, reps: int = 256, cpu: bool = True):
    """The CPU-runnable case: undirected G(n=4096, p=8/4095), seed 1, source
    0.  `reps` BFS from source 0 back to back on the resident graph; the
    reference entry point bfs_gpu_td_csc_sc timed as its callers use it; the
    CPU baseline (the reference's sequential BFS) timed in full."""
    g = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    roots = np.zeros(reps, np.int32)
    with tb.Graph(g.CP, g.IC, symmetric=True) as gr, tb.Context(g.n) as ctx:
        ctx.run_many(gr, roots[:8])
        ctx.mark(0)
        st = ctx.run_many(gr, roots)
        ctx.mark(1)
        ms = ctx.elapsed(0, 1)
    edges = sum(x["edges"] for x in st)
    out = {"workload": "Erdos-Renyi G(4096, 8/4095) seed 1, source 0",
           "n": g.n, "nz": g.nz, "levels": st[0]["levels"], "bfs": reps,
           "gteps": edges / 2 / (ms * 1e-3) / 1e9, "us_per_bfs": 1e3 * ms / reps}
    S = np.zeros(g.n, np.int32)
    sig = np.zeros(g.n, np.float32)
    tb.bfs_gpu_td_csc_sc(g.IC, g.CP, S, sig, 0, g.nz, g.n, 1)
    t0 = time.perf_counter()
    k = 20
    for _ in range(k):
        tb.bfs_gpu_td_csc_sc(g.IC, g.CP, S, sig, 0, g.nz, g.n, 1)
    out["dropin_bfs_gpu_td_csc_sc_ms_per_call"] = 1e3 * (time.perf_counter() - t0) / k
    tb.dropin_release()
    if cpu:  # source 0 repeated until ~2 s of CPU time, every BFS timed in full
        out["cpu_baseline"] = cpu_baseline(g, [0] * 1000, 2.0)
    return out


def bench_batched(tb, graphs, reps: int = 3, tune: bool = True, loop: bool = True):
    """Batched sources vs the per-source loop on the same GPU: R-MAT scale 18
    (ef16, a=.57 b=.19 c=.19, symmetrized, relabelled, seed 4), 1024 sources
    of nonzero degree (seed 4) in one batch, 64 per frontier word."""
    g = graphs.rmat(18, 16, 0.57, 0.19, 0.19, seed=4, relabel=True)
    src = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 1024, seed=4)
    with tb.Graph(g.CP, g.IC, symmetric=True) as gr:
        return _bench_batched(tb, g, gr, src, reps, tune, loop)


def _bench_batched(tb, g, gr, src, reps, tune, loop):
    b = tb.Batch(g.n, len(src))
    b.run(gr, src)  # warm-up
    runs = [b.run(gr, src) for _ in range(reps)]
    ms_default = float(np.median([r["ms"] for r in runs]))
    # autotuned launch grid and push / pull threshold for this graph
    best = b.autotune(gr, src)[0] if tune else tb.default_opts()
    runs = [b.run(gr, src, best) for _ in range(reps)]
    ms = float(np.median([r["ms"] for r in runs]))
    st = runs[0]
    batch_sps = len(src) / (ms * 1e-3)
    b.close()
    if loop:
        ctx = tb.Context(g.n)
        ctx.run_many(gr, src[:64])  # warm-up
        ctx.mark(0)
        per = ctx.run_many(gr, src)
        ctx.mark(1)
        loop_ms = ctx.elapsed(0, 1)
        loop_sps = len(src) / (loop_ms * 1e-3)
        loop_gteps = sum(x["edges"] for x in per) / 2 / (loop_ms * 1e-3) / 1e9
        ctx.close()
    else:
        loop_sps = loop_gteps = float("nan")
    # the algorithmic bytes of the 1024 single-source BFS the batch replaces
    # (4(n+1) + 4 m_s + 12 n each): what one launch stands in for, not its
    # HBM traffic (the batch reads each in-list once per 64 sources)
    b_eq = len(src) * (4 * (g.n + 1) + 12 * g.n) + 4 * st["edges"]
    return {
        "workload": "R-MAT scale 18 ef16 symmetrized relabelled seed 4, 1024 sources (seed 4) "
                    "in one batch; every source's S and sigma exact",
        "n": g.n, "nz": g.nz, "sources": len(src),
        "sources_per_s": batch_sps,
        "batch_ms": ms,
        "batch_ms_default_opts": ms_default,
        "autotuned": {"grid": int(best.grid), "batch_alpha": float(best.batch_alpha)},
        "gteps": st["edges"] / 2 / (ms * 1e-3) / 1e9,
        "levels": st["levels"], "push_levels": st["push_levels"],
        "pull_levels": st["pull_levels"], "count_bits": st["count_bits"],
        "per_source_loop_sources_per_s": loop_sps,
        "per_source_loop": "tbfs_bfs_many, chained direction-optimizing loop, same GPU",
        "speedup_vs_per_source_loop": batch_sps / loop_sps,
        "per_source_loop_gteps": loop_gteps,
        "single_source_equivalent": {
            "bytes": b_eq, "gbs": b_eq / (ms * 1e-3) / 1e9,
            "frac_of_hbm_peak": b_eq / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "note": "algorithmic bytes of the single-source BFS the batch replaces per second; "
                    "not HBM traffic"},
    }


def bench_multigraph(tb, graphs, reps: int = 3):
    """512 uneven Watts-Strogatz graphs (n ~ U[1000, 64000], k=10, p=0.1) as
    diagonal blocks of one CSC, one source each, in one multi-root BFS; against
    a loop of single-source BFS over the same resident graphs."""
    rng = np.random.default_rng(5)
    parts = [graphs.watts_strogatz(int(rng.integers(1000, 64001)), 10, 0.1, seed=i)
             for i in range(512)]
    bd, off = graphs.block_diagonal(parts)
    roots = np.array([off[i] + int(rng.integers(0, parts[i].n)) for i in range(len(parts))],
                     np.int32)
    with tb.Graph(bd.CP, bd.IC, symmetric=True) as gb, tb.Context(bd.n) as ctx:
        ctx.run_multi(gb, roots)
        runs = [ctx.run_multi(gb, roots) for _ in range(reps)]
    ms = float(np.median([r["ms"] for r in runs]))
    grs = [tb.Graph(p.CP, p.IC, symmetric=True) for p in parts]
    try:
        with tb.Context(max(p.n for p in parts)) as cs:
            for i in range(8):
                cs.run(grs[i], int(roots[i] - off[i]))
            t0 = time.perf_counter()
            for i in range(len(parts)):
                cs.run(grs[i], int(roots[i] - off[i]))
            loop_ms = (time.perf_counter() - t0) * 1e3
    finally:
        for x in grs:
            x.close()
    return {
        "workload": "512 Watts-Strogatz graphs n~U[1000,64000] k=10 p=0.1 block-diagonal, "
                    "one source each",
        "n": bd.n, "nz": bd.nz, "graphs": len(parts),
        "multi_root_ms": ms, "graphs_per_s": len(parts) / (ms * 1e-3),
        "levels": runs[0]["levels"],
        "per_graph_loop_ms": loop_ms,
        "per_graph_loop": "tbfs_bfs per graph, resident graphs, host wall time",
        "speedup_vs_per_graph_loop": loop_ms / ms,
    }


def bench_grid(tb, graphs, nsrc: int = 16, names=("chained", "graph", "persistent_1024x64", "host"),
               cpu_passes: int = 0):
    """High-diameter single-source BFS: 2048x2048 4-neighbour grid with 10% of
    its edges removed (seed 3), vertex ids permuted, sources in the largest
    component (seed 3).  Per-level cost of each level loop; with cpu_passes,
    the restated sequential BFS timed over its first cpu_passes passes and
    its whole run extrapolated (one thread)."""
    g = graphs.grid(2048, 2048, 0.1, seed=3, relabel=True)
    lc = graphs.largest_component_vertices(g)
    src = graphs.sample_sources(lc, nsrc, seed=3)
    with tb.Graph(g.CP, g.IC, symmetric=True) as gr, tb.Context(g.n) as ctx, \
            tb.Batch(g.n, nsrc) as b:
        return _bench_grid(tb, g, gr, ctx, b, src, nsrc, names, cpu_passes)


def _bench_grid(tb, g, gr, ctx, b, src, nsrc, names, cpu_passes):
    from turbobfs_amd import _lib
    loops = [
        ("chained", dict(loop=_lib.LOOP_CHAINED)),
        ("graph", dict(loop=_lib.LOOP_GRAPH)),
        ("persistent_1024x64", dict(loop=_lib.LOOP_PERSISTENT, block=1024, grid=64)),
        ("host", dict(loop=_lib.LOOP_HOST)),
    ]
    out = {"workload": "2048x2048 grid, 10% of edges removed (seed 3), relabelled; "
                       f"{nsrc} sources in the largest component (seed 3)",
           "n": g.n, "nz": g.nz, "loops": {}}
    for name, kw in loops:
        if name not in names:
            continue
        o = tb.default_opts(mode=_lib.MODE_AUTO, pull=_lib.PULL_LB, **kw)
        k = nsrc if name != "host" else 4
        ctx.run_many(gr, src[:2], o)   # warm-up (level budget, graph capture)
        st = ctx.run_many(gr, src[:k], o)
        ms = sum(x["ms"] for x in st)
        lv = sum(x["levels"] for x in st)
        out["loops"][name] = {"sources": k, "ms_per_bfs": ms / k, "levels_per_bfs": lv / k,
                              "us_per_level": 1e3 * ms / lv,
                              "gteps": sum(x["edges"] for x in st) / 2 / (ms * 1e-3) / 1e9}
    b.run(gr, src)
    st = b.run(gr, src)
    out["batch_of_sources"] = {"sources": nsrc, "ms": st["ms"], "levels": st["levels"],
                               "sources_per_s": nsrc / (st["ms"] * 1e-3)}
    if cpu_passes > 0:  # the restated sequential BFS: a timed sample of its passes
        import oracle
        levels0 = ctx.run_many(gr, src[:1])[0]["levels"]
        t0 = time.perf_counter()
        made = oracle.seq_td_csc_passes(g.CP, g.IC, int(src[0]), cpu_passes)[3]
        per_pass = (time.perf_counter() - t0) / made
        out["cpu_baseline"] = {
            "kind": "port", "cores": 1, "us_per_level": per_pass * 1e6,
            "ms_per_bfs_extrapolated": per_pass * (levels0 + 1) * 1e3,
            "sample": f"first {made} of the {levels0 + 1} passes of the restated bfs_seq_td_csc "
                      f"(each pass sweeps every column) from the first source, single thread, "
                      f"{per_pass * made:.1f} s; the whole BFS = passes x time per pass"}
    return out


def bench_serving_process():
    """bench_serving in a process of its own, as a server runs in deployment.
    Measured in the process that ran the other sections, the same server
    loses ~13% of its saturated rate (~2070 -> ~1800 queries/s) and its
    latency at 80% load grows several-fold: after work has run on other HIP
    streams of the process (even streams since destroyed), the four slot
    streams no longer get the hardware queues to themselves.  The child's
    JSON object is returned as it printed it."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--only", "serve"],
                       stdout=subprocess.PIPE, text=True, check=True)
    out = json.loads(r.stdout.strip().splitlines()[-1])["serve"]
    out["process"] = "measured in a child process of its own (bench.py --only serve)"
    return out


def bench_serving(tb, graphs, seconds: float = 1.5, rates=(0.25, 0.5, 0.8)):
    """Query serving: 8 resident graphs (R-MAT scales 16..22, ef16,
    symmetrized, relabelled, seed = scale; a 1024x1024 grid with 10% of its
    edges removed, seed 3), (graph, source) uniform (seed 5).  First the
    saturated rate (a burst of queries submitted at once), then Poisson
    arrivals at fractions of it; latency = submit -> S, sigma, overflow flags
    in the query's pinned host buffers.  The same streams then go to a
    per-source loop: the same server with one slot and one query per group
    (each query a single-source BFS, in arrival order)."""
    cs = [graphs.rmat(sc, 16, 0.57, 0.19, 0.19, seed=sc, relabel=True) for sc in range(16, 23)]
    cs.append(graphs.grid(1024, 1024, 0.1, seed=3, relabel=True))
    names = [f"rmat{sc}" for sc in range(16, 23)] + ["grid1024"]
    gs = [tb.Graph(g.CP, g.IC, symmetric=True) for g in cs]
    try:
        return _bench_serving(tb, cs, names, gs, seconds, rates)
    finally:
        for x in gs:
            x.close()


def _bench_serving(tb, cs, names, gs, seconds, rates):
    # result buffers: a ring of pinned rows per graph, reused round-robin
    rings = []
    for g in cs:
        k = int(min(64, max(4, (256 << 20) // (13 * g.n))))
        rings.append((k, tb.host_array(k * g.n, np.int32), tb.host_array(k * g.n, np.uint64),
                      tb.host_array(k * g.n, np.uint8)))
    use = [0] * len(cs)
    graph_of = {}  # ticket -> graph
    # a seeded sample of the queries (every 97th submitted) gets result
    # buffers of its own, kept for the check at the end
    sampled = []
    nsub = [0]

    def submit(srv, gi, r):
        n = cs[gi].n
        nsub[0] += 1
        if nsub[0] % 97 == 0:
            bufs = (tb.host_array(n, np.int32), tb.host_array(n, np.uint64),
                    tb.host_array(n, np.uint8))
            t = srv.submit(gi, r, *bufs)
            sampled.append((gi, r) + bufs)
        else:
            k, S, sig, ovf = rings[gi]
            j = use[gi] % k
            use[gi] += 1
            t = srv.submit(gi, r, S[j * n:(j + 1) * n], sig[j * n:(j + 1) * n],
                           ovf[j * n:(j + 1) * n])
        graph_of[t] = gi
        return t

    def draw(rng, count):
        gi = rng.integers(0, len(cs), size=count)
        rt = [int(rng.integers(0, cs[int(x)].n)) for x in gi]
        return [(int(a), b) for a, b in zip(gi, rt)]

    def collect(srv, expect):
        out = []
        while len(out) < expect:
            out += srv.poll(4096, wait_us=100000)
        return out

    def run(srv, burst_n, rates_qps):
        rng = np.random.default_rng(5)
        # warm-up: every graph once (first-use slice tables, kernel loading)
        for gi in range(len(cs)):
            submit(srv, gi, 0)
        srv.drain()
        srv.poll(4096)
        # saturated: everything at once
        burst = draw(rng, burst_n)
        for gi, r in burst:
            submit(srv, gi, r)
        done = collect(srv, len(burst))
        t_first = min(d["submit_us"] for d in done)
        t_last = max(d["done_us"] for d in done)
        sat_qps = len(done) / ((t_last - t_first) * 1e-6)
        lat = np.array([d["done_us"] - d["submit_us"] for d in done]) * 1e-3
        out = {"saturated": {"queries": len(done), "queries_per_s": sat_qps,
                             "p50_ms": float(np.percentile(lat, 50)),
                             "p99_ms": float(np.percentile(lat, 99)),
                             "mean_batch": float(np.mean([d["batch"] for d in done]))},
               "poisson": []}
        for rate in (rates_qps(sat_qps) if callable(rates_qps) else rates_qps):
            count = max(64, int(rate * seconds))
            arr = np.cumsum(rng.exponential(1.0 / rate, size=count))
            qs = draw(rng, count)
            t0 = time.perf_counter()
            for (gi, r), ta in zip(qs, arr):
                while time.perf_counter() - t0 < ta:
                    pass
                submit(srv, gi, r)
            done = collect(srv, count)
            lat = np.array([d["done_us"] - d["submit_us"] for d in done]) * 1e-3
            gof = np.array([graph_of[d["ticket"]] for d in done])
            wait = np.array([d["start_us"] - d["submit_us"] for d in done]) * 1e-3
            serv = np.array([d["done_us"] - d["start_us"] for d in done]) * 1e-3
            per_graph = {names[gi]: {"queries": int((gof == gi).sum()),
                                     "p50_ms": float(np.percentile(lat[gof == gi], 50)),
                                     "p99_ms": float(np.percentile(lat[gof == gi], 99)),
                                     "wait_p50_ms": float(np.percentile(wait[gof == gi], 50)),
                                     "service_p50_ms": float(np.percentile(serv[gof == gi], 50))}
                         for gi in range(len(cs)) if (gof == gi).any()}
            span = (max(d["done_us"] for d in done) - min(d["submit_us"] for d in done)) * 1e-6
            out["poisson"].append({
                "offered_qps": rate, "queries": count, "achieved_qps": count / span,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_batch": float(np.mean([d["batch"] for d in done])),
                "admitted_mid_batch": int(sum(1 for d in done if d["level"] > 0)),
                "per_graph": per_graph,
            })
        return out

    srv = tb.Server(gs)
    caps = [srv.info(i)["batch_cap"] for i in range(len(gs))]
    batched = run(srv, 1536, lambda sat: [f * sat for f in rates])
    srv.close()
    for p, f in zip(batched["poisson"], rates):
        p["offered_fraction"] = f
    # the per-source loop at the lowest offered rate (and saturated)
    loop_srv = tb.Server(gs, slots=1, max_batch=1, open_depth=-1)
    loop = run(loop_srv, 256, [batched["poisson"][0]["offered_qps"]])
    loop_srv.close()
    # the sampled served results against the single-source engine (tbfs_bfs:
    # the chained single-source kernels, not the batch kernels that served
    # most queries); the tests check served results against the oracle
    exact = 0
    with tb.Context(max(g.n for g in cs)) as ctx:
        for gi, r, S, sig, ovf in sampled:
            ctx.run(gs[gi], r)
            S2, sig2, ovf2 = ctx.read(cs[gi].n)
            exact += int(np.array_equal(S, S2) and np.array_equal(sig, sig2)
                         and np.array_equal(ovf, ovf2))
    result = {
        "workload": "8 resident graphs: R-MAT scales 16-22 (ef16) and a 1024x1024 grid with "
                    "10% of edges removed; (graph, source) uniform, seed 5; results copied to "
                    "pinned host buffers",
        "results_checked": {"checked": len(sampled), "exact": exact,
                            "against": "single-source tbfs_bfs on the same resident graph "
                                       "(every 97th submitted query)"},
        "graphs": {nm: {"n": g.n, "nz": g.nz, "batch_cap": c} for nm, g, c in zip(names, cs, caps)},
        "saturated": batched["saturated"],
        "poisson": batched["poisson"],
        "per_source_loop": dict(loop, server="one slot, one query per group: a single-source "
                                             "BFS per query in arrival order"),
        "saturated_speedup_vs_per_source_loop":
            batched["saturated"]["queries_per_s"] / loop["saturated"]["queries_per_s"],
    }
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--mode", default="auto_chain",
                    choices=["auto_persistent", "auto_chain", "auto_host", "auto_sc", "auto_wa", "pull_sc",
                             "pull_wa", "push", "cooc"])
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--fb-lds", type=int, default=0, help="-1: no LDS-staged frontier bitmap")
    ap.add_argument("--grid", type=int, default=0, help="workgroups of the persistent kernel (0: all co-resident)")
    ap.add_argument("--alpha", type=float, default=0.0)
    ap.add_argument("--beta", type=float, default=0.0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-sources and multi-graph measurements")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-er", action="store_true",
                    help="skip the Erdos-Renyi n=4096 (CPU-runnable case) measurement")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in entry point's PCIe-inclusive timing")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the per-variant (TD / BU / TDBU / COOC) GTEPS table")
    ap.add_argument("--variants", default="",
                    help="comma-separated subset of the per-variant table (default: all)")
    ap.add_argument("--only", choices=["batch", "multigraph", "grid", "serve", "variants",
                                       "er4096"],
                    default=None,
                    help="run only that measurement and print its JSON object (for profiling)")
    ap.add_argument("--no-tune", action="store_true",
                    help="default options only: no autotuning of the single-source or batched runs")
    ap.add_argument("--grid-loops", default="chained,graph,persistent_1024x64,host",
                    help="level loops timed by --only grid (comma-separated)")
    ap.add_argument("--no-loop", action="store_true",
                    help="batched run without the per-source loop it is compared with")
    ap.add_argument("--calib-mib", type=int, default=512,
                    help="size of the plain copy kernel run first (byte-counter reference)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the single-source BFS is measured on one MI355X", file=sys.stderr)
        sys.exit(2)

    import turbobfs_amd as tb
    from turbobfs_amd import _lib, graphs

    tb.load()
    if tb.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    if args.only == "variants":
        g = graphs.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
        sources = graphs.sample_sources(graphs.nonzero_degree_vertices(g), args.sources, seed=2)
        with tb.Graph(g.CP, g.IC, symmetric=True) as gr, tb.Context(g.n) as ctx:
            if args.calib_mib > 0:  # reference dispatch for byte counters
                ctx.calibrate_copy(args.calib_mib << 20, 1)
            names = [x for x in args.variants.split(",") if x]
            out = {"variants": bench_variants(tb, g, gr, ctx, sources, names=names)}
        print(json.dumps(out))
        return
    if args.only:
        fn = {"batch": lambda: bench_batched(tb, graphs, tune=not args.no_tune,
                                             loop=not args.no_loop),
              "multigraph": lambda: bench_multigraph(tb, graphs),
              "grid": lambda: bench_grid(tb, graphs, names=tuple(args.grid_loops.split(",")),
                                         cpu_passes=0 if args.no_cpu else 24),
              "serve": lambda: bench_serving(tb, graphs),
              "er4096": lambda: bench_er4096(tb, graphs, cpu=not args.no_cpu)}[args.only]
        print(json.dumps({args.only: fn()}))
        return
    g = graphs.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    sources = graphs.sample_sources(graphs.nonzero_degree_vertices(g), args.sources, seed=2)
    H, P = _lib.LOOP_HOST, _lib.LOOP_PERSISTENT
    modes = {
        "auto_persistent": dict(mode=_lib.MODE_AUTO, pull=_lib.PULL_LB, loop=P),
        "auto_host": dict(mode=_lib.MODE_AUTO, pull=_lib.PULL_LB, loop=H),
        "auto_chain": dict(mode=_lib.MODE_AUTO, pull=_lib.PULL_LB, loop=_lib.LOOP_CHAINED),
        "auto_sc": dict(mode=_lib.MODE_AUTO, pull=_lib.PULL_SC, loop=H),
        "auto_wa": dict(mode=_lib.MODE_AUTO, pull=_lib.PULL_WA, loop=H),
        "pull_sc": dict(mode=_lib.MODE_PULL_SC, loop=H),
        "pull_wa": dict(mode=_lib.MODE_PULL_WA, loop=H),
        "push": dict(mode=_lib.MODE_PUSH, pull=_lib.PULL_LB, loop=P),
        "cooc": dict(mode=_lib.MODE_COOC, loop=H),
    }
    kw = dict(modes[args.mode])
    if args.grid:
        kw["grid"] = args.grid
    if args.block:
        kw["block"] = args.block
    if args.fb_lds:
        kw["fb_lds"] = args.fb_lds
    if args.alpha:
        kw["alpha"] = args.alpha
    if args.beta:
        kw["beta"] = args.beta
    opts = tb.default_opts(**kw)
    # every handle is closed before the interpreter exits (the HIP runtime,
    # and a profiler's library when one is loaded, tear down at exit)
    with tb.Graph(g.CP, g.IC, symmetric=True) as gr, tb.Context(g.n) as ctx:
        line = bench_single(tb, graphs, args, g, sources, gr, ctx, opts)
    if not args.no_er:
        line["er4096"] = bench_er4096(tb, graphs, cpu=not args.no_cpu)
    if not args.no_batch:
        line["batched"] = bench_batched(tb, graphs)
        line["multi_graph"] = bench_multigraph(tb, graphs)
        line["high_diameter"] = bench_grid(tb, graphs, cpu_passes=0 if args.no_cpu else 24)
    if not args.no_serve:
        line["serving"] = bench_serving_process()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(g, sources, args.cpu_budget)
    print(json.dumps(line))


def bench_single(tb, graphs, args, g, sources, gr, ctx, opts):
    """The headline measurement: args.steps single-source BFS on the resident
    R-MAT graph, sources in turn; plus the per-variant table and the drop-in
    entry point's PCIe-inclusive timing."""
    # a plain copy of known size: achievable HBM bandwidth, and the reference
    # dispatch for the FETCH_SIZE / WRITE_SIZE counters of a profiled run
    calib_bytes = args.calib_mib << 20
    copy_ms = ctx.calibrate_copy(calib_bytes, 3) if args.calib_mib > 0 else 0.0

    roots = np.array([sources[i % len(sources)] for i in range(args.steps)], np.int32)
    if args.warmup:
        ctx.run_many(gr, roots[:args.warmup] if args.warmup <= len(roots) else
                     np.resize(roots, args.warmup), opts)

    def timed(o):
        # HIP events on the context's stream (the stream every BFS kernel runs on)
        ctx.mark(0)
        st = ctx.run_many(gr, roots, o)
        ctx.mark(1)
        return st, ctx.elapsed(0, 1)

    stats, total_ms = timed(opts)
    tune_info = None
    if not args.no_tune and args.mode in ("auto_chain",):
        # the autotuner on sources the timed steps do not use (the last 8 of
        # the sample), then the same timed steps again: runs with the
        # tunables at 0 take the tuned values for this graph's class
        tune_roots = np.asarray(sources[-8:], np.int32)
        best, tms, dms = ctx.autotune(gr, tune_roots, opts)
        ctx.run_many(gr, roots[:max(1, args.warmup)], opts)
        stats_t, total_t = timed(opts)
        tune_info = {
            "tuning_sources": len(tune_roots),
            "best": {"alpha": best.alpha, "beta": best.beta, "stage_ratio": best.stage_ratio,
                     "lb_grid": best.lb_grid},
            "tuning_ms_per_bfs": {"best": tms, "defaults": dms},
            "timed_ms_per_step": {"tuned": total_t / args.steps,
                                  "defaults": total_ms / args.steps},
            "gain": total_ms / total_t,
            "value_from": "tuned" if total_t < total_ms else "defaults",
        }
        if total_t < total_ms:
            stats, total_ms = stats_t, total_t
    edges = sum(st["edges"] for st in stats)
    kern_ms = sum(st["ms"] for st in stats)
    levels = [st["levels"] for st in stats]
    gteps = (edges / 2) / (total_ms * 1e-3) / 1e9

    # roofline: algorithmic bytes of one BFS = 4(n+1) + 4 m_s + 12 n
    n = g.n
    b_alg = args.steps * (4 * (n + 1) + 12 * n) + 4 * edges
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": gteps,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 depth + u64 path count",
        "data": "synthetic",
        "config": {
            "workload": f"single-source BFS with exact sigma, R-MAT scale {args.scale} "
                        f"ef16 a.57 b.19 c.19 symmetrized relabelled seed 2",
            "n": n, "nz": g.nz, "sources": len(sources), "source_seed": 2,
            # the timed steps take the sources in turn: fewer steps than
            # sources time only the first ones
            "sources_timed": int(len(set(roots.tolist()))),
            "mode": args.mode, "mean_levels": float(np.mean(levels)),
            "push_levels": float(np.mean([st["push_levels"] for st in stats])),
            "pull_levels": float(np.mean([st["pull_levels"] for st in stats])),
            "bfs_ms_median": float(np.median([st["ms"] for st in stats])),
            # per-source GTEPS from each BFS's own event pair
            "gteps_median": float(np.median([st["edges"] / 2 / (st["ms"] * 1e-3) / 1e9
                                             for st in stats])),
            "gteps_harmonic_mean": float(len(stats) / sum((st["ms"] * 1e-3) * 1e9 /
                                                          (st["edges"] / 2) for st in stats)),
            "in_degree": {k: v for k, v in gr.degree_stats("in").items() if k != "bins"},
        },
        "autotune": tune_info,
        "copy_kernel_gbs": (2 * calib_bytes / (copy_ms * 1e-3) / 1e9) if copy_ms else None,
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.scale, args.mode),
            "kernel": {"auto_chain": "k_chain_level (the level launches of one BFS; "
                                     "time = device time of the BFS incl. its init kernel)",
                       "auto_persistent": "k_persist (one launch = one whole BFS)",
                       "push": "k_persist (one launch = one whole BFS)"}.get(
                           args.mode, "whole BFS (all level kernels of one source)"),
        },
    }
    if not args.no_variants:
        line["variants"] = bench_variants(tb, g, gr, ctx, sources,
                                          names=[x for x in args.variants.split(",") if x])
    if not args.no_dropin:
        line["dropin_pcie_inclusive"] = bench_dropin(tb, g, sources)
    return line


if __name__ == "__main__":
    main()
