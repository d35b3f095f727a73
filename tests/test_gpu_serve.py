"""Serving (tbfs_server_*): resident graphs, queries batched as they arrive.

Every query's S / sigma / overflow flags must equal the oracle's single-source
BFS from its root, and must be the same whether the query ran alone or in a
mix of concurrent queries on several graphs (tolerance 0).
"""
import threading
import time

import numpy as np
import pytest

import oracle
import turbobfs_amd as tb
from turbobfs_amd import graphs, mtx
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _graph_set():
    gs = [
        (graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1), True),
        (graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=2, relabel=True), True),
        (graphs.grid(40, 30, 0.1, seed=3, relabel=True), True),
        (graphs.grid(48, 48, 0.0, seed=0, relabel=False), True),   # counts past 2^64
        (mtx.read_csc(golden("Tina_AskCal.mtx")), False),          # directed
    ]
    return [g for g, _ in gs], [tb.Graph(g.CP, g.IC, symmetric=s) for g, s in gs]


def _submit(srv, g, gi, root, store):
    S = tb.host_array(g.n, np.int32)
    sig = tb.host_array(g.n, np.uint64)
    ovf = tb.host_array(g.n, np.uint8)
    t = srv.submit(gi, root, S, sig, ovf)
    store[t] = (gi, root, S, sig, ovf)
    return t


def _check(cs, store, done):
    for d in done:
        assert d["status"] == 0, d
        gi, root, S, sig, ovf = store[d["ticket"]]
        eS, esig, eovf, _ = oracle.textbook(cs[gi].CP, cs[gi].IC, root)
        assert np.array_equal(S, eS), (gi, root)
        assert np.array_equal(sig, esig), (gi, root)
        assert np.array_equal(ovf, eovf), (gi, root)
        assert d["submit_us"] <= d["start_us"] <= d["done_us"]


def test_serve_mixed_stream_exact_and_batched(gpu):
    cs, gs = _graph_set()
    srv = tb.Server(gs, mem_budget=1 << 30, slots=3)
    rng = np.random.default_rng(5)
    store = {}
    print("[serve] mixed stream of 600 queries", flush=True)

    def producer():
        for i in range(600):
            gi = int(rng.integers(0, len(cs)))
            _submit(srv, cs[gi], gi, int(rng.integers(0, cs[gi].n)), store)
            if i % 50 == 0:
                time.sleep(0.002)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    srv.drain()
    done = srv.poll(1000)
    assert len(done) == 600
    _check(cs, store, done)
    assert max(d["batch"] for d in done) > 1          # queries were batched
    info = srv.info()
    assert info["batched"] == 600 and info["batches"] < 600
    srv.close()


def test_serve_alone_equals_in_mix(gpu):
    cs, gs = _graph_set()
    probe = [(1, 5), (3, 0), (4, 10), (0, 77)]
    alone = {}
    srv = tb.Server(gs, mem_budget=1 << 30, slots=2)
    for gi, r in probe:
        store = {}
        t = _submit(srv, cs[gi], gi, r, store)
        srv.drain()
        (d,) = srv.poll(10)
        assert d["ticket"] == t and d["batch"] == 1
        alone[(gi, r)] = store[t][2:]
    store = {}
    rng = np.random.default_rng(9)
    tickets = {}
    for i in range(300):
        if i % 60 == 0:
            gi, r = probe[(i // 60) % len(probe)]
        else:
            gi = int(rng.integers(0, len(cs)))
            r = int(rng.integers(0, cs[gi].n))
        tickets[_submit(srv, cs[gi], gi, r, store)] = (gi, r)
    srv.drain()
    done = srv.poll(1000)
    _check(cs, store, done)
    seen = 0
    for d in done:
        key = tickets[d["ticket"]]
        if key in alone and d["batch"] > 1:
            for a, b in zip(alone[key], store[d["ticket"]][2:]):
                assert np.array_equal(a, b)
            seen += 1
    assert seen >= 1
    srv.close()


def test_serve_small_budget_limits_batches(gpu):
    cs, gs = _graph_set()
    # room for two 64-source words of the largest graph per slot
    srv = tb.Server(gs, mem_budget=tb.server_budget_for(gs, 2, 2), slots=2)
    assert srv.info(0)["batch_cap"] == 128
    store = {}
    rng = np.random.default_rng(3)
    for _ in range(200):
        _submit(srv, cs[0], 0, int(rng.integers(0, cs[0].n)), store)
    srv.drain()
    done = srv.poll(1000)
    assert len(done) == 200 and max(d["batch"] for d in done) <= 128
    _check(cs, store, done)
    srv.close()


def test_serve_invalid(gpu):
    cs, gs = _graph_set()
    srv = tb.Server(gs[:2], mem_budget=1 << 28, slots=1)
    with pytest.raises(tb.TbfsError):
        srv.submit(2, 0)
    with pytest.raises(tb.TbfsError):
        srv.submit(0, cs[0].n)
    t = srv.submit(0, 3)   # no host buffers: runs, copies nothing
    srv.drain()
    (d,) = srv.poll(4)
    assert d["ticket"] == t and d["status"] == 0
    srv.close()
    with pytest.raises(tb.TbfsError):
        tb.Server(gs, mem_budget=1000)   # not even one word of the largest graph


def _stream(srv, cs, plan, gap_s, store):
    """Submit (graph, root) pairs `gap_s` seconds apart."""
    t0 = time.perf_counter()
    for i, (gi, r) in enumerate(plan):
        while time.perf_counter() - t0 < i * gap_s:
            pass
        _submit(srv, cs[gi], gi, r, store)


def test_serve_continuous_batch_admits_mid_batch(gpu):
    """A deep graph (256 x 256 grid, 10% of edges removed: ~500 levels):
    overlapping queries share a continuous batch, and queries arriving while
    it runs are seeded at a later level boundary (level > 0).  Their rows are
    exactly the oracle's, and equal to the same root's run alone."""
    cs = [graphs.grid(256, 256, 0.1, seed=4, relabel=True),
          graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)]
    gs = [tb.Graph(g.CP, g.IC, symmetric=True) for g in cs]
    srv = tb.Server(gs, mem_budget=1 << 30, slots=2)
    rng = np.random.default_rng(11)
    roots = [int(x) for x in rng.integers(0, cs[0].n, size=40)]
    # alone: the first query learns the graph's depth (a batch of one), the
    # next ones run alone in continuous batches
    alone = {}
    for r in roots[:3]:
        store = {}
        t = _submit(srv, cs[0], 0, r, store)
        srv.drain()
        (d,) = srv.poll(10)
        assert d["ticket"] == t and d["status"] == 0
        alone[r] = store[t][2:]
    print("[serve] continuous batch: 40 grid queries 0.4 ms apart, R-MAT between", flush=True)
    plan = []
    for i, r in enumerate(roots):
        plan.append((0, r))
        if i % 4 == 0:
            plan.append((1, int(rng.integers(0, cs[1].n))))
    store = {}
    _stream(srv, cs, plan, 0.0004, store)
    srv.drain()
    done = srv.poll(1000)
    assert len(done) == len(plan)
    _check(cs, store, done)
    mid = [d for d in done if d["graph"] == 0 and d["level"] > 0]
    print(f"[serve] admitted mid-batch: {len(mid)} (levels {sorted(d['level'] for d in mid)[:8]}...)",
          flush=True)
    assert len(mid) >= 5
    for d in done:  # roots 0..2 ran alone first and again in the stream
        gi, r = store[d["ticket"]][:2]
        if gi == 0 and r in alone:
            for a, b in zip(alone[r], store[d["ticket"]][2:]):
                assert np.array_equal(a, b)
    # three queries seeded mid-batch, each run again alone: same arrays
    for d in mid[:3]:
        r = store[d["ticket"]][1]
        one = {}
        t = _submit(srv, cs[0], 0, r, one)
        srv.drain()
        (d1,) = srv.poll(10)
        assert d1["ticket"] == t and d1["status"] == 0 and d1["level"] == 0
        for a, b in zip(one[t][2:], store[d["ticket"]][2:]):
            assert np.array_equal(a, b)
    assert all(d["level"] == 0 for d in done if d["graph"] == 1)
    srv.close()


def test_serve_continuous_batches_on_every_graph(gpu):
    """open_depth=1 makes every graph run as continuous batches after its
    first query: pull levels, the degree-ordered copy of the R-MAT graph,
    directed graphs and counts past 2^64 (48 x 48 grid) all go through
    admission at level boundaries; every row matches the oracle."""
    cs, gs = _graph_set()
    srv = tb.Server(gs, mem_budget=1 << 30, slots=3, open_depth=1)
    rng = np.random.default_rng(21)
    store = {}
    # one query per graph first (learns the depths), then a stream
    for gi in range(len(cs)):
        _submit(srv, cs[gi], gi, 0, store)
    srv.drain()
    first = srv.poll(100)
    _check(cs, store, first)
    store = {}
    plan = []
    for _ in range(300):
        gi = int(rng.integers(0, len(cs)))
        plan.append((gi, int(rng.integers(0, cs[gi].n))))
    print("[serve] continuous batches on every graph: 300 queries", flush=True)
    _stream(srv, cs, plan, 0.0001, store)
    srv.drain()
    done = srv.poll(1000)
    assert len(done) == 300
    _check(cs, store, done)
    levels = [d["level"] for d in done]
    print(f"[serve] seeded mid-batch: {sum(1 for x in levels if x > 0)} of 300", flush=True)
    assert any(x > 0 for x in levels)
    srv.close()


def test_serve_hold_batches_a_group(gpu):
    """While dispatch is held nothing runs; a group submitted meanwhile is
    taken together once released (one batch per graph, up to its cap), every
    result exact; tbfs_server_drain releases a hold."""
    cs, gs = _graph_set()
    srv = tb.Server(gs[:2])
    rng = np.random.default_rng(9)
    store = {}
    srv.hold(True)
    for gi in (0, 1):
        for r in rng.integers(0, cs[gi].n, 40):
            _submit(srv, cs[gi], gi, int(r), store)
    assert srv.poll(64, wait_us=50000) == []  # held: nothing ran
    srv.hold(False)
    done = []
    while len(done) < len(store):
        done += srv.poll(256, wait_us=100000)
    _check(cs, store, done)
    for gi in (0, 1):
        sizes = {d["batch"] for d in done if d["graph"] == gi}
        assert sizes == {40}, (gi, sizes)  # the whole group in one batch
    # a drain while held releases the hold and returns with everything done
    srv.hold(True)
    more = {}
    for r in rng.integers(0, cs[0].n, 5):
        _submit(srv, cs[0], 0, int(r), more)
    srv.drain()
    got = srv.poll(64)
    assert len(got) == 5
    _check(cs, more, got)
    srv.close()
