"""Parity of the HIP path (through the C ABI) with the CPU oracle: every
stage bit-exact (integer work), at sizes the oracle finishes in seconds."""
import numpy as np
import pytest

import pyoracle as O
from gpsm_amd import _lib
from gpsm_amd import graphs as G
from helpers import small_cases
from test_oracle import hand_fixture

pytestmark = pytest.mark.gpu


def bits_to_sets(bm, n):
    out = []
    for row in bm:
        b = np.unpackbits(row.view(np.uint8), bitorder="little")[:n]
        out.append(np.nonzero(b)[0])
    return out


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1023, 1024, 1025, 4097, (1 << 20) + 3, (1 << 24) + 4099])
def test_scan(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1 << 20, size=n).astype(np.uint32)
    out, tot = ctx.scan(a)
    ref, rtot = O.exclusive_scan(a)
    assert np.array_equal(out, ref) and tot == rtot


def test_scan_wraps_like_u32(ctx):
    a = np.full(4096 * 3 + 5, 0xFFFFFFF0, dtype=np.uint32)
    out, tot = ctx.scan(a)
    ref, rtot = O.exclusive_scan(a)
    assert np.array_equal(out, ref) and tot == rtot


@pytest.mark.parametrize("nbits", [1, 31, 32, 33, 255, 256, 257, 8192, 8193, (1 << 20) + 3])
@pytest.mark.parametrize("density", [0.0, 0.01, 0.5, 1.0])
def test_compact(ctx, nbits, density):
    rng = np.random.default_rng(nbits)
    words = np.zeros((nbits + 31) // 32, dtype=np.uint32)
    on = np.nonzero(rng.random(nbits) < density)[0] if density < 1 else np.arange(nbits)
    for i in on:
        words[i >> 5] |= np.uint32(1 << (i & 31))
    words[-1] |= np.uint32(0xFFFFFFFF) << np.uint32(nbits & 31) if nbits & 31 else np.uint32(0)  # garbage past nbits
    assert np.array_equal(ctx.compact(words, nbits), O.compact_bitmap(words, nbits))


@pytest.mark.parametrize("density", [0.01, 0.5])
def test_compact_many_segments(ctx, density):
    """A bitmap of 260 segments of 2048 words (the per-row scan of segment
    counts takes two steps), with set bits past nbits that must be ignored."""
    nbits = (1 << 24) + 3 * 65536 + 37
    rng = np.random.default_rng(5)
    by = np.packbits(rng.random(nbits) < density, bitorder="little")
    words = np.concatenate([by, np.zeros((-len(by)) % 4, dtype=np.uint8)]).view(np.uint32).copy()
    assert len(words) == (nbits + 31) // 32 and (len(words) + 2047) // 2048 == 260
    words[-1] |= np.uint32(0xFFFFFFFF) << np.uint32(nbits & 31)
    assert np.array_equal(ctx.compact(words, nbits), O.compact_bitmap(words, nbits))


@pytest.mark.parametrize("name", ["hand_aba.json", "hand_refine.json", "hand_explore_prefix.json"])
def test_hand_fixture(ctx, name):
    fx, g, q = hand_fixture(name)
    rg = ctx.upload(g)
    bm, cnt = rg.filter(q)
    sets = bits_to_sets(bm, g.n)
    for u, ids in enumerate(fx["cand_refined"]):
        assert sets[u].tolist() == ids
    edges, tables = rg.candidate_edges(q)
    assert edges == [tuple(e) for e in fx["edges"]]
    for (ids, off, lst), t in zip(tables, fx["cet"]):
        assert [ids.tolist(), off.tolist(), lst.tolist()] == t
    r = rg.match(q, want_tuples=True, tuple_cap=100)
    assert r.order == fx["order"]
    assert r.level_rows[1:] == fx["step_rows_injective"]
    assert r.tuples.tolist() == fx["embeddings"]


@pytest.mark.parametrize("case", range(16))
def test_filter_matches_oracle(ctx, case):
    g, q = small_cases(16, seed=31)[case]
    st, init, ref, fo = O.o2_filter(g, q)
    rg = ctx.upload(g)
    bm, cnt = rg.filter(q)
    if st != O.OK:
        # the reference stops at the first empty set; the device finishes the
        # schedule, so only the emptiness itself is compared
        assert (cnt == 0).any()
        return
    sets = bits_to_sets(bm, g.n)
    for u in range(q.n):
        assert np.array_equal(sets[u], np.nonzero(ref[u])[0]), u
        assert cnt[u] == ref[u].sum()


@pytest.mark.parametrize("case", range(12))
def test_candidate_edges_match_oracle(ctx, case):
    g, q = small_cases(12, seed=41)[case]
    ref = O.o2_cet(g, q)
    rg = ctx.upload(g)
    if ref is None:
        return
    edges, tables = rg.candidate_edges(q)
    assert edges == ref[0]
    for (a, b, c), (x, y, z) in zip(tables, ref[1]):
        assert np.array_equal(a, x) and np.array_equal(b, y) and np.array_equal(c, z)


@pytest.mark.parametrize("case", range(24))
def test_match_matches_oracle(ctx, case):
    g, q = small_cases(24, seed=51)[case]
    r2 = O.o2_match(g, q, mode=0)
    r1 = O.o1_enumerate(g, q)
    rg = ctx.upload(g)
    r = rg.match(q)
    assert (r.count, r.hash) == (r1["count"], r1["hash"]) == (r2["count"], r2["hash"])
    if r2["status"] == O.OK and r2["count"] > 0:
        assert r.order == r2["order"]
        assert r.level_rows[1:] == r2["level_rows"][1:q.n]
        assert r.cand == r2["cand"]
        assert r.join_work == r2["join_work"] == sum(r.level_work)
        assert r.level_work[1:] == r2["level_work"][1:q.n]


def test_er10k_walk6_full_tuple_list(ctx):
    """ER 10k/40k, 8 labels, 6-vertex walk query: the sorted
    int32 tuple list is identical to the independent enumerator's."""
    g, (q,) = G.er10k_walk6()
    r1 = O.o1_enumerate(g, q, tuple_cap=10 ** 6)
    r2 = O.o2_match(g, q, mode=0, tuple_cap=10 ** 6)
    r = ctx.upload(g).match(q, want_tuples=True, tuple_cap=10 ** 6)
    assert r.count == r1["count"] == r2["count"] >= 1
    assert r.tuples.tobytes() == r1["tuples"].tobytes() == r2["tuples"].tobytes()
    assert r.hash == r1["hash"]


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_closed_forms(ctx, k):
    import math
    rg = ctx.upload(G.complete(8))
    assert rg.match(G.clique_query(k)).count == math.perm(8, k)
    rc = ctx.upload(G.cycle(11))
    assert rc.match(G.path_query(k + 1)).count == 22


def test_absent_label_and_empty_results(ctx):
    g = G.small_graph(3, n=200, m=600, labels=2)
    rg = ctx.upload(g)
    q = G.path_query(3, label=7)  # label absent from the data graph
    r = rg.match(q)
    assert r.count == 0 and r.hash == 0
    # bipartite data graph, odd cycle query
    a = np.repeat(np.arange(5), 5)
    b = np.tile(np.arange(5, 10), 5)
    gb = G.from_edges(10, a, b, np.ones(10))
    assert ctx.upload(gb).match(G.cycle_query(5)).count == 0
    assert ctx.upload(gb).match(G.cycle_query(4)).count == 5 * 4 * 5 * 4 * 2


def test_invalid_queries_are_rejected(ctx):
    g = G.small_graph(3, n=200, m=600, labels=2)
    rg = ctx.upload(g)
    disconnected = G.Query(np.ones(4, dtype=np.uint32), np.array([0, 2], dtype=np.uint32),
                           np.array([1, 3], dtype=np.uint32))
    with pytest.raises(_lib.GpsmError) as e:
        rg.match(disconnected)
    assert e.value.code == -7
    big = G.path_query(65)
    with pytest.raises(_lib.GpsmError) as e:
        rg.match(big)
    assert e.value.code == -4


def test_batch_equals_single_and_oracle(ctx):
    g = G.small_graph(9, n=3000, m=12000, labels=3, kind="cl")
    qs = G.queries(g, 24, 3, 7, _lib.Q_WALK_HALF, 13, 10 ** 7)
    rg = ctx.upload(g)
    batch = rg.match_batch(qs)
    graph = rg.match_batch(qs, use_graph=True)
    for q, b, gr in zip(qs, batch, graph):
        s = rg.match(q)
        r1 = O.o1_enumerate(g, q)
        assert (b.count, b.hash) == (s.count, s.hash) == (r1["count"], r1["hash"]) == (gr.count, gr.hash)
        assert b.level_rows == s.level_rows
        assert b.join_work == s.join_work == gr.join_work
        assert b.level_work == s.level_work == gr.level_work


def test_cached_graphs_replay_and_update(ctx):
    """The captured phases run through the context's cached executable
    graphs: the same batch again relaunches them without a capture, eager
    batches and synchronisations in between change nothing, a batch of
    another shape is captured again (an in-place update or a new
    instantiation), and a batch of the same shapes with other labels (whose
    launches may have the same arguments as the last one's, reading other
    data) gives its own results; every run equals the eager result."""
    g = G.small_graph(21, n=4000, m=20000, labels=3, kind="cl")
    rg = ctx.upload(g)
    qa = G.queries(g, 16, 3, 7, _lib.Q_WALK_HALF, 31, 10 ** 7)
    qb = G.queries(g, 9, 4, 9, _lib.Q_WALK_INDUCED, 32, 10 ** 7)
    # qa's shapes with every label moved to the next one of the graph's three
    qc = [G.Query(labels=np.array([l % 3 + 1 for l in q.labels], dtype=np.uint32), src=q.src.copy(),
                  dst=q.dst.copy()) for q in qa]
    want_a = [(r.count, r.hash, r.join_work) for r in rg.match_batch(qa)]
    want_b = [(r.count, r.hash, r.join_work) for r in rg.match_batch(qb)]
    want_c = [(r.count, r.hash, r.join_work) for r in rg.match_batch(qc)]
    i0, u0, r0 = ctx.graph_cache_stats()
    for rep in range(4):
        got = [(r.count, r.hash, r.join_work) for r in rg.match_batch(qa, use_graph=True)]
        assert got == want_a, rep
        assert [(r.count, r.hash, r.join_work) for r in rg.match_batch(qa)] == want_a  # eager between replays
    i1, u1, r1 = ctx.graph_cache_stats()
    # two phases: captured once (instantiated or updated), then relaunched
    assert i1 - i0 + u1 - u0 == 2 and r1 - r0 == 6
    for rep in range(2):
        assert [(r.count, r.hash, r.join_work) for r in rg.match_batch(qb, use_graph=True)] == want_b
    i2, u2, r2 = ctx.graph_cache_stats()
    assert i2 - i1 + u2 - u1 >= 2 and r2 - r1 >= 2
    assert [(r.count, r.hash, r.join_work) for r in rg.match_batch(qa, use_graph=True)] == want_a
    _, _, r3 = ctx.graph_cache_stats()
    for rep in range(2):
        assert [(r.count, r.hash, r.join_work) for r in rg.match_batch(qc, use_graph=True)] == want_c, rep
    _, _, r4 = ctx.graph_cache_stats()
    assert r4 - r3 >= 2  # at least the second run relaunches both phases


def test_wide_tree_and_cyclic_queries_share_the_level_budget(ctx):
    """64-vertex paths (tree plans) and cycles (cyclic plans) in one batch on
    a 64-vertex cycle: the two join launches run side by side, and at 2048
    rows per level their level buffers (16.5 MB per workgroup at 64 query
    vertices, 2048 workgroups' worth of tasks) exceed the 16 GiB budget, so
    both launches drop workgroups in proportion. Closed forms: a path or the
    whole cycle has 2n placements, a shorter cycle none; every hash equals
    the enumerator's, and the cut launches equal the batch without it."""
    g = G.cycle(64)
    rg = ctx.upload(g)
    qs = [G.path_query(64)] * 8 + [G.cycle_query(64)] * 8 + [G.path_query(33), G.cycle_query(32)]
    res = rg.match_batch(qs, join_cap=2048)
    assert [r.count for r in res] == [128] * 17 + [0]
    want = {}
    for q, r in zip(qs, res):
        key = (len(q.labels), len(q.src))
        if key not in want:
            o = O.o1_enumerate(g, q)
            want[key] = (o["count"], o["hash"])
        assert (r.count, r.hash) == want[key]
    small = rg.match_batch(qs[16:], join_cap=2048)
    assert [(r.count, r.hash) for r in small] == [(r.count, r.hash) for r in res[16:]]
    assert [(r.count, r.hash) for r in rg.match_batch(qs, join_cap=2048, use_graph=True)] == \
        [(r.count, r.hash) for r in res]


def test_small_join_buffers_force_depth_first_spills(ctx):
    """A join buffer of 512 rows forces many partial windows and re-entries."""
    g = G.small_graph(10, n=2000, m=16000, labels=1, kind="er")
    rg = ctx.upload(g)
    for q in G.queries(g, 4, 4, 6, _lib.Q_WALK_TREE, 3):
        r1 = O.o1_enumerate(g, q)
        r = rg.match(q, join_cap=512)
        assert (r.count, r.hash) == (r1["count"], r1["hash"])
        # windows cut short and reloaded are not counted twice
        full = rg.match(q)
        assert r.join_work == full.join_work and r.level_rows == full.level_rows
        assert r.level_work == full.level_work


def test_explicit_orders_agree(ctx):
    g, q = small_cases(6, seed=61)[4]
    rg = ctx.upload(g)
    base = rg.match(q)
    rng = np.random.default_rng(0)
    adj = [set() for _ in range(q.n)]
    for a, b in zip(q.src, q.dst):
        adj[a].add(int(b))
        adj[b].add(int(a))
    for _ in range(5):
        start = int(rng.integers(q.n))
        order, seen = [start], {start}
        while len(order) < q.n:
            w = int(rng.choice(sorted({x for v in order for x in adj[v]} - seen)))
            order.append(w)
            seen.add(w)
        r = rg.match(q, order_mode=_lib.ORDER_EXPLICIT, explicit_order=order)
        r2 = O.o2_match(g, q, mode=0, order=order)
        assert (r.count, r.hash) == (base.count, base.hash) == (r2["count"], r2["hash"])
        if r2["count"]:
            assert r.level_rows[1:] == r2["level_rows"][1:q.n]
            assert r.join_work == r2["join_work"]
            assert r.level_work[1:] == r2["level_work"][1:q.n]


def test_profile_modes_keep_results_and_count_bytes(ctx):
    """Timing-only and counting profile modes time every launch family and
    leave results unchanged; only the counting mode adds the filter's
    measured work to its bytes.  The join-only mode times the join alone."""
    g = G.small_graph(12, n=3000, m=12000, labels=3, kind="cl")
    qs = G.queries(g, 16, 3, 6, _lib.Q_WALK_HALF, 19, 10 ** 7)
    rg = ctx.upload(g)
    base = rg.match_batch(qs)
    stats = {}
    for mode in (1, 2, 3):
        ctx.profile(mode)
        res = rg.match_batch(qs)
        stats[mode] = ctx.kernel_stats()
        assert [(r.count, r.hash) for r in res] == [(r.count, r.hash) for r in base]
    ctx.profile(0)
    for k in ("filter_step", "join", "cet_count", "cet_write"):
        assert stats[1][k]["launches"] == stats[2][k]["launches"] > 0 and stats[1][k]["ms"] > 0
    assert stats[2]["filter_step"]["bytes"] > stats[1]["filter_step"]["bytes"] > 0
    timed = {k for k, v in stats[3].items() if not k.startswith("wall:") and v["launches"]}
    assert timed == {"join"} and stats[3]["join"]["launches"] == stats[1]["join"]["launches"]


def test_work_budget_is_exact(ctx):
    """A budget equal to a query's join work lets it finish with the same
    result; one less abandons it (GPSM_EBUDGET, no partial count), alone
    and next to other queries in a batch."""
    g = G.small_graph(14, n=3000, m=15000, labels=2, kind="cl")
    qs = [q for q in G.queries(g, 12, 4, 6, _lib.Q_WALK_HALF, 23, 10 ** 7) if q.n > 2]
    rg = ctx.upload(g)
    base = rg.match_batch(qs)
    i = max(range(len(qs)), key=lambda j: base[j].join_work)
    q, b = qs[i], base[i]
    assert b.join_work > 1000
    ok = rg.match(q, work_budget=b.join_work)
    assert (ok.status, ok.count, ok.hash, ok.join_work) == (0, b.count, b.hash, b.join_work)
    with pytest.raises(_lib.GpsmError) as err:  # the single-query call returns the status
        rg.match(q, work_budget=b.join_work - 1)
    assert err.value.code == _lib.EBUDGET
    budgets = [r.join_work for r in base]
    res = rg.match_batch(qs, work_budget=b.join_work - 1)
    for r, r0, w in zip(res, base, budgets):
        if w <= b.join_work - 1:
            assert (r.status, r.count, r.hash) == (0, r0.count, r0.hash)
        else:
            assert (r.status, r.count, r.hash) == (_lib.EBUDGET, 0, 0)
    assert res[i].status == _lib.EBUDGET


def test_hub_bitmaps_give_the_same_results(ctx):
    """The join's edge tests through hub adjacency bitmaps (here every vertex
    of degree > 8 is a hub) give the same counts, hashes and work as binary
    searches only, and as the enumerator, on cliques and cycles (every
    vertex after the first edge has back edges) over a power-law graph."""
    g = G.small_graph(15, n=3000, m=24000, labels=2, kind="cl")
    qs = G.queries(g, 6, 4, 5, _lib.Q_CLIQUE, 29, 0) + G.queries(g, 6, 4, 4, _lib.Q_CYCLE, 30, 0)
    try:
        ctx.set_hub_policy(8, 1 << 30)
        rh = ctx.upload(g)
        ctx.set_hub_policy(512, 0)
        rn = ctx.upload(g)
    finally:
        ctx.set_hub_policy()  # defaults
    nh, d = rh.hubs()
    assert nh > 100 and d == 8 and rn.hubs() == (0, 0)
    assert rh.device_bytes > rn.device_bytes
    a = rh.match_batch(qs)
    b = rn.match_batch(qs)
    for q, x, y in zip(qs, a, b):
        r1 = O.o1_enumerate(g, q)
        assert (x.count, x.hash) == (y.count, y.hash) == (r1["count"], r1["hash"])
        assert x.level_work == y.level_work and x.level_rows == y.level_rows
    assert sum(x.count for x in a) > 0
    # the doubling rule: a cap below the bitmaps of degree > 8 raises the degree
    try:
        ctx.set_hub_policy(8, nh * ((g.n + 31) // 32) * 4 - 1)
        rc = ctx.upload(g)
    finally:
        ctx.set_hub_policy()
    nh2, d2 = rc.hubs()
    assert d2 >= 16 and nh2 < nh and nh2 * ((g.n + 31) // 32) * 4 <= nh * ((g.n + 31) // 32) * 4 - 1


def test_device_csr_build_equals_host_build(ctx):
    """gpsm_graph_upload_edges: an edge list with both orientations,
    duplicates and self-loops gives the host builder's CSR, and the same
    match results; an endpoint out of range is refused."""
    rng = np.random.default_rng(3)
    n, m = 4000, 30000
    a = rng.integers(0, n, m).astype(np.uint32)
    b = (a + rng.zipf(1.6, m) % n).astype(np.uint32) % n  # skewed, with self-loops (offset 0 mod n)
    a = np.concatenate([a, b[:5000]]).astype(np.uint32)   # reversed duplicates
    b = np.concatenate([b, a[:5000]]).astype(np.uint32)
    labels = (rng.integers(0, 3, n) + 1).astype(np.uint32)
    host = G.from_edges(n, a, b, labels)
    assert (a == b).any()
    dev = ctx.upload_edges(n, a, b, labels)
    assert np.array_equal(dev.g.row_offset, host.row_offset) and np.array_equal(dev.g.col, host.col)
    rh = ctx.upload(host)
    qs = G.queries(host, 12, 3, 6, _lib.Q_WALK_HALF, 31, 10 ** 7)
    x = rh.match_batch(qs)
    y = dev.match_batch(qs)
    assert [(r.count, r.hash, r.join_work) for r in x] == [(r.count, r.hash, r.join_work) for r in y]
    assert sum(r.count for r in x) > 0
    with pytest.raises(_lib.GpsmError):
        ctx.upload_edges(n, np.array([0, n], dtype=np.uint32), np.array([1, 2], dtype=np.uint32), labels)
    empty = ctx.upload_edges(5, np.zeros(0, dtype=np.uint32), np.zeros(0, dtype=np.uint32),
                             np.ones(5, dtype=np.uint32))
    assert empty.g.adj == 0 and np.array_equal(empty.g.row_offset, np.zeros(6, dtype=np.uint32))


def test_cli_counts_from_text_and_from_cache_agree(tmp_path):
    """GpSM --count on a text data graph and on its binary cache prints the
    same counts and hashes, equal to the enumerator's."""
    import os
    import subprocess
    g = G.small_graph(16, n=2500, m=10000, labels=3, kind="cl")
    qs = G.queries(g, 4, 3, 5, _lib.Q_WALK_HALF, 37, 10 ** 6)
    data, query, cache = tmp_path / "d.g", tmp_path / "q.g", tmp_path / "d.csr"
    src, dst = g.edges()
    _lib.lib().gpsm_io_write_graph(str(data).encode(), 0, 0, g.n, _lib.ptr(g.labels), len(src), _lib.ptr(src),
                                   _lib.ptr(dst))
    _lib.lib().gpsm_io_write_end(str(data).encode())
    for i, q in enumerate(qs):
        q.struct()
        _lib.lib().gpsm_io_write_graph(str(query).encode(), i, i, q.n, _lib.ptr(q.labels), q.m, _lib.ptr(q.src),
                                       _lib.ptr(q.dst))
    _lib.lib().gpsm_io_write_end(str(query).encode())
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "GpSM")
    assert subprocess.run([exe, "--write-cache", str(data), str(cache)], timeout=60).returncode == 0
    outs = []
    for d in (data, cache):
        r = subprocess.run([exe, "--count", str(d), str(query), str(tmp_path / "ans.txt")], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append([l for l in r.stdout.splitlines() if l.startswith("query ")])
    assert outs[0] == outs[1] and len(outs[0]) == len(qs)
    for line, q in zip(outs[0], qs):
        ref = O.o1_enumerate(g, q)
        assert line.endswith("%d embeddings, hash %016x" % (ref["count"], ref["hash"])), line


def test_cli_matches_every_data_graph_of_a_file(tmp_path):
    """A data file of several graphs (the reference's main loop,
    main/run.cpp:60-98): the driver reads the next graph on another thread
    while the current one is matched, and frees each graph's device copy
    before the next can reuse its address; every graph's counts equal the
    enumerator's."""
    import os
    import subprocess
    graphs = [G.small_graph(20 + i, n=1500 + 500 * i, m=6000 + 1000 * i, labels=3, kind="cl") for i in range(3)]
    data, query = tmp_path / "d.g", tmp_path / "q.g"
    for i, g in enumerate(graphs):
        src, dst = g.edges()
        _lib.lib().gpsm_io_write_graph(str(data).encode(), i, i, g.n, _lib.ptr(g.labels), len(src),
                                       _lib.ptr(src), _lib.ptr(dst))
    _lib.lib().gpsm_io_write_end(str(data).encode())
    qs = G.queries(graphs[0], 3, 3, 4, _lib.Q_WALK_HALF, 41, 10 ** 6)
    for i, q in enumerate(qs):
        q.struct()
        _lib.lib().gpsm_io_write_graph(str(query).encode(), i, i, q.n, _lib.ptr(q.labels), q.m, _lib.ptr(q.src),
                                       _lib.ptr(q.dst))
    _lib.lib().gpsm_io_write_end(str(query).encode())
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "GpSM")
    r = subprocess.run([exe, "--count", str(data), str(query), str(tmp_path / "ans.txt")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("query ")]
    assert len(lines) == len(graphs) * len(qs)
    for gi, g in enumerate(graphs):
        for qi, q in enumerate(qs):
            ref = O.o1_enumerate(g, q)
            assert lines[gi * len(qs) + qi] == "query %d: %d embeddings, hash %016x" % (qi, ref["count"], ref["hash"])
