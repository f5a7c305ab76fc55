"""The CPU oracle (oracle/) pinned against closed forms, brute force, a
hand-worked fixture and library routines.  CPU only."""
import itertools
import json
import math
import os

import numpy as np
import pytest

import pyoracle as O
from gpsm_amd import _lib
from gpsm_amd import graphs as G
from helpers import small_cases, sorted_rows

HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1


def brute(g, q):
    """All injective label- and edge-preserving maps, by enumeration."""
    adj = set()
    for v in range(g.n):
        for w in g.col[g.row_offset[v]:g.row_offset[v + 1]]:
            adj.add((v, int(w)))
    rows = []
    for f in itertools.permutations(range(g.n), q.n):
        if any(g.labels[f[i]] != q.labels[i] for i in range(q.n)):
            continue
        if all((f[a], f[b]) in adj for a, b in zip(q.src, q.dst)):
            rows.append(f)
    h = 0
    for r in rows:
        h = (h + O.tuple_hash(r)) & MASK
    return len(rows), h, np.array(sorted(rows), dtype=np.uint32).reshape(-1, q.n)


def both(g, q):
    r2 = O.o2_match(g, q, mode=0)
    r2h = O.o2_match(g, q, mode=1)
    r1 = O.o1_enumerate(g, q)
    return r1, r2, r2h


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6])
def test_clique_in_complete(k):
    g = G.complete(7)
    for r in both(g, G.clique_query(k)):
        assert r["count"] == math.perm(7, k)


@pytest.mark.parametrize("k", [2, 3, 5, 8])
def test_path_in_cycle(k):
    g = G.cycle(9)
    for r in both(g, G.path_query(k)):
        assert r["count"] == 18


@pytest.mark.parametrize("k", [3, 4, 7])
def test_cycle_in_cycle(k):
    g = G.cycle(k)
    for r in both(g, G.cycle_query(k)):
        assert r["count"] == 2 * k


def test_star_counts():
    g = G.small_graph(5, n=60, m=200, labels=1)
    deg = np.diff(g.row_offset.astype(np.int64))
    for leaves in (1, 2, 3):
        expect = int(sum(math.perm(int(d), leaves) for d in deg))
        for r in both(g, G.star_query(leaves)):
            assert r["count"] == expect


def test_odd_cycle_in_bipartite_is_empty():
    a = np.repeat(np.arange(6), 6)
    b = np.tile(np.arange(6, 12), 6)
    g = G.from_edges(12, a, b, np.ones(12))
    for r in both(g, G.cycle_query(5)):
        assert r["count"] == 0
    for r in both(g, G.cycle_query(4)):
        assert r["count"] == 6 * 5 * 6 * 5 * 2  # ordered 4-cycles of K_{6,6}, both orientations


def test_labelled_product():
    # triangle with labels 1,2,3 in K_n with label classes of size a,b,c: a*b*c maps
    n = 9
    labels = np.array([1, 1, 2, 2, 2, 3, 3, 3, 3])
    a, b = np.triu_indices(n, 1)
    g = G.from_edges(n, a, b, labels)
    q = G.Query(np.array([1, 2, 3], dtype=np.uint32), np.array([0, 1, 0], dtype=np.uint32),
                np.array([1, 2, 2], dtype=np.uint32))
    for r in both(g, q):
        assert r["count"] == 2 * 3 * 4


@pytest.mark.parametrize("case", range(30))
def test_brute_force_tiny(case):
    rng = np.random.default_rng(case)
    n = int(rng.integers(5, 8))
    m = int(rng.integers(n, n * (n - 1) // 2 + 1))
    g = G.erdos_renyi(n, m, 100 + case)
    G.set_labels(g, int(rng.integers(1, 3)), 0.0, case)
    q = G.queries(g, 1, 2, min(5, n), [_lib.Q_WALK_TREE, _lib.Q_WALK_INDUCED][case % 2], 7 + case)[0]
    cnt, h, rows = brute(g, q)
    r1 = O.o1_enumerate(g, q, tuple_cap=10000)
    r2 = O.o2_match(g, q, mode=0, tuple_cap=10000)
    r2h = O.o2_match(g, q, mode=1)
    assert (r1["count"], r1["hash"]) == (cnt, h)
    assert (r2["count"], r2["hash"]) == (cnt, h)
    assert (r2h["count"], r2h["hash"]) == (cnt, h)
    assert np.array_equal(r1["tuples"], rows)
    assert np.array_equal(r2["tuples"], rows)


def hand_fixture(name="hand_aba.json"):
    with open(os.path.join(HERE, "golden", name)) as f:
        fx = json.load(f)
    g = G.from_edges(fx["data"]["n"], np.array(fx["data"]["src"]), np.array(fx["data"]["dst"]),
                     np.array(fx["data"]["labels"]))
    q = G.Query(np.array(fx["query"]["labels"], dtype=np.uint32), np.array(fx["query"]["src"], dtype=np.uint32),
                np.array(fx["query"]["dst"], dtype=np.uint32))
    return fx, g, q


@pytest.mark.parametrize("name", ["hand_aba.json", "hand_refine.json", "hand_explore_prefix.json"])
def test_hand_worked_fixture(name):
    """Every intermediate of the reference's steps, worked by hand: a case
    where refine changes nothing, and one where explore and refine remove
    candidates and the join order skips a larger first table."""
    fx, g, q = hand_fixture(name)
    st, init, ref, fo = O.o2_filter(g, q)
    assert st == O.OK
    assert fo == fx["filter_order"]
    for u, ids in enumerate(fx["cand_init"]):
        assert list(np.nonzero(init[u])[0]) == ids
    for u, ids in enumerate(fx["cand_refined"]):
        assert list(np.nonzero(ref[u])[0]) == ids
    edges, tables = O.o2_cet(g, q)
    assert edges == [tuple(e) for e in fx["edges"]]
    for (ids, off, lst), t in zip(tables, fx["cet"]):
        assert list(ids) == t[0] and list(off) == t[1] and list(lst) == t[2]
    r1 = O.o2_match(g, q, mode=1, tuple_cap=100)
    assert r1["step_rows"] == fx["step_rows_homomorphic"]
    r0 = O.o2_match(g, q, mode=0, tuple_cap=100)
    assert r0["step_rows"] == fx["step_rows_injective"]
    assert r0["order"] == fx["order"]
    assert r0["tuples"].tolist() == fx["embeddings"]
    assert r1["tuples"].tolist() == fx["embeddings"]


@pytest.mark.parametrize("case", range(12))
def test_o1_equals_o2_random(case):
    g, q = small_cases(12, seed=3)[case]
    r1 = O.o1_enumerate(g, q, tuple_cap=200000)
    r2 = O.o2_match(g, q, mode=0, tuple_cap=200000)
    assert (r1["count"], r1["hash"]) == (r2["count"], r2["hash"])
    if r1["count"] <= 200000:
        assert np.array_equal(r1["tuples"], r2["tuples"])
    # the set hash is the sum of tuple hashes
    if 0 < r1["count"] <= 2000:
        assert O.set_hash(r1["tuples"]) == r1["hash"]


@pytest.mark.parametrize("case", range(8))
def test_filter_invariants(case):
    """Soundness (every embedding image stays a candidate), monotone refine,
    CET = {(mu, mv) in E : mu in C(u), mv in C(v)} grouped and sorted."""
    g, q = small_cases(8, seed=11)[case]
    st, init, ref, fo = O.o2_filter(g, q)
    r = O.o1_enumerate(g, q, tuple_cap=100000)
    if st == O.OK:
        assert np.all(ref <= init)
        for t in r["tuples"]:
            for u, v in enumerate(t):
                assert ref[u, v]
        edges, tables = O.o2_cet(g, q)
        for (u, v), (ids, off, lst) in zip(edges, tables):
            assert list(ids) == list(np.nonzero(ref[u])[0])
            for j, mu in enumerate(ids):
                nb = g.col[g.row_offset[mu]:g.row_offset[mu + 1]]
                assert list(lst[off[j]:off[j + 1]]) == [w for w in nb if ref[v, w]]
        # |CET(u,v)| = |CET(v,u)|
        for (u, v), (ids, off, lst) in zip(edges, tables):
            rev = sum(int(ref[u, w]) for mv in np.nonzero(ref[v])[0]
                      for w in g.col[g.row_offset[mv]:g.row_offset[mv + 1]])
            assert rev == len(lst)
    else:
        assert r["count"] == 0


@pytest.mark.parametrize("case", range(6))
def test_order_independence(case):
    """Any valid vertex order yields the same set (the join order never
    changes the result)."""
    g, q = small_cases(6, seed=21)[case]
    base = O.o2_match(g, q, mode=0)
    rng = np.random.default_rng(case)
    adj = [set() for _ in range(q.n)]
    for a, b in zip(q.src, q.dst):
        adj[a].add(int(b))
        adj[b].add(int(a))
    for _ in range(4):
        start = int(rng.integers(q.n))
        order, seen = [start], {start}
        while len(order) < q.n:
            front = sorted({w for v in order for w in adj[v]} - seen)
            w = int(rng.choice(front))
            order.append(w)
            seen.add(w)
        r = O.o2_match(g, q, mode=0, order=order)
        assert (r["count"], r["hash"]) == (base["count"], base["hash"])


def join_work_by_levels(g, q, cand, order, edges, size):
    """First-table rows plus, for every later position k, the candidates the
    expansion visits: |N(r[parent]) & C(w)| summed over the rows of level
    k-1, where rows are the injective partial maps of order[:k] inside the
    candidate sets with every query edge among them present.  The parent of
    w is its smallest-|CET| edge (lowest index on ties) to an earlier vertex."""
    nbr = [set(int(x) for x in g.col[g.row_offset[v]:g.row_offset[v + 1]]) for v in range(g.n)]
    pos = {v: k for k, v in enumerate(order)}
    a, b = order[0], order[1]
    rows = [(x, y) for x in cand[a] for y in nbr[x] if y in cand[b]]
    levels = [0, len(rows)]
    for k in range(2, q.n):
        w = order[k]
        mine = [i for i, (u, v) in enumerate(edges) if (u == w and pos[v] < k) or (v == w and pos[u] < k)]
        pe = min(mine, key=lambda i: (size[i], i))
        par = pos[edges[pe][0] if edges[pe][1] == w else edges[pe][1]]
        backs = [pos[u if v == w else v] for i, (u, v) in enumerate(edges) if i in mine and i != pe]
        nxt = []
        levels.append(0)
        for r in rows:
            ext = [x for x in nbr[r[par]] if x in cand[w]]
            levels[k] += len(ext)
            nxt += [r + (x,) for x in ext if x not in r and all(x in nbr[r[p]] for p in backs)]
        rows = nxt
    return levels, len(rows)


@pytest.mark.parametrize("case", range(16))
def test_join_work_by_levels(case):
    """O2's join work (first-table rows + list entries its expand steps
    visit) equals a level-by-level enumeration over its own candidate sets,
    for the heuristic order and for random connected orders."""
    g, q = small_cases(16, seed=71)[case]
    base = O.o2_match(g, q, mode=0)
    if base["status"] != O.OK or base["count"] == 0 or base["count"] > 200000:
        return
    st, _, ref, _ = O.o2_filter(g, q)
    cand = [set(np.nonzero(ref[u])[0].tolist()) for u in range(q.n)]
    edges, size = base["edges"], base["cet_size"]
    orders = [base["order"]]
    rng = np.random.default_rng(case)
    adj = [set() for _ in range(q.n)]
    for a, b in zip(q.src, q.dst):
        adj[a].add(int(b))
        adj[b].add(int(a))
    for _ in range(3):
        start = int(rng.integers(q.n))
        order, seen = [start], {start}
        while len(order) < q.n:
            w = int(rng.choice(sorted({x for v in order for x in adj[v]} - seen)))
            order.append(w)
            seen.add(w)
        orders.append(order)
    for order in orders:
        r = O.o2_match(g, q, mode=0, order=order)
        levels, count = join_work_by_levels(g, q, cand, order, edges, size)
        assert count == r["count"] == base["count"]
        assert r["level_work"][1:] == levels[1:], (order, r["level_work"], levels)
        assert r["join_work"] == sum(levels)


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 1024, 1025, (1 << 20) + 3])
def test_scan_oracle(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1000, size=n).astype(np.uint32)
    out, tot = O.exclusive_scan(a)
    ref = np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]).astype(np.uint32) if n else a
    assert np.array_equal(out, ref)
    assert tot == int(a.astype(np.uint64).sum())


@pytest.mark.parametrize("nbits", [1, 31, 32, 33, 1000, 65536 + 17])
def test_compact_oracle(nbits):
    rng = np.random.default_rng(nbits)
    bits = (rng.random(nbits) < 0.3)
    words = np.zeros((nbits + 31) // 32, dtype=np.uint32)
    for i in np.nonzero(bits)[0]:
        words[i >> 5] |= np.uint32(1 << (i & 31))
    assert np.array_equal(O.compact_bitmap(words, nbits), np.nonzero(bits)[0].astype(np.uint32))


def _params(seed):
    rng = np.random.default_rng(seed)
    F, H = 32, 128
    p = [rng.normal(0, 1 / np.sqrt(F), H * F), rng.normal(0, 0.1, H), rng.normal(0, 1 / np.sqrt(H), H * H),
         rng.normal(0, 0.1, H), rng.normal(0, 1 / np.sqrt(H), H), rng.normal(0, 0.1, 1)]
    return np.concatenate(p).astype(np.float32)


def test_mlp_oracle_finite_differences():
    p = _params(1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(16, 32)).astype(np.float32)
    t = rng.normal(size=16).astype(np.float32)
    loss, g = O.mlp_grad(p, x, t, bf16=False)
    idx = rng.choice(len(p), 40, replace=False)
    for i in idx:
        eps = 1e-2
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = O.mlp_grad(pp, x, t, bf16=False)
        lm, _ = O.mlp_grad(pm, x, t, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_oracle_matches_torch():
    torch = pytest.importorskip("torch")
    p = _params(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(32, 32)).astype(np.float32)
    t = rng.normal(size=32).astype(np.float32)
    F, H = 32, 128
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    o = 0
    W1 = P[o:o + H * F].view(H, F); o += H * F
    b1 = P[o:o + H]; o += H
    W2 = P[o:o + H * H].view(H, H); o += H * H
    b2 = P[o:o + H]; o += H
    W3 = P[o:o + H]; o += H
    b3 = P[o:o + 1]
    X = torch.tensor(x, dtype=torch.float64)
    y = torch.relu(torch.relu(X @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    L = 0.5 * ((y - torch.tensor(t, dtype=torch.float64)) ** 2).mean()
    L.backward()
    loss, g = O.mlp_grad(p, x, t, bf16=False)
    yo = O.mlp_forward(p, x, bf16=False)
    assert np.allclose(yo, y.detach().numpy(), rtol=1e-5, atol=1e-5)
    assert abs(loss - L.item()) < 1e-9
    assert np.allclose(g, P.grad.numpy(), rtol=1e-7, atol=1e-9)


def test_mlp_oracle_bf16_rounding():
    """BF16 mode rounds to nearest even: 1 + 2^-8 (a tie) rounds to 1, and
    1 + 3*2^-9 rounds up to 1 + 2^-7."""
    p = np.zeros(32 * 128 + 128 + 128 * 128 + 128 + 128 + 1, dtype=np.float32)
    # y = W3[0] * relu(W2[0,0] * relu(W1[0,0] * x0)); all weights 1
    p[0] = 1.0
    o = 32 * 128 + 128
    p[o] = 1.0
    o += 128 * 128 + 128
    p[o] = 1.0
    x = np.zeros((2, 32), dtype=np.float32)
    x[0, 0] = 1 + 2.0 ** -8
    x[1, 0] = 1 + 3 * 2.0 ** -9
    y = O.mlp_forward(p, x, bf16=True)
    assert y[0] == 1.0
    assert y[1] == np.float32(1 + 2.0 ** -7)


def test_o2_many_equals_single():
    """The all-cores CPU baseline (O2 across queries) gives each query's
    single-threaded O2 result."""
    g = G.small_graph(5, n=3000, m=12000, labels=3, kind="cl")
    qs = G.queries(g, 24, 3, 7, _lib.Q_WALK_HALF, 17, 10 ** 7)
    many = O.o2_match_many(g, qs, threads=4)
    for q, (c, h) in zip(qs, many):
        r = O.o2_match(g, q, mode=0)
        assert (c, h) == (r["count"], r["hash"])
