"""The C-ABI library: loads, exports every declared symbol, and its host-only
parts (generators, text writer, tuple hash) behave.  CPU only: no call here
touches a GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
from gpsm_amd import _lib
from gpsm_amd import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("gpsm.h", "gpsm_gen.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(gpsm_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert declared() <= set(_lib.SIGNATURES), sorted(declared() - set(_lib.SIGNATURES))


def test_init_without_device_fails_cleanly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = C.c_void_p()
    assert _lib.lib().gpsm_init(0, C.byref(h)) == -6  # GPSM_ENODEV, no exit()


def _check_csr(g):
    ro, col = g.row_offset.astype(np.int64), g.col.astype(np.int64)
    assert ro[0] == 0 and np.all(np.diff(ro) >= 0)
    src = np.repeat(np.arange(g.n), np.diff(ro))
    assert np.all(src != col)
    for v in range(0, g.n, max(1, g.n // 200)):
        r = col[ro[v]:ro[v + 1]]
        assert np.all(np.diff(r) > 0)
    # symmetric
    fwd = set(zip(src.tolist(), col.tolist()))
    assert all((b, a) in fwd for a, b in list(fwd)[:5000])


def test_generators_deterministic_and_valid():
    a = G.erdos_renyi(5000, 20000, 7)
    b = G.erdos_renyi(5000, 20000, 7)
    assert np.array_equal(a.col, b.col) and a.adj == 40000
    _check_csr(a)
    c = G.chung_lu(20000, 80000, 2.1, 3)
    assert c.adj == 160000
    _check_csr(c)
    r = G.rmat(12, 20000, 0.57, 0.19, 0.19, 5)
    assert r.adj == 40000
    _check_csr(r)
    G.set_labels(c, 16, 1.0, 9)
    cnt = np.bincount(c.labels, minlength=17)[1:]
    assert cnt[0] > cnt[3] > cnt[15] > 0  # Zipf(1.0)


@pytest.mark.parametrize("kind", [_lib.Q_WALK_TREE, _lib.Q_WALK_INDUCED, _lib.Q_CYCLE, _lib.Q_CLIQUE])
def test_queries_are_data_subgraphs(kind):
    g = G.small_graph(4, n=3000, m=9000, labels=3, kind="cl")
    k = 4 if kind == _lib.Q_CLIQUE else 6
    qs = G.queries(g, 6, k, k, kind, 11)
    assert qs == qs and len(qs) == 6
    for q in qs:
        assert q.n == k
        # connected, simple
        assert len(set(zip(q.src.tolist(), q.dst.tolist()))) == q.m
        # at least one embedding (the walk itself)
        assert O.o1_enumerate(g, q)["count"] >= 1
        if kind == _lib.Q_CYCLE:
            assert q.m == k
        if kind == _lib.Q_CLIQUE:
            assert q.m == k * (k - 1) // 2


def test_query_sets_do_not_depend_on_thread_count():
    g = G.small_graph(4, n=3000, m=9000, labels=3, kind="cl")
    a = G.queries(g, 8, 4, 8, _lib.Q_WALK_HALF, 5, 10 ** 7)
    os.environ["OMP_NUM_THREADS"] = "1"
    b = G.queries(g, 8, 4, 8, _lib.Q_WALK_HALF, 5, 10 ** 7)
    for x, y in zip(a, b):
        assert np.array_equal(x.labels, y.labels) and np.array_equal(x.src, y.src)


def test_tuple_hash_matches_oracle():
    rng = np.random.default_rng(0)
    for k in (2, 5, 17):
        t = rng.integers(0, 1 << 31, size=k).astype(np.uint32)
        assert _lib.lib().gpsm_tuple_hash(_lib.ptr(t), k) == O.tuple_hash(t)


def test_text_writer_format(tmp_path):
    g = G.complete(4)
    src, dst = g.edges()
    p = str(tmp_path / "g.txt").encode()
    assert _lib.lib().gpsm_io_write_graph(p, 0, 0, g.n, _lib.ptr(g.labels), len(src), _lib.ptr(src),
                                          _lib.ptr(dst)) == 0
    assert _lib.lib().gpsm_io_write_end(p) == 0
    lines = open(p.decode()).read().splitlines()
    assert lines[0] == "t # 0" and lines[1].split()[:2] == ["4", "6"]
    assert lines[2] == "v 0 1" and lines[6] == "e 0 1 0" and lines[-1] == "t # -1"


def _count(path):
    n, v, e = C.c_uint32(0), C.c_uint64(0), C.c_uint64(0)
    rc = _lib.lib().gpsm_io_count_graphs(str(path).encode(), C.byref(n), C.byref(v), C.byref(e))
    return rc, n.value, v.value, e.value


def test_reader_roundtrip_and_errors(tmp_path):
    """The text reader reads what the writer writes, and rejects a missing
    terminator or a malformed line instead of looping (io/IO.cpp:53-97)."""
    p = tmp_path / "two.g"
    for gid, g in enumerate([G.complete(5), G.cycle(7)]):
        src, dst = g.edges()
        _lib.lib().gpsm_io_write_graph(str(p).encode(), gid, gid, g.n, _lib.ptr(g.labels), len(src),
                                       _lib.ptr(src), _lib.ptr(dst))
    _lib.lib().gpsm_io_write_end(str(p).encode())
    assert _count(p) == (0, 2, 12, 17)
    text = p.read_text()
    bad = tmp_path / "unterminated.g"
    bad.write_text(text.replace("t # -1\n", ""))
    assert _count(bad)[0] == -8
    bad2 = tmp_path / "malformed.g"
    bad2.write_text(text.replace("e 0 1 0", "x 0 1 0", 1))
    assert _count(bad2)[0] == -8
    bad3 = tmp_path / "range.g"
    bad3.write_text(text.replace("e 0 1 0", "e 0 99 0", 1))
    assert _count(bad3)[0] == -8


def test_binary_csr_cache_roundtrip_and_errors(tmp_path):
    """A graph written to the binary cache maps back with identical arrays;
    damaged files are refused with GPSM_EFORMAT."""
    g = G.small_graph(3, n=5000, m=20000, labels=7, kind="cl")
    p = tmp_path / "g.csr"
    G.save_csr(g, p)
    m = G.load_csr(p)
    assert m.n == g.n and m.adj == g.adj
    assert np.array_equal(m.row_offset, g.row_offset) and np.array_equal(m.col, g.col)
    assert np.array_equal(m.labels, g.labels)
    m.close()
    raw = p.read_bytes()
    assert raw[:8] == b"GPSMCSR\x01" and len(raw) == 64 + 4 * (2 * g.n + 1 + g.adj)
    bad = tmp_path / "bad.csr"
    cases = [raw[:-4],                                   # truncated
             b"GPSMCSR\x02" + raw[8:],                   # unknown version
             raw[:64] + (1).to_bytes(4, "little") + raw[68:]]  # row_offset[0] != 0
    ro_end = 64 + 4 * g.n  # row_offset[n]
    cases.append(raw[:ro_end] + (g.adj - 1).to_bytes(4, "little") + raw[ro_end + 4:])  # offsets end early
    for data in cases:
        bad.write_bytes(data)
        with pytest.raises(_lib.GpsmError) as e:
            G.load_csr(bad)
        assert e.value.code == -8
    with pytest.raises(_lib.GpsmError):
        G.load_csr(tmp_path / "missing.csr")


def test_cli_writes_the_binary_cache(tmp_path):
    """GpSM --write-cache turns the first data graph of a text file into a
    binary cache whose arrays equal the CSR the reader builds."""
    g = G.small_graph(4, n=2000, m=9000, labels=5, kind="er")
    p = tmp_path / "data.g"
    src, dst = g.edges()
    _lib.lib().gpsm_io_write_graph(str(p).encode(), 0, 0, g.n, _lib.ptr(g.labels), len(src), _lib.ptr(src),
                                   _lib.ptr(dst))
    _lib.lib().gpsm_io_write_end(str(p).encode())
    out = tmp_path / "data.csr"
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "GpSM")
    r = subprocess.run([exe, "--write-cache", str(p), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    m = G.load_csr(out)
    assert np.array_equal(m.row_offset, g.row_offset) and np.array_equal(m.col, g.col)
    assert np.array_equal(m.labels, g.labels)
    m.close()
    bad = subprocess.run([exe, "--write-cache", str(tmp_path / "missing.g"), str(out)], capture_output=True,
                         timeout=60)
    assert bad.returncode != 0


def test_cli_directed_refuses_count_and_cache(tmp_path):
    """--directed needs tuples and the text format's arc lists: with --count
    or a binary cache the driver stops with a usage error before touching a
    device."""
    g = G.small_graph(5, n=300, m=900, labels=2, kind="er")
    p = tmp_path / "data.g"
    src, dst = g.edges()
    _lib.lib().gpsm_io_write_graph(str(p).encode(), 0, 0, g.n, _lib.ptr(g.labels), len(src), _lib.ptr(src),
                                   _lib.ptr(dst))
    _lib.lib().gpsm_io_write_end(str(p).encode())
    cache = tmp_path / "data.csr"
    G.save_csr(g, cache)
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "GpSM")
    for args in (["--directed", "--count", str(p), str(p)], ["--directed", str(cache), str(p)]):
        r = subprocess.run([exe] + args + [str(tmp_path / "ans.txt")], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "--directed" in r.stderr, r.stderr


def test_generators_follow_their_laws():
    """Chung-Lu: vertex i's expected degree is 2m * min(w_i * s, cap) / sum,
    w_i = (i + 1)^(-1/(gamma - 1)), s scaling the weights to 2m, cap =
    sqrt(2m); mean degrees of id slices match it (a wrong exponent or a
    missing cap fails).  R-MAT: the top-level quadrant of the deduplicated
    undirected edges follows (a, b, c, d), the dense corner first."""
    n, m, gamma = 20000, 80000, 2.1
    c = G.chung_lu(n, m, gamma, 3)
    deg = np.diff(c.row_offset).astype(np.float64)
    w = (np.arange(n) + 1.0) ** (-1.0 / (gamma - 1.0))
    cw = np.minimum(w * (2.0 * m / w.sum()), np.sqrt(2.0 * m))
    expect = 2.0 * m * cw / cw.sum()
    for lo, hi in ((100, 200), (1000, 2000), (10000, 20000)):
        assert abs(deg[lo:hi].mean() / expect[lo:hi].mean() - 1) < 0.1, (lo, hi)
    assert deg.max() < 2 * np.sqrt(2.0 * m)  # the cap: uncapped, vertex 0 would expect ~9,600
    r = G.rmat(12, 20000, 0.57, 0.19, 0.19, 5)
    src = np.repeat(np.arange(r.n), np.diff(r.row_offset).astype(np.int64))
    dst = r.col.astype(np.int64)
    keep = src < dst
    s, d, half = src[keep], dst[keep], r.n // 2
    both_low, both_high = np.mean((s < half) & (d < half)), np.mean((s >= half) & (d >= half))
    assert 0.40 < both_low < 0.52 and 0.06 < both_high < 0.12


@pytest.mark.parametrize("nl,s", [(16, 1.0), (32, 0.8)])
def test_label_shares_follow_zipf(nl, s):
    """Labels 1..L with P(k) proportional to k^-s (16 labels at s = 1.0:
    P(1) = 0.296, P(16) = 0.018)."""
    g = G.chung_lu(50000, 100000, 2.1, 11)
    G.set_labels(g, nl, s, 4)
    share = np.bincount(g.labels, minlength=nl + 1)[1:] / g.n
    p = np.arange(1, nl + 1, dtype=np.float64) ** -s
    p /= p.sum()
    assert np.all(np.abs(share - p) < 4 * np.sqrt(p * (1 - p) / g.n) + 1e-3)
    if nl == 16:
        assert abs(p[0] - 0.296) < 0.001 and abs(p[15] - 0.018) < 0.001


def test_walk_query_kinds():
    """A tree walk keeps the k - 1 first-visit edges; an induced walk keeps
    every data edge among its k vertices, so some embedding's image spans
    exactly the query's edges; the half/half mix alternates them by slot."""
    g = G.small_graph(4, n=3000, m=9000, labels=6, kind="cl")
    ro, col = g.row_offset, g.col

    def edges_among(vs):
        s = set(int(v) for v in vs)
        return sum(int(np.isin(col[ro[v]:ro[v + 1]], list(s)).sum()) for v in s) // 2

    for q in G.queries(g, 6, 5, 7, _lib.Q_WALK_TREE, 21):
        assert q.m == q.n - 1
    induced = G.queries(g, 6, 4, 5, _lib.Q_WALK_INDUCED, 22)  # small: O1 enumerates every embedding
    assert any(q.m > q.n - 1 for q in induced)  # some walk closes a cycle (not only first-visit edges)
    for q in induced:
        assert q.m >= q.n - 1
        r = O.o1_enumerate(g, q, tuple_cap=300000)
        assert any(edges_among(t) == q.m for t in r["tuples"])
    half = G.queries(g, 8, 5, 7, _lib.Q_WALK_HALF, 23)
    assert all(q.m == q.n - 1 for q in half[0::2])
