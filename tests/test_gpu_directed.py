"""The directed, edge-labelled final check (IO::verify, io/IO.cpp:146-180,
with Graph::isEdgeContained, graph/Graph.cpp:107-118) on the device.

The expected set is independent of the matcher: every injective tuple of
the data vertices, enumerated in numpy, whose vertex labels equal the query's
and whose images of every query arc (a, b, label) are data arcs with that
label.  The matcher's undirected tuples, filtered by gpsm_verify_directed,
must equal it exactly (the undirected result is a superset: a labelled arc
implies the undirected edge)."""
import itertools
import os
import subprocess

import numpy as np
import pytest

from gpsm_amd import _lib
from gpsm_amd import graphs as G

pytestmark = pytest.mark.gpu


def directed_case(seed, n=22, arcs=300, vlabels=2, elabels=3):
    """A directed multigraph with edge labels: reciprocal arcs with other
    labels, repeated arcs and self-loops included."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, size=arcs)
    b = rng.integers(0, n, size=arcs)
    lab = rng.integers(0, elabels, size=arcs)
    k = arcs // 6
    a = np.concatenate([a, b[:k], a[:k], [3, 5]])          # reverse arcs, repeats, self-loops
    b = np.concatenate([b, a[:k], b[:k], [3, 5]])
    lab = np.concatenate([lab, (lab[:k] + 1) % elabels, lab[:k], [0, 1]])
    vlab = rng.integers(1, vlabels + 1, size=n)
    return n, a.astype(np.uint32), b.astype(np.uint32), lab.astype(np.uint32), vlab.astype(np.uint32)


def query_arcs(kind):
    """(n, vertex labels, [(a, b, edge label)]): no self-loops (the matcher
    rejects them); a 2-cycle needs both arcs."""
    return {
        "path": (3, [1, 2, 1], [(0, 1, 0), (2, 1, 1)]),
        "twoway": (3, [1, 1, 2], [(0, 1, 0), (1, 0, 1), (1, 2, 2)]),
        "triangle": (3, [1, 1, 1], [(0, 1, 0), (1, 2, 0), (2, 0, 1)]),
        "star": (4, [2, 1, 1, 1], [(0, 1, 0), (2, 0, 1), (0, 3, 2)]),
        "square": (4, [1, 2, 1, 2], [(0, 1, 1), (1, 2, 0), (3, 2, 2), (3, 0, 0)]),
    }[kind]


def brute_force(n, a, b, lab, vlab, qn, qlab, qarcs):
    has = np.zeros((int(lab.max()) + 2, n, n), dtype=bool)
    has[lab, a, b] = True
    P = np.array(list(itertools.permutations(range(n), qn)), dtype=np.int64)
    ok = np.ones(len(P), dtype=bool)
    for j in range(qn):
        ok &= vlab[P[:, j]] == qlab[j]
    for (x, y, l) in qarcs:
        ok &= has[l, P[:, x], P[:, y]] if l < has.shape[0] else False
    return {tuple(r) for r in P[ok].tolist()}


def as_query(qn, qlab, qarcs):
    return G.Query(np.array(qlab, dtype=np.uint32), np.array([x for x, _, _ in qarcs], dtype=np.uint32),
                   np.array([y for _, y, _ in qarcs], dtype=np.uint32))


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("kind", ["path", "twoway", "triangle", "star", "square"])
def test_verify_directed_equals_brute_force(ctx, seed, kind):
    n, a, b, lab, vlab = directed_case(seed)
    qn, qlab, qarcs = query_arcs(kind)
    g = ctx.upload(G.from_edges(n, a, b, vlab))
    g.set_directed(a, b, lab)
    q = as_query(qn, qlab, qarcs)
    r = g.match(q, want_tuples=True, tuple_cap=1 << 20)
    tuples = r.tuples.reshape(-1, qn)
    keep = g.verify_directed(q, qarcs, tuples)
    got = {tuple(t) for t in tuples[keep].tolist()}
    ref = brute_force(n, a, b, lab, vlab, qn, qlab, qarcs)
    print(kind, seed, "undirected", len(tuples), "directed", len(got))
    assert got == ref


def test_verify_directed_rows_and_edge_cases(ctx):
    n, a, b, lab, vlab = directed_case(3)
    g = ctx.upload(G.from_edges(n, a, b, vlab))
    qn, qlab, qarcs = query_arcs("twoway")
    q = as_query(qn, qlab, qarcs)
    with pytest.raises(_lib.GpsmError):
        g.verify_directed(q, qarcs, np.zeros((1, qn), dtype=np.uint32))  # no arcs attached yet
    g.set_directed(a, b, lab)
    # hand-made rows: a repeated vertex, a vertex id >= n, a missing arc, a valid row
    x, y = int(a[0]), int(b[0])
    valid = next(t for t in itertools.permutations(range(n), 3)
                 if all(((a == t[s]) & (b == t[d]) & (lab == l)).any() for s, d, l in qarcs))
    rows = np.array([[x, x, y], [0, 1, n], list(valid), [0, 1, 2]], dtype=np.uint32)
    keep = g.verify_directed(q, qarcs, rows)
    assert keep[0] == 0 and keep[1] == 0 and keep[2] == 1
    assert keep[3] == all(((a == rows[3][s]) & (b == rows[3][d]) & (lab == l)).any() for s, d, l in qarcs)
    assert len(g.verify_directed(q, qarcs, np.zeros((0, qn), dtype=np.uint32))) == 0
    # a query arc endpoint outside the query
    with pytest.raises(_lib.GpsmError):
        g.verify_directed(q, [(0, 3, 0)], rows)
    # unsorted rows are refused by the raw entry point
    row = np.array([0, 2] + [2] * (n - 1), dtype=np.uint32)
    col = np.array([5, 4], dtype=np.uint32)
    el = np.array([0, 0], dtype=np.uint32)
    assert _lib.lib().gpsm_graph_set_directed(g.h, _lib.ptr(row), _lib.ptr(col), _lib.ptr(el)) == -1  # GPSM_EINVAL
    # many rows: the same answer when the tuples are repeated 50 000 times
    big = np.tile(rows, (50_000, 1))
    kb = g.verify_directed(q, qarcs, big)
    assert np.array_equal(kb, np.tile(keep, 50_000))


def write_text(path, graphs):
    """graphs: [(vertex labels, [(a, b, label)])] in the reference's text format."""
    with open(path, "w") as f:
        for i, (vl, arcs) in enumerate(graphs):
            f.write("t # %d\n%d %d %d %d\n" % (i, len(vl), len(arcs), 0, 0))
            for v, l in enumerate(vl):
                f.write("v %d %d\n" % (v, l))
            for x, y, l in arcs:
                f.write("e %d %d %d\n" % (x, y, l))
        f.write("t # -1\n")


def read_answers(path):
    blocks, cur = [], None
    for line in open(path):
        if line.startswith("query graph:"):
            cur = set()
            blocks.append(cur)
        elif line.startswith("("):
            cur.add(tuple(int(p.split(",")[1].strip(" )")) for p in line.split(") ") if p.strip()))
    return blocks


def test_cli_directed(tmp_path):
    """GpSM --directed writes exactly the brute-force directed, labelled set;
    without it the output is the larger undirected set."""
    n, a, b, lab, vlab = directed_case(4)
    data, query = tmp_path / "d.g", tmp_path / "q.g"
    write_text(data, [(vlab.tolist(), list(zip(a.tolist(), b.tolist(), lab.tolist())))])
    kinds = ["twoway", "triangle", "square"]
    write_text(query, [(query_arcs(k)[1], query_arcs(k)[2]) for k in kinds])
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "GpSM")
    outs = {}
    for flag in ("--directed", None):
        ans = tmp_path / ("ans_%s.txt" % bool(flag))
        cmd = [exe] + ([flag] if flag else []) + [str(data), str(query), str(ans)]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs[flag] = read_answers(ans)
    for i, k in enumerate(kinds):
        qn, qlab, qarcs = query_arcs(k)
        ref = brute_force(n, a, b, lab, vlab, qn, qlab, qarcs)
        assert outs["--directed"][i] == ref, k
        assert ref < outs[None][i], k  # the check removes something
