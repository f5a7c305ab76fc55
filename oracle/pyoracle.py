"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never as the measured or
shipped path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MAXQ = 64
MAXE = 2048

OK, EMPTY_INIT, EMPTY_REFINE, EMPTY_JOIN = 0, 1, 2, 3


class Graph(C.Structure):
    _fields_ = [("n", C.c_uint32), ("row_offset", C.c_void_p), ("col", C.c_void_p), ("labels", C.c_void_p)]


class Query(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("labels", C.c_void_p), ("src", C.c_void_p),
                ("dst", C.c_void_p)]


class O2Out(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("count", C.c_uint64),
        ("hash", C.c_uint64),
        ("filter_order", C.c_uint32 * MAXQ),
        ("n_edges", C.c_uint32),
        ("edge_from", C.c_uint32 * MAXE),
        ("edge_to", C.c_uint32 * MAXE),
        ("cet_size", C.c_uint32 * MAXE),
        ("cand", C.c_uint32 * MAXQ),
        ("order", C.c_uint32 * MAXQ),
        ("level_rows", C.c_uint64 * MAXQ),
        ("n_steps", C.c_uint32),
        ("step_edge", C.c_uint32 * MAXE),
        ("step_join", C.c_uint8 * MAXE),
        ("step_rows", C.c_uint64 * MAXE),
        ("tuples", C.c_void_p),
        ("tuple_cap", C.c_uint64),
        ("n_tuples", C.c_uint64),
        ("join_work", C.c_uint64),
        ("level_work", C.c_uint64 * MAXQ),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_o2_match.argtypes = [C.POINTER(Graph), C.POINTER(Query), C.c_int, C.c_void_p, C.POINTER(O2Out)]
        L.orc_o2_match_many.argtypes = [C.POINTER(Graph), C.POINTER(Query), C.c_uint32, C.c_int, C.c_void_p,
                                        C.c_void_p]
        L.orc_o2_filter.argtypes = [C.POINTER(Graph), C.POINTER(Query), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_o2_cet.argtypes = [C.POINTER(Graph), C.POINTER(Query), C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]
        L.orc_o2_cet.restype = C.c_int64
        L.orc_o1_enumerate.argtypes = [C.POINTER(Graph), C.POINTER(Query), C.c_int, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
        L.orc_tuple_hash.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_tuple_hash.restype = C.c_uint64
        L.orc_exclusive_scan_u32.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.orc_exclusive_scan_u32.restype = C.c_uint64
        L.orc_compact_bitmap.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.orc_compact_bitmap.restype = C.c_uint64
        L.orc_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
        L.orc_mlp_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                   C.POINTER(C.c_double), C.c_void_p]
        _LIB = L
    return _LIB


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class _Keep:
    """Keeps numpy buffers alive for the lifetime of a ctypes struct."""

    def __init__(self, *arrs):
        self.arrs = arrs


def make_graph(g):
    ro, col, lab = _u32(g.row_offset), _u32(g.col), _u32(g.labels)
    s = Graph(len(lab), _ptr(ro), _ptr(col), _ptr(lab))
    return s, _Keep(ro, col, lab)


def make_query(q):
    lab, src, dst = _u32(q.labels), _u32(q.src), _u32(q.dst)
    s = Query(len(lab), len(src), _ptr(lab), _ptr(src), _ptr(dst))
    return s, _Keep(lab, src, dst)


def o2_match(g, q, mode=0, order=None, tuple_cap=0):
    gs, k1 = make_graph(g)
    qs, k2 = make_query(q)
    out = O2Out()
    buf = None
    if tuple_cap:
        buf = np.zeros(tuple_cap * len(q.labels), dtype=np.uint32)
        out.tuples = buf.ctypes.data
        out.tuple_cap = tuple_cap
    o = _u32(order) if order is not None else None
    rc = lib().orc_o2_match(C.byref(gs), C.byref(qs), mode, _ptr(o), C.byref(out))
    if rc != 0:
        raise ValueError("orc_o2_match failed (%d)" % rc)
    res = {
        "status": out.status,
        "count": out.count,
        "hash": out.hash,
        "filter_order": list(out.filter_order[: len(q.labels)]),
        "edges": [(out.edge_from[i], out.edge_to[i]) for i in range(out.n_edges)],
        "cet_size": list(out.cet_size[: out.n_edges]),
        "cand": list(out.cand[: len(q.labels)]),
        "order": list(out.order[: len(q.labels)]),
        "level_rows": list(out.level_rows[: len(q.labels)]),
        "step_rows": list(out.step_rows[: out.n_steps]),
        "step_edge": list(out.step_edge[: out.n_steps]),
        "step_join": list(out.step_join[: out.n_steps]),
        "join_work": out.join_work,
        "level_work": list(out.level_work[: len(q.labels)]),
    }
    if tuple_cap:
        nt = min(out.n_tuples, tuple_cap)
        res["tuples"] = buf[: nt * len(q.labels)].reshape(nt, len(q.labels)).copy()
    return res


def o2_filter(g, q):
    gs, k1 = make_graph(g)
    qs, k2 = make_query(q)
    n, k = len(g.labels), len(q.labels)
    init = np.zeros(k * n, dtype=np.uint8)
    ref = np.zeros(k * n, dtype=np.uint8)
    fo = np.zeros(MAXQ, dtype=np.uint32)
    st = lib().orc_o2_filter(C.byref(gs), C.byref(qs), _ptr(init), _ptr(ref), _ptr(fo))
    return st, init.reshape(k, n), ref.reshape(k, n), list(fo[:k])


def o2_cet(g, q):
    gs, k1 = make_graph(g)
    qs, k2 = make_query(q)
    toff = np.zeros(MAXE, dtype=np.uint64)
    ef = np.zeros(MAXE, dtype=np.uint32)
    et = np.zeros(MAXE, dtype=np.uint32)
    cn = np.zeros(MAXE, dtype=np.uint32)
    ne = C.c_uint32(0)
    args = (_ptr(toff), _ptr(ef), _ptr(et), _ptr(cn), C.byref(ne))
    total = lib().orc_o2_cet(C.byref(gs), C.byref(qs), None, 0, *args)
    if total < 0:
        return None
    buf = np.zeros(max(total, 1), dtype=np.uint32)
    lib().orc_o2_cet(C.byref(gs), C.byref(qs), _ptr(buf), total, *args)
    tables = []
    for i in range(ne.value):
        m = int(cn[i])
        t = buf[int(toff[i]):]
        s = int(t[2 * m])
        tables.append((t[:m].copy(), t[m:2 * m + 1].copy(), t[2 * m + 1:2 * m + 1 + s].copy()))
    return [(int(ef[i]), int(et[i])) for i in range(ne.value)], tables


def o2_match_many(g, qs, threads=0):
    """O2 over a query list, OpenMP across queries: [(count, hash)]."""
    gs, k1 = make_graph(g)
    arr = (Query * len(qs))()
    keep = []
    for i, q in enumerate(qs):
        s, k = make_query(q)
        arr[i] = s
        keep.append(k)
    cnt = np.zeros(len(qs), dtype=np.uint64)
    hs = np.zeros(len(qs), dtype=np.uint64)
    rc = lib().orc_o2_match_many(C.byref(gs), arr, len(qs), threads, _ptr(cnt), _ptr(hs))
    if rc < 0:
        raise ValueError("orc_o2_match_many failed (%d)" % rc)
    return [(int(c), int(h)) for c, h in zip(cnt, hs)]


def o1_enumerate(g, q, threads=0, tuple_cap=0):
    gs, k1 = make_graph(g)
    qs, k2 = make_query(q)
    cnt, hs, nt = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    buf = np.zeros(max(tuple_cap, 1) * len(q.labels), dtype=np.uint32) if tuple_cap else None
    rc = lib().orc_o1_enumerate(C.byref(gs), C.byref(qs), threads, C.byref(cnt), C.byref(hs), _ptr(buf),
                                tuple_cap, C.byref(nt))
    if rc != 0:
        raise ValueError("orc_o1_enumerate failed (%d)" % rc)
    res = {"count": cnt.value, "hash": hs.value}
    if tuple_cap:
        k = min(cnt.value, tuple_cap)
        res["tuples"] = buf[: k * len(q.labels)].reshape(k, len(q.labels)).copy()
    return res


def tuple_hash(t):
    t = _u32(t)
    return lib().orc_tuple_hash(_ptr(t), len(t))


def set_hash(tuples):
    h = 0
    for t in tuples:
        h = (h + tuple_hash(t)) & 0xFFFFFFFFFFFFFFFF
    return h


def exclusive_scan(a):
    a = _u32(a)
    out = np.zeros_like(a)
    tot = lib().orc_exclusive_scan_u32(_ptr(a), _ptr(out), len(a))
    return out, tot


def compact_bitmap(bits, nbits):
    bits = _u32(bits)
    ids = np.zeros(max(nbits, 1), dtype=np.uint32)
    c = lib().orc_compact_bitmap(_ptr(bits), nbits, _ptr(ids))
    return ids[:c].copy()


def mlp_forward(params, x, bf16=True):
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(x.shape[0], dtype=np.float32)
    lib().orc_mlp_forward(_ptr(p), _ptr(x), x.shape[0], int(bf16), _ptr(y))
    return y


def mlp_grad(params, x, t, bf16=True):
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(p.shape[0], dtype=np.float64)
    loss = C.c_double(0)
    lib().orc_mlp_grad(_ptr(p), _ptr(x), _ptr(t), x.shape[0], int(bf16), C.byref(loss), _ptr(g))
    return loss.value, g
