"""Python face of the matcher: a device context, a resident data graph and
the match entry points, all through the C ABI of libgpsm.so.

Mirrors the reference's host interface (bookug/GpSM match/Match.h:80-125:
Match::initGPU, Match(query, data), Match::match) with an explicit context and
a graph uploaded once.  There is no CPU path: without a HIP device every call
raises GpsmError.
"""
import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, ptr, u32


@dataclass
class MatchResult:
    count: int
    hash: int
    order: list
    level_rows: list
    cand: list
    device_ms: float
    status: int = 0
    join_work: int = 0
    level_work: list = field(default=None)
    tuples: np.ndarray = field(default=None)


def _result(r, q):
    return MatchResult(int(r.count), int(r.hash), list(r.order[:q]), [int(x) for x in r.level_rows[:q]],
                       list(r.cand[:q]), float(r.device_ms), int(r.status), int(r.join_work),
                       [int(x) for x in r.level_work[:q]])


class _BatchResult:
    """One query's result of a batch, read from the library's result struct
    when a field is used: a batch of hundreds of queries then converts only
    what the caller reads (MatchResult's fields; tuples is always None)."""
    __slots__ = ("_r", "_q")
    tuples = None

    def __init__(self, r, q):
        self._r = r  # a view into the batch's result array, which it keeps alive
        self._q = q

    count = property(lambda s: int(s._r.count))
    hash = property(lambda s: int(s._r.hash))
    order = property(lambda s: list(s._r.order[:s._q]))
    level_rows = property(lambda s: [int(x) for x in s._r.level_rows[:s._q]])
    cand = property(lambda s: list(s._r.cand[:s._q]))
    device_ms = property(lambda s: float(s._r.device_ms))
    status = property(lambda s: int(s._r.status))
    join_work = property(lambda s: int(s._r.join_work))
    level_work = property(lambda s: [int(x) for x in s._r.level_work[:s._q]])


def _opts(order_mode=_lib.ORDER_HEURISTIC, explicit_order=None, want_tuples=False, tuple_cap=0, use_graph=False,
          block=0, join_cap=0, work_budget=0):
    o = _lib.Opts()
    o.order_mode = order_mode
    keep = None
    if explicit_order is not None:
        keep = u32(explicit_order)
        o.explicit_order = ptr(keep)
    o.want_tuples = int(want_tuples)
    o.tuple_cap = tuple_cap
    o.use_graph = int(use_graph)
    o.block = block
    o.join_cap = join_cap
    o.work_budget = int(work_budget)
    return o, keep


class Context:
    """One HIP device and stream (Match::initGPU, match/Match.cu:15-70)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(_lib.lib().gpsm_init(device, C.byref(h)), "gpsm_init")
        self.h = h
        self.planner = None
        self.graphs = weakref.WeakSet()  # resident graphs, freed before the context

    @property
    def device_name(self):
        return _lib.lib().gpsm_device_name(self.h).decode()

    def close(self):
        if self.h:
            for g in list(self.graphs):
                g.free()
            if self.planner is not None:
                self.planner.free()
                self.planner = None
            _lib.lib().gpsm_shutdown(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, g):
        return ResidentGraph(self, g)

    def upload_edges(self, n, src, dst, labels):
        """A resident graph built on the device from an edge list
        (gpsm_graph_upload_edges); its CSR is read back into .g."""
        from .graphs import DataGraph
        src, dst, lab = u32(src), u32(dst), u32(labels)
        h = C.c_void_p()
        check(_lib.lib().gpsm_graph_upload_edges(self.h, n, len(src), ptr(src), ptr(dst), ptr(lab), C.byref(h)),
              "gpsm_graph_upload_edges")
        nn, adj = C.c_uint32(), C.c_uint64()
        _lib.lib().gpsm_graph_info(h, C.byref(nn), C.byref(adj))
        ro = np.zeros(nn.value + 1, dtype=np.uint32)
        col = np.zeros(max(adj.value, 1), dtype=np.uint32)
        rc = _lib.lib().gpsm_graph_csr(h, ptr(ro), ptr(col))
        rg = ResidentGraph(self, DataGraph(ro, col[:adj.value], lab.copy()), handle=h)
        check(rc, "gpsm_graph_csr")
        return rg

    def profile(self, mode=1):
        """0 off, 1 per-launch timing, 2 timing + the filter's work counters
        (algorithmic bytes), 3 timing of the join launches only; setting a
        mode clears the statistics."""
        check(_lib.lib().gpsm_profile_enable(self.h, int(mode)), "gpsm_profile_enable")

    def kernel_stats(self):
        arr = (_lib.KStat * 16)()
        n = C.c_uint32(0)
        check(_lib.lib().gpsm_profile_read(self.h, arr, 16, C.byref(n)), "gpsm_profile_read")
        return {arr[i].name.decode(): {"launches": int(arr[i].launches), "ms": float(arr[i].total_ms),
                                       "bytes": float(arr[i].bytes)} for i in range(n.value)}

    def set_hub_policy(self, min_degree=64, max_bytes=48 << 30):
        """Hub adjacency bitmaps for graphs uploaded afterwards (vertices of
        degree > D, D doubled from min_degree until they fit max_bytes)."""
        check(_lib.lib().gpsm_set_hub_policy(self.h, min_degree, max_bytes), "gpsm_set_hub_policy")

    def graph_cache_stats(self):
        """(instantiations, in-place updates, relaunches without a capture) of
        the context's cached hipGraphs."""
        i, u, r = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(_lib.lib().gpsm_graph_cache_stats(self.h, C.byref(i), C.byref(u), C.byref(r)),
              "gpsm_graph_cache_stats")
        return i.value, u.value, r.value

    def tuning(self):
        """The context's launch tuning (defaults until gpsm_autotune runs)."""
        v = [C.c_uint32() for _ in range(4)]
        check(_lib.lib().gpsm_tuning(self.h, *[C.byref(x) for x in v]), "gpsm_tuning")
        return dict(zip(["filter_block", "join_cap", "join_wg_per_cu", "tasks_per_wg"], [x.value for x in v]))

    def set_tuning(self, filter_block, join_cap, join_wg_per_cu, tasks_per_wg):
        check(_lib.lib().gpsm_set_tuning(self.h, filter_block, join_cap, join_wg_per_cu, tasks_per_wg),
              "gpsm_set_tuning")

    # --- parity entry points -------------------------------------------------
    def scan(self, a):
        a = u32(a)
        out = np.zeros_like(a)
        tot = C.c_uint64(0)
        check(_lib.lib().gpsm_scan_u32(self.h, ptr(a), ptr(out), len(a), C.byref(tot)), "gpsm_scan_u32")
        return out, tot.value

    def compact(self, bits, nbits):
        bits = u32(bits)
        ids = np.zeros(max(nbits, 1), dtype=np.uint32)
        cnt = C.c_uint64(0)
        check(_lib.lib().gpsm_compact_bitmap(self.h, ptr(bits), nbits, ptr(ids), C.byref(cnt)),
              "gpsm_compact_bitmap")
        return ids[:cnt.value].copy()


class ResidentGraph:
    """A data graph uploaded once to HBM (replaces Match::copyGraphToGPU,
    which re-uploaded it for every query, match/Match.cu:128-143)."""

    def __init__(self, ctx, g, handle=None):
        self.ctx = ctx
        self.g = g
        if handle is None:
            handle = C.c_void_p()
            ro, col, lab = u32(g.row_offset), u32(g.col), u32(g.labels)
            check(_lib.lib().gpsm_graph_upload(ctx.h, len(lab), ptr(ro), ptr(col), ptr(lab), C.byref(handle)),
                  "gpsm_graph_upload")
        self.h = handle
        ctx.graphs.add(self)

    def free(self):
        # a graph whose context is already closed was freed with it
        if self.h and self.ctx.h:
            _lib.lib().gpsm_graph_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def match(self, q, **kw):
        o, keep = _opts(**kw)
        qs = q.struct()
        r = _lib.Result()
        rc = _lib.lib().gpsm_match(self.ctx.h, self.h, C.byref(qs), C.byref(o), C.byref(r))
        if rc not in (0,):
            raise _lib.GpsmError(rc, "gpsm_match")
        res = _result(r, q.n)
        if kw.get("want_tuples"):
            nt = int(r.n_tuples)
            res.tuples = np.ctypeslib.as_array(r.tuples, shape=(max(nt * q.n, 1),))[:nt * q.n].reshape(
                nt, q.n).copy() if nt else np.zeros((0, q.n), dtype=np.uint32)
            _lib.lib().gpsm_result_free(C.byref(r))
        return res

    def match_batch(self, queries, explicit_orders=None, **kw):
        """explicit_orders: one vertex order per query (with order_mode=ORDER_EXPLICIT)."""
        if explicit_orders is not None:
            arr_o = np.zeros((len(queries), _lib.MAX_QV), dtype=np.uint32)
            for i, od in enumerate(explicit_orders):
                arr_o[i, :len(od)] = od
            kw["explicit_order"] = arr_o.reshape(-1)
            kw.setdefault("order_mode", _lib.ORDER_EXPLICIT)
        o, keep = _opts(**kw)
        arr = (_lib.QueryS * len(queries))()
        for i, q in enumerate(queries):
            arr[i] = q.struct()
        res = (_lib.Result * len(queries))()
        check(_lib.lib().gpsm_match_batch(self.ctx.h, self.h, len(queries), arr, C.byref(o), res),
              "gpsm_match_batch")
        return [_BatchResult(res[i], q.n) for i, q in enumerate(queries)]

    def plan_features(self, q, cand, cet_size, order, level_rows):
        """Planner training rows of one measured run (gpsm_plan_features)."""
        qs = q.struct()
        cand = u32(cand)
        cs = np.ascontiguousarray(cet_size, dtype=np.uint64)
        od = u32(order)
        lr = np.ascontiguousarray(level_rows, dtype=np.uint64)
        feat = np.zeros((max(q.n - 2, 1), _lib.FEAT), dtype=np.float32)
        tgt = np.zeros(max(q.n - 2, 1), dtype=np.float32)
        n = _lib.lib().gpsm_plan_features(self.h, C.byref(qs), ptr(cand), ptr(cs), ptr(od), ptr(lr), ptr(feat),
                                          ptr(tgt))
        if n < 0:
            raise _lib.GpsmError(n, "gpsm_plan_features")
        return feat[:n], tgt[:n]

    def hubs(self):
        """(number of hubs, degree above which a vertex is one)."""
        nh, d = C.c_uint32(), C.c_uint32()
        check(_lib.lib().gpsm_graph_hubs(self.h, C.byref(nh), C.byref(d)), "gpsm_graph_hubs")
        return nh.value, d.value

    @property
    def device_bytes(self):
        return int(_lib.lib().gpsm_graph_device_bytes(self.h))

    def set_directed(self, src, dst, elabel):
        """Attaches the directed, edge-labelled arcs src[i] -> dst[i] (label
        elabel[i]) of the data graph, the out-lists that Graph::isEdgeContained
        scans (graph/Graph.cpp:107-118), for verify_directed."""
        src = np.asarray(src, dtype=np.uint32)
        dst = np.asarray(dst, dtype=np.uint32)
        lab = np.asarray(elabel, dtype=np.uint32)
        nn, adj = C.c_uint32(), C.c_uint64()
        check(_lib.lib().gpsm_graph_info(self.h, C.byref(nn), C.byref(adj)), "gpsm_graph_info")
        n = nn.value
        if len(src) and int(max(src.max(), dst.max())) >= n:
            raise ValueError("arc endpoint >= n")
        order = np.lexsort((lab, dst, src))
        row = np.zeros(n + 1, dtype=np.uint32)
        np.cumsum(np.bincount(src, minlength=n), out=row[1:])
        col = np.ascontiguousarray(dst[order])
        lab = np.ascontiguousarray(lab[order])
        check(_lib.lib().gpsm_graph_set_directed(self.h, ptr(row), ptr(col), ptr(lab)), "gpsm_graph_set_directed")

    def verify_directed(self, q, arcs, tuples):
        """IO::verify (io/IO.cpp:146-180) on the device: a bool mask over the
        rows of tuples (k x q.n, column j = query vertex j) that are injective
        and map every query arc (src, dst, label) of arcs to a data arc with
        that label.  Needs set_directed first."""
        a = np.asarray(arcs, dtype=np.uint32).reshape(-1, 3)
        src, dst, lab = (np.ascontiguousarray(a[:, i]) for i in range(3))
        t = np.ascontiguousarray(np.asarray(tuples, dtype=np.uint32).reshape(-1, q.n))
        keep = np.zeros(max(len(t), 1), dtype=np.uint8)
        kept = C.c_uint64()
        check(_lib.lib().gpsm_verify_directed(self.ctx.h, self.h, q.n, len(a), ptr(src), ptr(dst), ptr(lab), ptr(t),
                                              len(t), ptr(keep), C.byref(kept)), "gpsm_verify_directed")
        mask = keep[:len(t)].astype(bool)
        assert int(mask.sum()) == kept.value
        return mask

    def cet_edges(self, q):
        """Query edges (from < to, the planner's edge ids) and |CET| of each
        (gpsm_candidate_edges without the lists)."""
        qs = q.struct()
        ef = np.zeros(max(q.m, 1), dtype=np.uint32)
        et = np.zeros(max(q.m, 1), dtype=np.uint32)
        size = np.zeros(max(q.m, 1), dtype=np.uint32)
        check(_lib.lib().gpsm_candidate_edges(self.ctx.h, self.h, C.byref(qs), ptr(ef), ptr(et), ptr(size), None, 0,
                                              None), "gpsm_candidate_edges")
        return ef[:q.m], et[:q.m], size[:q.m].astype(np.uint64)

    def cet_sizes(self, q):
        """|CET| of every query edge."""
        return self.cet_edges(q)[2]

    def autotune(self, queries, **kw):
        """Runs gpsm_autotune on a sample batch; returns (the join kernel's ms with
        the chosen setting, tuning)."""
        o, keep = _opts(**kw)
        arr = (_lib.QueryS * len(queries))()
        for i, q in enumerate(queries):
            arr[i] = q.struct()
        ms = C.c_double(0)
        check(_lib.lib().gpsm_autotune(self.ctx.h, self.h, len(queries), arr, C.byref(o), C.byref(ms)),
              "gpsm_autotune")
        return ms.value, self.ctx.tuning()

    def filter(self, q):
        W = (self.g.n + 31) // 32
        bm = np.zeros(q.n * W, dtype=np.uint32)
        cnt = np.zeros(q.n, dtype=np.uint32)
        qs = q.struct()
        check(_lib.lib().gpsm_filter(self.ctx.h, self.h, C.byref(qs), ptr(bm), ptr(cnt)), "gpsm_filter")
        return bm.reshape(q.n, W), cnt

    def candidate_edges(self, q):
        """Forward candidate-edge tables in the reference layout
        [C(u) ids | offsets | lists] (match/Match.cu:1119-1121)."""
        _, cand = self.filter(q)
        qs = q.struct()
        E = q.m
        ef = np.zeros(E, dtype=np.uint32)
        et = np.zeros(E, dtype=np.uint32)
        size = np.zeros(E, dtype=np.uint32)
        check(_lib.lib().gpsm_candidate_edges(self.ctx.h, self.h, C.byref(qs), ptr(ef), ptr(et), ptr(size), None, 0,
                                              None), "gpsm_candidate_edges")
        m = [int(cand[ef[e]]) for e in range(E)]
        cap_words = sum(2 * m[e] + 1 + int(size[e]) for e in range(E))
        buf = np.zeros(max(cap_words, 1), dtype=np.uint32)
        toff = np.zeros(max(E, 1), dtype=np.uint64)
        check(_lib.lib().gpsm_candidate_edges(self.ctx.h, self.h, C.byref(qs), ptr(ef), ptr(et), ptr(size),
                                              ptr(buf), cap_words, ptr(toff)), "gpsm_candidate_edges")
        out = []
        for e in range(E):
            t = buf[int(toff[e]):]
            k, s = m[e], int(size[e])
            out.append((t[:k].copy(), t[k:2 * k + 1].copy(), t[2 * k + 1:2 * k + 1 + s].copy()))
        return [(int(ef[i]), int(et[i])) for i in range(E)], out
