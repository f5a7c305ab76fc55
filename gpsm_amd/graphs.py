"""Data graphs and queries as numpy arrays, and the seeded generators of the
benchmark workloads (host code in libgpsm.so, SplitMix64 streams)."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr, u32


@dataclass
class DataGraph:
    """Undirected CSR: sorted rows, no duplicates, no self-loops, labels >= 1."""
    row_offset: np.ndarray
    col: np.ndarray
    labels: np.ndarray

    @property
    def n(self):
        return len(self.labels)

    @property
    def adj(self):
        return int(self.row_offset[-1])

    def degree(self):
        return np.diff(self.row_offset.astype(np.int64))

    def edges(self):
        """Undirected edge list (a < b)."""
        src = np.repeat(np.arange(self.n, dtype=np.uint32), self.degree())
        keep = src < self.col
        return src[keep], self.col[keep]


@dataclass
class Query:
    """A query graph: labels and its edge list in file order (src, dst)."""
    labels: np.ndarray
    src: np.ndarray
    dst: np.ndarray

    @property
    def n(self):
        return len(self.labels)

    @property
    def m(self):
        return len(self.src)

    def struct(self):
        """ctypes gpsm_query pointing at this query's arrays (keep self alive).
        Cached while the three arrays stay the same objects."""
        key = (id(self.labels), id(self.src), id(self.dst))
        if self.__dict__.get("_skey") == key:
            return self.__dict__["_s"]
        self.labels = u32(self.labels)
        self.src = u32(self.src)
        self.dst = u32(self.dst)
        s = _lib.QueryS(self.n, self.m, ptr(self.labels), ptr(self.src), ptr(self.dst))
        self.__dict__["_skey"] = (id(self.labels), id(self.src), id(self.dst))
        self.__dict__["_s"] = s
        return s


class MappedGraph(DataGraph):
    """A DataGraph whose arrays are a read-only mapping of a binary CSR
    cache file (gpsm_io_map_csr); close() unmaps it."""

    def __init__(self, s):
        n, adj = int(s.n), int(s.adj)
        super().__init__(np.ctypeslib.as_array(s.row_offset, shape=(n + 1,)),
                         np.ctypeslib.as_array(s.col, shape=(max(adj, 1),))[:adj],
                         np.ctypeslib.as_array(s.labels, shape=(n,)))
        self._s = s

    def close(self):
        if self._s is not None:
            self.row_offset = self.col = self.labels = None
            _lib.lib().gpsm_io_unmap_csr(C.byref(self._s))
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def save_csr(g, path):
    """Writes g as a binary CSR cache file (gpsm_io_write_csr)."""
    ro, col, lab = u32(g.row_offset), u32(g.col), u32(g.labels)
    check(_lib.lib().gpsm_io_write_csr(str(path).encode(), len(lab), ptr(ro), ptr(col), ptr(lab)),
          "gpsm_io_write_csr")


def load_csr(path):
    """Maps a binary CSR cache file (gpsm_io_map_csr): a MappedGraph."""
    s = _lib.CSR()
    check(_lib.lib().gpsm_io_map_csr(str(path).encode(), C.byref(s)), "gpsm_io_map_csr")
    return MappedGraph(s)


def _take_csr(s):
    n = s.n
    adj = s.adj
    g = DataGraph(np.ctypeslib.as_array(s.row_offset, shape=(n + 1,)).copy(),
                  np.ctypeslib.as_array(s.col, shape=(max(adj, 1),))[:adj].copy(),
                  np.ctypeslib.as_array(s.labels, shape=(max(n, 1),))[:n].copy())
    _lib.lib().gpsm_gen_free_csr(C.byref(s))
    return g


def _as_csr(g):
    g.row_offset, g.col, g.labels = u32(g.row_offset), u32(g.col), u32(g.labels)
    s = _lib.CSR(g.n, g.adj, g.row_offset.ctypes.data_as(C.POINTER(C.c_uint32)),
                 g.col.ctypes.data_as(C.POINTER(C.c_uint32)), g.labels.ctypes.data_as(C.POINTER(C.c_uint32)))
    return s


def from_edges(n, a, b, labels=None):
    """CSR from an undirected edge list (duplicates and self-loops dropped)."""
    a, b = u32(a), u32(b)
    lab = u32(labels) if labels is not None else None
    s = _lib.CSR()
    check(_lib.lib().gpsm_gen_csr_from_edges(n, len(a), ptr(a), ptr(b), ptr(lab), C.byref(s)), "csr_from_edges")
    return _take_csr(s)


def erdos_renyi(n, m, seed):
    s = _lib.CSR()
    check(_lib.lib().gpsm_gen_er(n, m, seed, C.byref(s)), "gen_er")
    return _take_csr(s)


def chung_lu(n, m, gamma, seed):
    s = _lib.CSR()
    check(_lib.lib().gpsm_gen_chung_lu(n, m, gamma, seed, C.byref(s)), "gen_chung_lu")
    return _take_csr(s)


def rmat(scale, m, a, b, c, seed):
    s = _lib.CSR()
    check(_lib.lib().gpsm_gen_rmat(scale, m, a, b, c, seed, C.byref(s)), "gen_rmat")
    return _take_csr(s)


def set_labels(g, L, zipf_s, seed):
    s = _as_csr(g)
    check(_lib.lib().gpsm_gen_labels(C.byref(s), L, zipf_s, seed), "gen_labels")
    return g


def queries(g, nq, kmin, kmax, kind, seed, node_budget=0):
    s = _as_csr(g)
    qs = _lib.QSet()
    check(_lib.lib().gpsm_gen_queries(C.byref(s), nq, kmin, kmax, kind, seed, node_budget, C.byref(qs)),
          "gen_queries")
    nv = qs.v_off[nq]
    ne = qs.e_off[nq]
    v_off = np.ctypeslib.as_array(qs.v_off, shape=(nq + 1,)).copy()
    e_off = np.ctypeslib.as_array(qs.e_off, shape=(nq + 1,)).copy()
    lab = np.ctypeslib.as_array(qs.labels, shape=(nv + 1,))[:nv].copy()
    src = np.ctypeslib.as_array(qs.src, shape=(ne + 1,))[:ne].copy()
    dst = np.ctypeslib.as_array(qs.dst, shape=(ne + 1,))[:ne].copy()
    _lib.lib().gpsm_gen_free_qset(C.byref(qs))
    return [Query(lab[v_off[i]:v_off[i + 1]].copy(), src[e_off[i]:e_off[i + 1]].copy(),
                  dst[e_off[i]:e_off[i + 1]].copy()) for i in range(nq)]


# --- the benchmark workloads ---------------------------------------------------

def er10k_walk6(seed=1):
    """ER 10,000 vertices / 40,000 edges, 8 uniform labels, one 6-vertex walk query."""
    g = erdos_renyi(10_000, 40_000, seed)
    set_labels(g, 8, 0.0, seed)
    q = queries(g, 1, 6, 6, _lib.Q_WALK_TREE, seed)
    return g, q


def chunglu_1m(seed=2, n=1_000_000, m=8_000_000):
    """Chung-Lu power law (gamma 2.1), 16 labels Zipf(1.0)."""
    g = chung_lu(n, m, 2.1, seed)
    set_labels(g, 16, 1.0, seed)
    return g


def small_graph(seed, n=2000, m=8000, labels=4, zipf=0.0, kind="er"):
    if kind == "er":
        g = erdos_renyi(n, m, seed)
    else:
        g = chung_lu(n, m, 2.1, seed)
    return set_labels(g, labels, zipf, seed)


# --- closed-form fixtures -------------------------------------------------------

def complete(n, label=1):
    a, b = np.triu_indices(n, 1)
    return from_edges(n, a, b, np.full(n, label))


def cycle(n, label=1):
    a = np.arange(n)
    return from_edges(n, a, (a + 1) % n, np.full(n, label))


def path_query(k, label=1):
    a = np.arange(k - 1)
    return Query(u32(np.full(k, label)), u32(a), u32(a + 1))


def cycle_query(k, label=1):
    a = np.arange(k)
    b = (a + 1) % k
    return Query(u32(np.full(k, label)), u32(np.minimum(a, b)), u32(np.maximum(a, b)))


def clique_query(k, label=1):
    a, b = np.triu_indices(k, 1)
    return Query(u32(np.full(k, label)), u32(a), u32(b))


def star_query(leaves, label=1):
    return Query(u32(np.full(leaves + 1, label)), u32(np.zeros(leaves)), u32(np.arange(1, leaves + 1)))
