"""The benchmark workloads, built from seeded generators.

* er10k_walk6: ER 10k vertices / 40k edges, 8 uniform labels, one 6-vertex
  walk query (the size the CPU enumerator finishes in milliseconds).
* chunglu1m_walk8_16: Chung-Lu 1M/8M (gamma 2.1), 16 Zipf(1.0) labels, 100
  walk queries drawn with 8-16 vertices, half tree-like, half induced.
* chunglu1m_mix4_32: the same graph with 256 walk queries drawn with 4-32
  vertices (matched with the heuristic or the learned join order).
* rmat22_64m + rmat_shape_queries: R-MAT scale 22 (a, b, c, d = .57, .19,
  .19, .05), 64M undirected edges, 32 Zipf(0.8) labels; cycles C4-C8,
  cliques K4-K8 and induced walks of up to 50 vertices.

Queries pass the feasibility rule of gpsm_gen_queries (an exact upper bound
on the embedding count, `budget`); slots shrink until it holds, so realised
sizes are reported with every result.
"""
from . import _lib
from . import graphs as G

EMBEDDING_BUDGET = 10 ** 8


def feasible_queries(rg, nq, kmin, kmax, kinds, seed, work_budget, per_slot=6):
    """Queries that pass an exact device-measured feasibility rule.

    For output slot i, per_slot candidates of kind kinds[i % len(kinds)] are
    drawn (sizes uniform in [kmin, kmax], no host-side rule), matched in one
    batch with a join work budget, and the first candidate whose join finishes
    within the budget is kept.  A slot whose candidates all exceed it keeps its
    smallest-work candidate's host-rule replacement (gpsm_gen_queries with the
    same budget as an embedding bound).  Every decision is an exact property
    of the query, so the set does not depend on the device or thread count.
    Returns (queries, number of slots that needed the host fallback)."""
    out, fallbacks = [], 0
    by_kind = {}
    for ki, kind in enumerate(kinds):
        slots = list(range(ki, nq, len(kinds)))
        cands = G.queries(rg.g, len(slots) * per_slot, kmin, kmax, kind, seed + 7919 * ki, 0)
        res = rg.match_batch(cands, work_budget=work_budget)
        for si, slot in enumerate(slots):
            pick = None
            for j in range(per_slot):
                if res[si * per_slot + j].status == 0:
                    pick = cands[si * per_slot + j]
                    break
            by_kind[slot] = pick
    for slot in range(nq):
        q = by_kind[slot]
        if q is None:
            fallbacks += 1
            kind = kinds[slot % len(kinds)]
            q = G.queries(rg.g, 1, kmin, kmax, kind, seed + 104729 * (slot + 1), work_budget)[0]
        out.append(q)
    return out, fallbacks


def er10k_walk6(seed=1):
    return G.er10k_walk6(seed)


def chunglu1m_walk8_16(seed=2, budget=EMBEDDING_BUDGET, nq=100):
    g = G.chunglu_1m(seed)
    return g, G.queries(g, nq, 8, 16, _lib.Q_WALK_HALF, seed, budget)


def chunglu1m_mix4_32(seed=2, budget=EMBEDDING_BUDGET, nq=256, g=None):
    g = g if g is not None else G.chunglu_1m(seed)
    return g, G.queries(g, nq, 4, 32, _lib.Q_WALK_HALF, seed + 1000, budget)


def rmat22_64m(seed=5, scale=22, m=64_000_000):
    g = G.rmat(scale, m, 0.57, 0.19, 0.19, seed)
    return G.set_labels(g, 32, 0.8, seed)


def rmat_shape_queries(g, seed=5, budget=EMBEDDING_BUDGET, per_kind=8):
    qs = []
    qs += G.queries(g, per_kind, 4, 8, _lib.Q_CYCLE, seed, budget)
    qs += G.queries(g, per_kind, 4, 8, _lib.Q_CLIQUE, seed + 1, budget)
    qs += G.queries(g, per_kind, 8, 50, _lib.Q_WALK_INDUCED, seed + 2, budget)
    return qs
