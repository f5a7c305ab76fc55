"""The R-MAT shape workload (cycles C4-C8, cliques K4-K8, induced walks of
8-50 vertices kept by the device-exact feasibility rule) on a scale-16 R-MAT
graph with the stress configuration's parameters (a, b, c, d = .57, .19,
.19, .05; 32 Zipf(0.8) labels): batched heuristic and learned orders and the
plain path against the host restatement O2, query by query.

O2 is the comparator here, not the independent enumerator O1: O1 extends
partial maps over whole adjacency lists with no candidate filtering, which
around R-MAT hubs runs for minutes on queries with a few million
embeddings; O2 is itself checked against O1 and brute force on smaller
graphs (test_oracle.py)."""
import pytest

import pyoracle as O
from gpsm_amd import _lib
from gpsm_amd import graphs as G
from gpsm_amd import planner
from gpsm_amd.workloads import feasible_queries

pytestmark = pytest.mark.gpu


def test_rmat_shapes_match_o2(ctx):
    g = G.rmat(16, 1_000_000, 0.57, 0.19, 0.19, 5)
    G.set_labels(g, 32, 0.8, 5)
    rg = ctx.upload(g)
    qs, dropped = [], 0
    for kind, kmin, kmax, seed in ((_lib.Q_CYCLE, 4, 8, 5), (_lib.Q_CLIQUE, 4, 8, 6), (_lib.Q_WALK_INDUCED, 8, 50, 7)):
        q, d, _ = feasible_queries(rg, 4, kmin, kmax, [kind], seed, 2 * 10 ** 6, per_round=6, rounds=3)
        qs += q
        dropped += d
    kinds = sorted({q.m > q.n - 1 for q in qs})
    print("%d queries (sizes %s), %d slots dropped" % (len(qs), [q.n for q in qs], dropped))
    assert len(qs) >= 4 and kinds == [True]  # cycles, cliques and induced walks all have cycles
    heur = rg.match_batch(qs, order_mode=_lib.ORDER_HEURISTIC)
    planner.install(ctx)
    learn = rg.match_batch(qs, order_mode=_lib.ORDER_LEARNED)
    for q, h, l in zip(qs, heur, learn):
        r2 = O.o2_match(g, q, mode=0)
        assert (h.count, h.hash) == (l.count, l.hash) == (r2["count"], r2["hash"]), q.n
        p = rg.match(q)
        assert (p.count, p.hash, p.join_work) == (h.count, h.hash, h.join_work)
    assert sum(h.count for h in heur) > 0
