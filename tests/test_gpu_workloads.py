"""Batched workloads and the learned planner on the GPU.

* A mix of walk queries drawn with 4-32 vertices on a 100K-vertex Chung-Lu
  graph, batched with hipGraph-captured phases, against the
  independent CPU enumerator query by query.
* Learned vs heuristic join order: identical counts and hashes, and the
  trained network's effect on the weighted join work reported and bounded
  (at most 1.1x the heuristic's; the gain itself is measured on the R-MAT
  cliques in test_gpu_configs.py and on the benchmark mix by bench.py).
"""
import os

import numpy as np
import pytest

import pyoracle as O
from gpsm_amd import _lib
from gpsm_amd import graphs as G
from gpsm_amd import planner
from gpsm_amd.calibrate_planner import backs_per_level

BACK = 3.9  # kBackWeight in gpsm_amd/csrc/host/plan.h

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix():
    g = G.chung_lu(100_000, 800_000, 2.1, 77)
    G.set_labels(g, 16, 1.0, 77)
    qs = G.queries(g, 64, 4, 32, _lib.Q_WALK_HALF, 78, 2 * 10 ** 6)
    return g, qs


def test_batched_mix_matches_enumerator(ctx, mix):
    g, qs = mix
    rg = ctx.upload(g)
    res = rg.match_batch(qs, use_graph=True)
    for q, r in zip(qs, res):
        ref = O.o1_enumerate(g, q)
        assert (r.count, r.hash) == (ref["count"], ref["hash"]), q.n


def test_autotuner_keeps_results(ctx, mix):
    g, qs = mix
    rg = ctx.upload(g)
    before = rg.match_batch(qs)
    ms, tune = rg.autotune(qs[:32])
    assert ms > 0 and tune["filter_block"] in (64, 128, 256) and tune["tasks_per_wg"] in (8, 16, 32, 64)
    assert tune["join_wg_per_cu"] in (2, 4, 6, 8) and tune["join_cap"] in (2048, 4096, 8192)
    after = rg.match_batch(qs)
    assert [(r.count, r.hash) for r in before] == [(r.count, r.hash) for r in after]
    # a saved tuning is set back exactly; out-of-range values are refused
    ctx.set_tuning(128, 2048, 2, 32)
    assert ctx.tuning() == {"filter_block": 128, "join_cap": 2048, "join_wg_per_cu": 2, "tasks_per_wg": 32}
    with pytest.raises(_lib.GpsmError):
        ctx.set_tuning(100, 4096, 4, 8)
    with pytest.raises(_lib.GpsmError):
        ctx.set_tuning(256, 1024, 4, 8)
    small = rg.match_batch(qs)
    assert [(r.count, r.hash) for r in small] == [(r.count, r.hash) for r in before]
    ctx.set_tuning(128, 4096, 8, 32)  # the defaults, for the tests after this one


@pytest.mark.skipif(not os.path.exists(planner.WEIGHTS), reason="no trained planner weights")
def test_learned_order_same_results_and_bounded_join_work(ctx, mix):
    g, qs = mix
    rg = ctx.upload(g)
    heur = rg.match_batch(qs, order_mode=_lib.ORDER_HEURISTIC)
    planner.install(ctx)
    learn = rg.match_batch(qs, order_mode=_lib.ORDER_LEARNED)
    assert [(r.count, r.hash) for r in heur] == [(r.count, r.hash) for r in learn]
    # join work as the device counts it per level, each examined candidate
    # weighted by the edge-membership checks of its level (the planner's
    # cost unit, with the weight measured for a back-edge check)
    def weighted(q, r):
        tabs = rg.cet_edges(q)
        backs = backs_per_level(q, r.order, *tabs)
        return sum(r.level_work[k] * (1 + BACK * backs[k]) for k in range(1, q.n))

    wh = sum(weighted(q, r) for q, r in zip(qs, heur))
    wl = sum(weighted(q, r) for q, r in zip(qs, learn))
    jh = sum(r.join_work for r in heur)
    jl = sum(r.join_work for r in learn)
    changed = sum(h.order != l.order for h, l in zip(heur, learn))
    print("weighted join work: heuristic %d, learned %d (%.2fx); candidates %d vs %d; %d of %d plans changed"
          % (wh, wl, wh / max(wl, 1), jh, jl, changed, len(qs)))
    assert all(r.join_work > 0 for r in heur if r.count)
    # the heuristic order is always a candidate and is replaced only when the
    # network predicts a clearly cheaper plan
    assert wl <= 1.10 * wh
