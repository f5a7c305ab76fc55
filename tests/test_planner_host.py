"""The learned planner's host logic on the CPU (gpsm_amd/csrc/host/plan.cpp):
candidate orders, the greedy search and the choice among candidates by
chained predicted cost.  The network is replaced by synthetic predictors
whose chained costs are recomputed here from first principles, and the
candidate list is checked against the mirror in train_planner (the orders
the training data covers)."""
import math
import os
import subprocess
import types

import numpy as np
import pytest

from gpsm_amd.train_planner import candidate_list

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MARGIN = 0.9  # kLearnedMargin in plan.h
BACK = 3.9  # kBackWeight in plan.h


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("planner") / "planner_driver")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "gpsm_amd", "csrc"), "-I", os.path.join(ROOT, "gpsm_amd", "csrc", "host"),
                    os.path.join(ROOT, "tests", "planner_driver.cpp"),
                    os.path.join(ROOT, "gpsm_amd", "csrc", "host", "plan.cpp"), "-o", exe], check=True)
    return exe


def random_query(rng):
    n = int(rng.integers(3, 13))
    pairs = set()
    for v in range(1, n):  # random spanning tree
        pairs.add((int(rng.integers(v)), v))
    extra = int(rng.integers(0, n))
    for _ in range(extra):
        a, b = sorted(int(x) for x in rng.choice(n, 2, replace=False))
        pairs.add((a, b))
    edges = sorted(pairs)
    rng.shuffle(edges)
    edges = [(b, a) if rng.random() < 0.5 else (a, b) for a, b in edges]
    labels = rng.integers(0, 8, n)
    cand = rng.integers(1, 5000, n)
    cet = rng.choice(10 ** 6, len(edges), replace=False) + 1
    return n, edges, labels, cand, cet


def run(driver, n, edges, labels, cand, cet, mode):
    lines = ["%d %d" % (n, len(edges)), " ".join(map(str, labels))]
    lines += ["%d %d" % e for e in edges]
    lines += [" ".join(map(str, cand)), " ".join(map(str, cet)), mode]
    p = subprocess.run([driver], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    out = {"cand": []}
    for line in p.stdout.splitlines():
        tag, *rest = line.split()
        if tag == "edges":
            out["edges"] = [tuple(int(x) for x in r.split(":")) for r in rest]
        elif tag == "cand":
            out["cand"].append([int(x) for x in rest])
        elif tag == "chosen":
            out["chosen"] = int(rest[0])
        else:
            out[tag] = [int(x) for x in rest]
    return out


def expected_cost(order, edges, cand, mode):
    """R_1 + sum over k = 2..n-1 of R_{k-1} * |CET(e_k)| / |C(parent)| * (1 +
    BACK * back edges of position k): R_1 = |CET| of the first edge, later R_k
    through the predictor, each fed the previous prediction; the final level
    adds its examined candidates only."""
    n = len(order)
    pos = {v: k for k, v in enumerate(order)}
    first = next(i for i, (a, b, _) in enumerate(edges) if {a, b} == {order[0], order[1]})
    rows = float(edges[first][2])
    cost, lr = rows, math.log1p(rows)
    for k in range(2, n):
        w = order[k]
        best, nb = None, 0
        for i, (a, b, s) in enumerate(edges):
            other = b if a == w else a if b == w else None
            if other is None or pos[other] >= k:
                continue
            nb += 1
            if best is None or s < edges[best][2]:
                best = i
        a, b, s = edges[best]
        parent = b if a == w else a
        cost += rows * s / max(int(cand[parent]), 1) * (1 + BACK * (nb - 1))  # no parallel edges here
        if k == n - 1:
            break
        if mode == "zero":
            y = 0.0
        elif mode == "cet":
            y = math.log1p(s)
        else:
            y = lr + math.log1p(s) - math.log1p(int(cand[parent]))
        y = max(0.0, y)
        rows, lr = math.expm1(y), y
    return cost


def connected_prefixes(order, edges):
    for k in range(1, len(order)):
        if not any({a, b} & {order[k]} and ({a, b} - {order[k]}) <= set(order[:k]) for a, b, _ in edges):
            return False
    return True


@pytest.mark.parametrize("seed", range(40))
@pytest.mark.parametrize("mode", ["zero", "cet", "chain"])
def test_plan_choice(driver, seed, mode):
    rng = np.random.default_rng(seed)
    n, edges, labels, cand, cet = random_query(rng)
    out = run(driver, n, edges, labels, cand, cet, mode)
    E = out["edges"]
    cands = out["cand"]
    # candidates: the heuristic first, valid and distinct
    assert cands[0] == out["heur"]
    assert len({tuple(c) for c in cands}) == len(cands)
    for od in cands + [out["greedy"]]:
        assert sorted(od) == list(range(n)) and connected_prefixes(od, E)
    # ... and exactly the continuations the training data covers
    ef = np.array([a for a, _, _ in E]), np.array([b for _, b, _ in E]), np.array([s for _, _, s in E])
    mirror = [out["heur"]]
    for od in candidate_list(types.SimpleNamespace(n=n), *ef):
        if od not in mirror:
            mirror.append(od)
    assert mirror == cands
    # the choice: cheapest by chained cost, the heuristic unless clearly beaten
    full = cands + ([out["greedy"]] if out["greedy"] not in cands else [])
    costs = [expected_cost(od, E, cand, mode) for od in full]
    best = min(range(len(full)), key=lambda i: (costs[i], i))
    if best and not costs[best] < MARGIN * costs[0]:
        best = 0
    near_tie = sum(abs(c - costs[best]) <= 1e-4 * max(costs[best], 1) for c in costs) > 1 or \
        abs(costs[best] - MARGIN * costs[0]) <= 1e-4 * max(costs[0], 1)
    assert out["plan"] == full[out["chosen"]]
    if near_tie:
        assert costs[out["chosen"]] <= costs[0] * (1 + 1e-4)
    else:
        assert out["chosen"] == best, (costs, out["chosen"], best)
