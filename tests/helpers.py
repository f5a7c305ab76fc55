"""Shared inputs for the tests: small seeded graphs and queries."""
import numpy as np

from gpsm_amd import _lib
from gpsm_amd import graphs as G


def small_cases(n_cases=24, seed=0):
    """Random small ER / Chung-Lu graphs with 1-4 labels and walk queries."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n_cases):
        n = int(rng.integers(20, 300))
        m = int(rng.integers(n, 4 * n))
        kind = "er" if t % 2 == 0 else "cl"
        g = G.small_graph(1000 + t, n=n, m=min(m, n * (n - 1) // 4), labels=int(rng.integers(1, 5)), kind=kind)
        qkind = [_lib.Q_WALK_TREE, _lib.Q_WALK_INDUCED][t % 2]
        k = int(rng.integers(2, 7))
        qs = G.queries(g, 1, k, k, qkind, 77 + t)
        out.append((g, qs[0]))
    return out


def sorted_rows(a):
    a = np.asarray(a)
    if len(a) == 0:
        return a
    idx = np.lexsort(a.T[::-1])
    return a[idx]
