// Seeded synthetic data graphs and queries (no reference counterpart: the
// reference reads its datasets from files, io/IO.cpp:46-100, and ships none).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <parallel/algorithm>

#include "gpsm.h"
#include "gpsm_gen.h"
#include "rng.h"

using gpsm::SplitMix64;

namespace {

// Open-addressing set of undirected edge keys (lo << 32 | hi), preserving the
// order of first insertion in `order`.
struct EdgeSet {
  std::vector<uint64_t> slot;
  uint64_t mask;
  std::vector<uint64_t> order;
  explicit EdgeSet(uint64_t m) {
    uint64_t cap = 1;
    while (cap < 2 * m + 16) cap <<= 1;
    slot.assign(cap, ~0ULL);
    mask = cap - 1;
    order.reserve(m);
  }
  bool insert(uint32_t a, uint32_t b) {
    if (a == b) return false;
    if (a > b) std::swap(a, b);
    uint64_t key = ((uint64_t)a << 32) | b;
    uint64_t h = gpsm::fmix64(key) & mask;
    while (slot[h] != ~0ULL) {
      if (slot[h] == key) return false;
      h = (h + 1) & mask;
    }
    slot[h] = key;
    order.push_back(key);
    return true;
  }
};

int csr_from_keys(uint32_t n, const std::vector<uint64_t>& keys, gpsm_csr* out) {
  std::vector<uint32_t> a(keys.size()), b(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) {
    a[i] = (uint32_t)(keys[i] >> 32);
    b[i] = (uint32_t)keys[i];
  }
  return gpsm_gen_csr_from_edges(n, keys.size(), a.data(), b.data(), nullptr, out);
}

inline bool has_edge(const gpsm_csr* g, uint32_t a, uint32_t b) {
  const uint32_t* lo = g->col + g->row_offset[a];
  const uint32_t* hi = g->col + g->row_offset[a + 1];
  return std::binary_search(lo, hi, b);
}

// Feasibility rule.  Screen: Knuth's estimator of the cost of a backtracking
// search (D. E. Knuth, "Estimating the efficiency of backtrack programs",
// Math. Comp. 29, 1975) that places the pattern's vertices in BFS order, each
// next to its parent's image, checking label, degree, injectivity and the
// other earlier edges; each probe follows one random root-to-leaf path and the
// product of branching factors along it, times the adjacency length scanned
// at each level, estimates the adjacency entries the whole search examines.
// Decision: an exact upper bound on the embedding count (tree_bound).  The
// probes draw from the slot's own stream, so the rule is deterministic.
struct Feasibility {
  const gpsm_csr* g;
  const std::vector<std::vector<uint32_t>>* by_label;  // label -> vertices
  // the pattern in search order: position -> label, degree, earlier neighbours
  // (positions, ascending; the first is the BFS parent)
  uint32_t k = 0;
  std::vector<uint32_t> plab, pdeg;
  std::vector<std::vector<uint32_t>> back;

  bool ok_cand(uint32_t i, uint32_t c, const std::vector<uint32_t>& f) const {
    if (g->labels[c] != plab[i] || g->row_offset[c + 1] - g->row_offset[c] < pdeg[i]) return false;
    for (uint32_t j = 0; j < i; ++j)
      if (f[j] == c) return false;
    for (size_t t = 1; t < back[i].size(); ++t)
      if (!has_edge(g, c, f[back[i][t]])) return false;
    return true;
  }

  // Search order: BFS (ascending neighbour ids) from the vertex with the
  // largest degree / label-frequency score, the start GpSM's filter uses.
  bool prepare(uint32_t k_, const uint32_t* qlab, const std::vector<std::pair<uint32_t, uint32_t>>& edges) {
    k = k_;
    std::vector<std::vector<uint32_t>> adj(k);
    for (auto& e : edges) {
      adj[e.first].push_back(e.second);
      adj[e.second].push_back(e.first);
    }
    uint32_t start = 0;
    double best = -1;
    for (uint32_t u = 0; u < k; ++u) {
      std::sort(adj[u].begin(), adj[u].end());
      size_t freq = qlab[u] < by_label->size() ? (*by_label)[qlab[u]].size() : 0;
      double s = freq ? (double)adj[u].size() / (double)freq : 1e300;
      if (s > best) {
        best = s;
        start = u;
      }
    }
    std::vector<uint32_t> ord{start}, pos(k, ~0u);
    pos[start] = 0;
    for (size_t h = 0; h < ord.size(); ++h)
      for (uint32_t w : adj[ord[h]])
        if (pos[w] == ~0u) {
          pos[w] = (uint32_t)ord.size();
          ord.push_back(w);
        }
    if (ord.size() != k) return false;
    plab.resize(k);
    pdeg.resize(k);
    back.assign(k, {});
    for (uint32_t i = 0; i < k; ++i) {
      plab[i] = qlab[ord[i]];
      pdeg[i] = (uint32_t)adj[ord[i]].size();
      for (uint32_t w : adj[ord[i]])
        if (pos[w] < i) back[i].push_back(pos[w]);
      std::sort(back[i].begin(), back[i].end());
    }
    return true;
  }

  double estimate(SplitMix64& rng, int probes) {
    if (plab[0] >= by_label->size()) return 0;
    const auto& roots = (*by_label)[plab[0]];
    if (roots.empty()) return 0;
    double total = 0;
    std::vector<uint32_t> f(k), cands;
    for (int p = 0; p < probes; ++p) {
      double weight = (double)roots.size();
      double cost = weight;
      f[0] = roots[rng.below(roots.size())];
      for (uint32_t i = 1; i < k; ++i) {
        uint32_t mp = f[back[i][0]];
        cost += weight * (g->row_offset[mp + 1] - g->row_offset[mp]);
        cands.clear();
        for (uint32_t e = g->row_offset[mp]; e < g->row_offset[mp + 1]; ++e)
          if (ok_cand(i, g->col[e], f)) cands.push_back(g->col[e]);
        if (cands.empty()) break;
        weight *= (double)cands.size();
        f[i] = cands[rng.below(cands.size())];
      }
      total += cost;
    }
    return total / probes;
  }

  // Upper bound on the embedding count: label- and degree-preserving
  // homomorphisms of the BFS spanning tree of the pattern, by dynamic
  // programming from the leaves (hom(T_u, v) = prod over children c of
  // sum over w in adj(v) of hom(T_c, w)).  Every embedding restricts to one
  // such homomorphism, so the count bounds the embeddings from above.
  double tree_bound() {
    std::vector<std::vector<uint32_t>> kids(k);
    for (uint32_t i = 1; i < k; ++i) kids[back[i][0]].push_back(i);
    std::vector<std::vector<float>> val(k);
    for (int64_t i = (int64_t)k - 1; i >= 0; --i) {
      val[i].assign(g->n, 0.0f);
      if (plab[i] >= by_label->size()) return 0;
      for (uint32_t v : (*by_label)[plab[i]]) {
        if (g->row_offset[v + 1] - g->row_offset[v] < pdeg[i]) continue;
        double prod = 1.0;
        for (uint32_t c : kids[i]) {
          double s = 0;
          for (uint32_t e = g->row_offset[v]; e < g->row_offset[v + 1]; ++e) s += val[c][g->col[e]];
          prod *= s;
          if (prod == 0) break;
        }
        val[i][v] = (float)std::min(prod, 1e30);
      }
      for (uint32_t c : kids[i]) std::vector<float>().swap(val[c]);
    }
    double tot = 0;
    for (float x : val[0]) tot += x;
    return tot;
  }

  bool feasible(uint32_t k_, const uint32_t* qlab, const std::vector<std::pair<uint32_t, uint32_t>>& edges,
                uint64_t budget_, SplitMix64& rng) {
    if (!prepare(k_, qlab, edges)) return false;
    if (estimate(rng, 64) > 4.0 * (double)budget_) return false;
    return tree_bound() <= (double)budget_;
  }
};

// Draws per slot before gpsm_gen_queries gives up (GPSM_EINVAL).
constexpr int kMaxDraws = 3200;

struct QueryOut {
  std::vector<uint32_t> labels;
  std::vector<std::pair<uint32_t, uint32_t>> edges;
};

bool walk_query(const gpsm_csr* g, SplitMix64& rng, uint32_t k, bool induced, QueryOut& q) {
  uint32_t start;
  int tries = 0;
  do {
    start = (uint32_t)rng.below(g->n);
    if (++tries > 1000) return false;
  } while (g->row_offset[start + 1] == g->row_offset[start]);
  std::vector<uint32_t> order{start};
  std::vector<std::pair<uint32_t, uint32_t>> tree;
  uint32_t cur = 0;  // index into order
  uint64_t steps = 0;
  while (order.size() < k && steps < 64ULL * k) {
    uint32_t v = order[cur];
    uint32_t d = g->row_offset[v + 1] - g->row_offset[v];
    uint32_t nb = g->col[g->row_offset[v] + rng.below(d)];
    uint32_t idx = (uint32_t)(std::find(order.begin(), order.end(), nb) - order.begin());
    if (idx == order.size()) {
      order.push_back(nb);
      tree.push_back({cur, idx});
    }
    cur = idx;
    ++steps;
  }
  if (order.size() < k) return false;
  q.labels.resize(k);
  for (uint32_t i = 0; i < k; ++i) q.labels[i] = g->labels[order[i]];
  q.edges.clear();
  if (induced) {
    for (uint32_t i = 0; i < k; ++i)
      for (uint32_t j = i + 1; j < k; ++j)
        if (has_edge(g, order[i], order[j])) q.edges.push_back({i, j});
  } else {
    for (auto& e : tree) q.edges.push_back({std::min(e.first, e.second), std::max(e.first, e.second)});
    std::sort(q.edges.begin(), q.edges.end());
  }
  return true;
}

bool cycle_query(const gpsm_csr* g, SplitMix64& rng, uint32_t k, QueryOut& q) {
  if (g->row_offset[g->n] == 0 || k < 3) return false;
  for (int attempt = 0; attempt < 20000; ++attempt) {
    // degree-biased start: endpoint of a random adjacency entry
    uint32_t start = g->col[rng.below(g->row_offset[g->n])];
    std::vector<uint32_t> path{start};
    bool ok = true;
    while (path.size() < k && ok) {
      uint32_t v = path.back();
      uint32_t d = g->row_offset[v + 1] - g->row_offset[v];
      ok = false;
      for (int t = 0; t < 8; ++t) {
        uint32_t nb = g->col[g->row_offset[v] + rng.below(d)];
        if (std::find(path.begin(), path.end(), nb) == path.end()) {
          path.push_back(nb);
          ok = true;
          break;
        }
      }
    }
    if (!ok || !has_edge(g, path.back(), start)) continue;
    q.labels.resize(k);
    for (uint32_t i = 0; i < k; ++i) q.labels[i] = g->labels[path[i]];
    q.edges.clear();
    for (uint32_t i = 0; i + 1 < k; ++i) q.edges.push_back({i, i + 1});
    q.edges.push_back({0, k - 1});
    std::sort(q.edges.begin(), q.edges.end());
    return true;
  }
  return false;
}

bool clique_query(const gpsm_csr* g, SplitMix64& rng, uint32_t k, QueryOut& q) {
  if (g->row_offset[g->n] == 0) return false;
  for (int attempt = 0; attempt < 20000; ++attempt) {
    uint32_t v0 = g->col[rng.below(g->row_offset[g->n])];
    if (g->row_offset[v0 + 1] - g->row_offset[v0] < k - 1) continue;
    std::vector<uint32_t> members{v0};
    std::vector<uint32_t> cand(g->col + g->row_offset[v0], g->col + g->row_offset[v0 + 1]);
    while (members.size() < k && !cand.empty()) {
      uint32_t pick = cand[rng.below(cand.size())];
      members.push_back(pick);
      std::vector<uint32_t> next;
      const uint32_t* lo = g->col + g->row_offset[pick];
      const uint32_t* hi = g->col + g->row_offset[pick + 1];
      for (uint32_t c : cand)
        if (c != pick && std::binary_search(lo, hi, c)) next.push_back(c);
      cand.swap(next);
    }
    if (members.size() < k) continue;
    q.labels.resize(k);
    for (uint32_t i = 0; i < k; ++i) q.labels[i] = g->labels[members[i]];
    q.edges.clear();
    for (uint32_t i = 0; i < k; ++i)
      for (uint32_t j = i + 1; j < k; ++j) q.edges.push_back({i, j});
    return true;
  }
  return false;
}

}  // namespace

extern "C" {

int gpsm_gen_csr_from_edges(uint32_t n, uint64_t m, const uint32_t* a, const uint32_t* b,
                            const uint32_t* labels, gpsm_csr* out) {
  if (!out || (m && (!a || !b))) return GPSM_EINVAL;
  std::memset(out, 0, sizeof(*out));
  std::vector<uint64_t> deg(n + 1, 0);
  for (uint64_t i = 0; i < m; ++i) {
    if (a[i] >= n || b[i] >= n) return GPSM_EINVAL;
    if (a[i] == b[i]) continue;
    deg[a[i]]++;
    deg[b[i]]++;
  }
  std::vector<uint64_t> off(n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) off[v + 1] = off[v] + deg[v];
  if (off[n] > 0xFFFFFFFFULL) return GPSM_EINVAL;
  std::vector<uint32_t> tmp(off[n]);
  std::vector<uint64_t> pos(off.begin(), off.end() - 1);
  for (uint64_t i = 0; i < m; ++i) {
    if (a[i] == b[i]) continue;
    tmp[pos[a[i]]++] = b[i];
    tmp[pos[b[i]]++] = a[i];
  }
  std::vector<uint32_t> newdeg(n);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; ++v) {
    uint32_t* lo = tmp.data() + off[v];
    uint32_t* hi = tmp.data() + off[v + 1];
    std::sort(lo, hi);
    newdeg[v] = (uint32_t)(std::unique(lo, hi) - lo);
  }
  out->n = n;
  out->row_offset = (uint32_t*)std::malloc(sizeof(uint32_t) * (n + 1));
  out->labels = (uint32_t*)std::malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!out->row_offset || !out->labels) return GPSM_ENOMEM;
  out->row_offset[0] = 0;
  for (uint32_t v = 0; v < n; ++v) out->row_offset[v + 1] = out->row_offset[v] + newdeg[v];
  out->adj = out->row_offset[n];
  out->col = (uint32_t*)std::malloc(sizeof(uint32_t) * (out->adj ? out->adj : 1));
  if (!out->col) return GPSM_ENOMEM;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; ++v)
    std::memcpy(out->col + out->row_offset[v], tmp.data() + off[v], sizeof(uint32_t) * newdeg[v]);
  for (uint32_t v = 0; v < n; ++v) out->labels[v] = labels ? labels[v] : 1;
  return GPSM_OK;
}

int gpsm_gen_er(uint32_t n, uint64_t m, uint64_t seed, gpsm_csr* out) {
  if (n < 2 || m > (uint64_t)n * (n - 1) / 2) return GPSM_EINVAL;
  SplitMix64 rng(seed, 1);
  EdgeSet es(m);
  while (es.order.size() < m) es.insert((uint32_t)rng.below(n), (uint32_t)rng.below(n));
  return csr_from_keys(n, es.order, out);
}

int gpsm_gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint64_t seed, gpsm_csr* out) {
  if (n < 2 || gamma <= 1.0 || m > (uint64_t)n * (n - 1) / 4) return GPSM_EINVAL;
  const double expo = -1.0 / (gamma - 1.0);
  std::vector<double> w(n);
  double sum = 0;
  for (uint32_t i = 0; i < n; ++i) sum += (w[i] = std::pow((double)i + 1.0, expo));
  const double scale = 2.0 * (double)m / sum;
  const double cap = std::sqrt(2.0 * (double)m);
  std::vector<double> cdf(n);
  double acc = 0;
  for (uint32_t i = 0; i < n; ++i) {
    acc += std::min(w[i] * scale, cap);
    cdf[i] = acc;
  }
  SplitMix64 rng(seed, 2);
  EdgeSet es(m);
  auto draw = [&]() -> uint32_t {
    double x = rng.uniform() * acc;
    uint32_t i = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    return i < n ? i : n - 1;
  };
  uint64_t guard = 0;
  while (es.order.size() < m) {
    uint32_t a = draw(), b = draw();
    es.insert(a, b);
    if (++guard > 64 * m + 1000000) return GPSM_EINVAL;  // weights too skewed to reach m distinct edges
  }
  return csr_from_keys(n, es.order, out);
}

int gpsm_gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed,
                  gpsm_csr* out) {
  if (scale < 1 || scale > 31 || a + b + c >= 1.0) return GPSM_EINVAL;
  const uint32_t n = 1u << scale;
  const double ab = a + b, abc = a + b + c;
  // Rounds of candidate edges drawn in fixed chunks, each from its own
  // stream (seed, round, chunk), then sorted and deduplicated: the edge set
  // does not depend on the thread count.  If a round overshoots, the m kept
  // edges are those of smallest fmix64(key ^ seed) (a uniform subset).
  const uint64_t CH = 1 << 20;
  std::vector<uint64_t> keys;
  for (uint32_t round = 0; keys.size() < m; ++round) {
    if (round > 64) return GPSM_EINVAL;  // too dense to reach m
    const uint64_t want = (m - keys.size()) + (m - keys.size()) / 5 + 1024;
    const uint64_t nch = (want + CH - 1) / CH;
    const size_t base = keys.size();
    keys.resize(base + nch * CH, ~0ULL);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < (int64_t)nch; ++ch) {
      SplitMix64 rng(seed, ((uint64_t)(round + 1) << 40) | (uint64_t)ch);
      uint64_t* dst = keys.data() + base + (uint64_t)ch * CH;
      for (uint64_t i = 0; i < CH; ++i) {
        uint32_t u = 0, v = 0;
        for (uint32_t l = 0; l < scale; ++l) {
          const double r = rng.uniform();
          const uint32_t bu = r >= ab, bv = (r >= a && r < ab) || r >= abc;
          u = (u << 1) | bu;
          v = (v << 1) | bv;
        }
        dst[i] = u == v ? ~0ULL : (((uint64_t)std::min(u, v) << 32) | std::max(u, v));
      }
    }
    __gnu_parallel::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (!keys.empty() && keys.back() == ~0ULL) keys.pop_back();
  }
  if (keys.size() > m) {
    std::vector<std::pair<uint64_t, uint64_t>> pri(keys.size());
#pragma omp parallel for
    for (int64_t i = 0; i < (int64_t)keys.size(); ++i) pri[i] = {gpsm::fmix64(keys[i] ^ seed), keys[i]};
    std::nth_element(pri.begin(), pri.begin() + m, pri.end());
    keys.resize(m);
    for (uint64_t i = 0; i < m; ++i) keys[i] = pri[i].second;
    __gnu_parallel::sort(keys.begin(), keys.end());
  }
  return csr_from_keys(n, keys, out);
}

int gpsm_gen_labels(gpsm_csr* g, uint32_t L, double zipf_s, uint64_t seed) {
  if (!g || L == 0) return GPSM_EINVAL;
  SplitMix64 rng(seed, 4);
  if (zipf_s == 0.0) {
    for (uint32_t v = 0; v < g->n; ++v) g->labels[v] = 1 + (uint32_t)rng.below(L);
    return GPSM_OK;
  }
  std::vector<double> cdf(L);
  double acc = 0;
  for (uint32_t k = 0; k < L; ++k) cdf[k] = (acc += std::pow((double)(k + 1), -zipf_s));
  for (uint32_t v = 0; v < g->n; ++v) {
    double x = rng.uniform() * acc;
    uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    g->labels[v] = 1 + (k < L ? k : L - 1);
  }
  return GPSM_OK;
}

int gpsm_gen_queries(const gpsm_csr* g, uint32_t nq, uint32_t kmin, uint32_t kmax, int kind,
                     uint64_t seed, uint64_t node_budget, gpsm_qset* out) {
  if (!g || !out || kmin < 2 || kmax < kmin || kmax > GPSM_MAX_QV) return GPSM_EINVAL;
  std::memset(out, 0, sizeof(*out));
  std::vector<QueryOut> qs(nq);
  std::vector<int> fail(nq, 0);
  uint32_t maxlab = 0;
  for (uint32_t v = 0; v < g->n; ++v) maxlab = std::max(maxlab, g->labels[v]);
  std::vector<std::vector<uint32_t>> by_label(maxlab + 1);
  for (uint32_t v = 0; v < g->n; ++v) by_label[g->labels[v]].push_back(v);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t slot = 0; slot < (int64_t)nq; ++slot) {
    SplitMix64 rng(seed, 1000 + (uint64_t)slot);
    int qkind = kind == GPSM_Q_WALK_HALF ? ((slot & 1) ? GPSM_Q_WALK_INDUCED : GPSM_Q_WALK_TREE) : kind;
    Feasibility fz;
    fz.g = g;
    fz.by_label = &by_label;
    bool done = false;
    // every draw takes a fresh size, uniform in [kmin, kmax]: a slot's query
    // follows the size distribution conditioned on passing the rule, and no
    // size is ever shrunk below kmin
    for (int attempt = 0; attempt < kMaxDraws && !done; ++attempt) {
      const uint32_t k = kmin + (uint32_t)rng.below(kmax - kmin + 1);
      QueryOut q;
      bool ok;
      if (qkind == GPSM_Q_CYCLE) ok = cycle_query(g, rng, k, q);
      else if (qkind == GPSM_Q_CLIQUE) ok = clique_query(g, rng, k, q);
      else ok = walk_query(g, rng, k, qkind == GPSM_Q_WALK_INDUCED, q);
      if (!ok) continue;
      if (node_budget && !fz.feasible(k, q.labels.data(), q.edges, node_budget, rng)) continue;
      qs[slot] = std::move(q);
      done = true;
    }
    if (!done) fail[slot] = 1;
  }
  for (uint32_t i = 0; i < nq; ++i)
    if (fail[i]) return GPSM_EINVAL;
  uint64_t nv = 0, ne = 0;
  for (auto& q : qs) {
    nv += q.labels.size();
    ne += q.edges.size();
  }
  out->nq = nq;
  out->v_off = (uint32_t*)std::malloc(sizeof(uint32_t) * (nq + 1));
  out->e_off = (uint32_t*)std::malloc(sizeof(uint32_t) * (nq + 1));
  out->labels = (uint32_t*)std::malloc(sizeof(uint32_t) * (nv + 1));
  out->src = (uint32_t*)std::malloc(sizeof(uint32_t) * (ne + 1));
  out->dst = (uint32_t*)std::malloc(sizeof(uint32_t) * (ne + 1));
  if (!out->v_off || !out->e_off || !out->labels || !out->src || !out->dst) return GPSM_ENOMEM;
  out->v_off[0] = out->e_off[0] = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    auto& q = qs[i];
    std::memcpy(out->labels + out->v_off[i], q.labels.data(), sizeof(uint32_t) * q.labels.size());
    for (size_t e = 0; e < q.edges.size(); ++e) {
      out->src[out->e_off[i] + e] = q.edges[e].first;
      out->dst[out->e_off[i] + e] = q.edges[e].second;
    }
    out->v_off[i + 1] = out->v_off[i] + (uint32_t)q.labels.size();
    out->e_off[i + 1] = out->e_off[i] + (uint32_t)q.edges.size();
  }
  return GPSM_OK;
}

void gpsm_gen_free_csr(gpsm_csr* g) {
  if (!g) return;
  std::free(g->row_offset);
  std::free(g->col);
  std::free(g->labels);
  std::memset(g, 0, sizeof(*g));
}

void gpsm_gen_free_qset(gpsm_qset* s) {
  if (!s) return;
  std::free(s->v_off);
  std::free(s->labels);
  std::free(s->e_off);
  std::free(s->src);
  std::free(s->dst);
  std::memset(s, 0, sizeof(*s));
}

int gpsm_io_write_graph(const char* path, int append, int gid, uint32_t n, const uint32_t* labels,
                        uint64_t m, const uint32_t* src, const uint32_t* dst) {
  FILE* fp = std::fopen(path, append ? "a" : "w");
  if (!fp) return GPSM_EINVAL;
  std::fprintf(fp, "t # %d\n%u %llu %u %u\n", gid, n, (unsigned long long)m, 0u, 0u);
  for (uint32_t v = 0; v < n; ++v) std::fprintf(fp, "v %u %u\n", v, labels[v]);
  for (uint64_t e = 0; e < m; ++e) std::fprintf(fp, "e %u %u 0\n", src[e], dst[e]);
  std::fclose(fp);
  return GPSM_OK;
}

int gpsm_io_write_end(const char* path) {
  FILE* fp = std::fopen(path, "a");
  if (!fp) return GPSM_EINVAL;
  std::fprintf(fp, "t # -1\n");
  std::fclose(fp);
  return GPSM_OK;
}

uint64_t gpsm_tuple_hash(const uint32_t* t, uint32_t q) {
  uint64_t h = 1;
  for (uint32_t j = 0; j < q; ++j) h *= gpsm::tuple_term(j, t[j]);
  return h;
}

}  // extern "C"
