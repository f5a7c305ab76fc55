// Host-side planning of one query (see plan.h).
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <queue>

#include "gpsm_dev.h"

namespace gpsm {

uint32_t GraphStats::dense(uint32_t label) const {
  auto it = std::lower_bound(label_value.begin(), label_value.end(), label);
  if (it == label_value.end() || *it != label) return 0;
  return (uint32_t)(it - label_value.begin()) + 1;
}

uint32_t GraphStats::count(uint32_t label) const {
  uint32_t d = dense(label);
  return d ? label_count[d - 1] : 0;
}

int HostQuery::build(const gpsm_query* q) {
  if (!q || q->n == 0 || !q->labels || (q->m && (!q->src || !q->dst))) return GPSM_EINVAL;
  if (q->n > GPSM_MAX_QV) return GPSM_ETOOBIG;
  n = q->n;
  lab.assign(q->labels, q->labels + n);
  in.assign(n, {});
  out.assign(n, {});
  for (uint32_t e = 0; e < q->m; ++e) {
    uint32_t a = q->src[e], b = q->dst[e];
    if (a >= n || b >= n || a == b) return GPSM_EINVAL;
    out[a].push_back(b);  // Graph::addEdge, graph/Graph.cpp:18-22
    in[b].push_back(a);
  }
  ro.assign(n + 1, 0);
  for (uint32_t u = 0; u < n; ++u) ro[u + 1] = ro[u] + (uint32_t)(in[u].size() + out[u].size());
  ci.resize(ro[n]);
  for (uint32_t u = 0; u < n; ++u) {
    uint32_t p = ro[u];
    for (uint32_t x : in[u]) ci[p++] = x;
    for (uint32_t x : out[u]) ci[p++] = x;
    std::sort(ci.begin() + ro[u], ci.begin() + ro[u + 1]);
  }
  ef.clear();
  et.clear();
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t src : in[i]) {
      ef.push_back(std::min(src, i));
      et.push_back(std::max(src, i));
    }
  if (ef.empty()) return GPSM_EDISCONN;
  // connectivity
  std::vector<uint8_t> seen(n, 0);
  std::vector<uint32_t> st{0};
  seen[0] = 1;
  uint32_t cnt = 1;
  while (!st.empty()) {
    uint32_t u = st.back();
    st.pop_back();
    for (uint32_t i = ro[u]; i < ro[u + 1]; ++i)
      if (!seen[ci[i]]) {
        seen[ci[i]] = 1;
        ++cnt;
        st.push_back(ci[i]);
      }
  }
  return cnt == n ? GPSM_OK : GPSM_EDISCONN;
}

void key_range(const HostQuery& q, const GraphStats& s, uint32_t u, uint32_t& lo, uint32_t& hi) {
  uint32_t d = s.dense(q.lab[u]);
  if (d == 0) {  // label absent from the data graph: empty range
    lo = 1;
    hi = 0;
    return;
  }
  uint32_t deg = std::min(q.deg(u), KEY_DEG_MASK);
  lo = (d << KEY_DEG_BITS) | deg;
  hi = (d << KEY_DEG_BITS) | KEY_DEG_MASK;
}

void make_filter_schedule(const HostQuery& q, const GraphStats& s, FilterSchedule& out) {
  out.steps.clear();
  // generateSpanningTree, match/Match.cu:347-389
  std::vector<float> score(q.n);
  for (uint32_t i = 0; i < q.n; ++i) {
    float num = (float)s.count(q.lab[i]);
    int degree = (int)q.deg(i);
    score[i] = (float)((degree + 0.0) / num);
  }
  float maxv = 0.0f;
  int maxi = 0;
  for (size_t i = 0; i < q.ef.size(); ++i) {
    float tmp = score[q.ef[i]] + score[q.et[i]];
    if (tmp > maxv) {
      maxv = tmp;
      maxi = (int)i;
    }
  }
  uint32_t start = q.ef[maxi];
  if (score[q.et[maxi]] > score[start]) start = q.et[maxi];
  // BFS over the sorted query CSR, match/Match.cu:391-424
  std::vector<uint8_t> visit(q.n, 0);
  out.filter_order.assign(1, start);
  visit[start] = 1;
  std::queue<uint32_t> tq;
  tq.push(start);
  while (!tq.empty()) {
    uint32_t t = tq.front();
    tq.pop();
    for (uint32_t i = q.ro[t]; i < q.ro[t + 1]; ++i) {
      uint32_t adj = q.ci[i];
      if (visit[adj]) continue;
      tq.push(adj);
      out.filter_order.push_back(adj);
      visit[adj] = 1;
    }
  }
  out.start = start;
  key_range(q, s, start, out.start_lo, out.start_hi);
  // InitializeCandidateVertices, match/Match.cu:670-715: explore every vertex
  // in filter order towards its not yet initialised neighbours.
  std::vector<uint8_t> init(q.n, 0);
  init[start] = 1;
  for (uint32_t u : out.filter_order) {
    FilterStep st;
    st.type = OP_EXPLORE;
    st.u = u;
    for (uint32_t i = q.ro[u]; i < q.ro[u + 1]; ++i) {
      uint32_t v = q.ci[i];
      if (init[v]) continue;
      uint32_t lo, hi;
      key_range(q, s, v, lo, hi);
      st.v.push_back(v);
      st.lo.push_back(lo);
      st.hi.push_back(hi);
    }
    if (!st.v.empty()) out.steps.push_back(std::move(st));
    for (uint32_t i = q.ro[u]; i < q.ro[u + 1]; ++i) init[q.ci[i]] = 1;
  }
  // generateSimplifiedGraph + RefineCandidateVertices, match/Match.cu:721-890:
  // vertices of degree > 1, neighbours from the in list then the out list,
  // four refining rounds in id order (the fifth only checks emptiness).
  std::vector<uint8_t> valid(q.n, 0);
  for (uint32_t i = 0; i < q.n; ++i) valid[i] = q.deg(i) > 1;
  for (int round = 0; round < 4; ++round)
    for (uint32_t u = 0; u < q.n; ++u) {
      if (!valid[u]) continue;
      FilterStep st;
      st.type = OP_REFINE;
      st.u = u;
      st.round = (uint32_t)round;
      for (uint32_t id : q.in[u])
        if (valid[id]) st.v.push_back(id);
      for (uint32_t id : q.out[u])
        if (valid[id]) st.v.push_back(id);
      if (!st.v.empty()) out.steps.push_back(std::move(st));
    }
}

namespace {

struct Builder {
  const HostQuery& q;
  const std::vector<uint64_t>& size;
  OrderPlan& out;
  std::vector<int> pos;
  std::vector<uint8_t> used_e;
  Builder(const HostQuery& q_, const std::vector<uint64_t>& s, OrderPlan& o) : q(q_), size(s), out(o) {
    pos.assign(q.n, -1);
    used_e.assign(q.ef.size(), 0);
    out.order.clear();
    out.par_pos.assign(q.n, 0);
    out.par_edge.assign(q.n, 0);
    out.backs.assign(q.n, {});
  }
  void place(uint32_t v) {
    pos[v] = (int)out.order.size();
    out.order.push_back(v);
  }
  // After placing v as position k (k >= 2) through edge pe: its remaining
  // edges to placed vertices are join-mode edges.
  void close(uint32_t v, uint32_t pe) {
    uint32_t k = (uint32_t)pos[v];
    used_e[pe] = 1;
    for (size_t i = 0; i < q.ef.size(); ++i) {
      if (used_e[i]) continue;
      uint32_t a = q.ef[i], b = q.et[i];
      if (pos[a] >= 0 && pos[b] >= 0) {
        used_e[i] = 1;
        uint32_t other = (a == v) ? b : a;
        if (a == v || b == v) out.backs[k].push_back((uint32_t)pos[other]);
      }
    }
    std::sort(out.backs[k].begin(), out.backs[k].end());
    out.backs[k].erase(std::unique(out.backs[k].begin(), out.backs[k].end()), out.backs[k].end());
    // a parallel query edge to the parent is a membership check of the same pair
    auto it = std::find(out.backs[k].begin(), out.backs[k].end(), out.par_pos[k]);
    if (it != out.backs[k].end()) out.backs[k].erase(it);
  }
};

}  // namespace

void heuristic_order(const HostQuery& q, const std::vector<uint64_t>& cet_size, uint64_t big, OrderPlan& out) {
  Builder b(q, cet_size, out);
  const size_t E = q.ef.size();
  uint64_t minv = big;
  uint32_t minx = 0;
  for (uint32_t i = 0; i < E; ++i)
    if (cet_size[i] < minv) {
      minv = cet_size[i];
      minx = i;
    }
  out.first_edge = minx;
  b.place(q.ef[minx]);
  b.place(q.et[minx]);
  b.used_e[minx] = 1;
  // edges parallel to the first one
  for (uint32_t i = 0; i < E; ++i)
    if (!b.used_e[i] && b.pos[q.ef[i]] >= 0 && b.pos[q.et[i]] >= 0) b.used_e[i] = 1;
  while (out.order.size() < q.n) {
    uint64_t nminv = big;
    uint32_t nminx = 0;
    bool found = false;
    for (uint32_t i = 0; i < E; ++i) {
      if (b.used_e[i]) continue;
      bool fa = b.pos[q.ef[i]] >= 0, fb = b.pos[q.et[i]] >= 0;
      if (fa == fb) continue;
      if (!found || cet_size[i] < nminv) {
        nminv = cet_size[i];
        nminx = i;
        found = true;
      }
    }
    if (!found) break;  // disconnected (rejected by HostQuery::build)
    uint32_t u = q.ef[nminx], v = q.et[nminx];
    uint32_t parent = b.pos[u] >= 0 ? u : v, child = b.pos[u] >= 0 ? v : u;
    b.place(child);
    uint32_t k = (uint32_t)b.pos[child];
    out.par_pos[k] = (uint32_t)b.pos[parent];
    out.par_edge[k] = nminx;
    b.close(child, nminx);
  }
}

bool order_from_vertices(const HostQuery& q, const std::vector<uint64_t>& cet_size, const uint32_t* order,
                         OrderPlan& out) {
  Builder b(q, cet_size, out);
  const size_t E = q.ef.size();
  std::vector<uint8_t> seen(q.n, 0);
  for (uint32_t i = 0; i < q.n; ++i) {
    if (order[i] >= q.n || seen[order[i]]) return false;
    seen[order[i]] = 1;
  }
  uint32_t a = order[0], c = order[1];
  bool found = false;
  for (uint32_t i = 0; i < E && !found; ++i)
    if ((q.ef[i] == a && q.et[i] == c) || (q.ef[i] == c && q.et[i] == a)) {
      out.first_edge = i;
      found = true;
    }
  if (!found) return false;
  b.place(a);
  b.place(c);
  for (uint32_t i = 0; i < E; ++i)
    if (b.pos[q.ef[i]] >= 0 && b.pos[q.et[i]] >= 0) b.used_e[i] = 1;
  for (uint32_t k = 2; k < q.n; ++k) {
    uint32_t w = order[k];
    uint64_t best = 0;
    int bi = -1;
    for (uint32_t i = 0; i < E; ++i) {
      if (b.used_e[i]) continue;
      bool hit = (q.ef[i] == w && b.pos[q.et[i]] >= 0) || (q.et[i] == w && b.pos[q.ef[i]] >= 0);
      if (hit && (bi < 0 || cet_size[i] < best)) {
        best = cet_size[i];
        bi = (int)i;
      }
    }
    if (bi < 0) return false;
    uint32_t parent = q.ef[bi] == w ? q.et[bi] : q.ef[bi];
    b.place(w);
    out.par_pos[k] = (uint32_t)b.pos[parent];
    out.par_edge[k] = (uint32_t)bi;
    b.close(w, (uint32_t)bi);
  }
  return true;
}

// ---------------------------------------------------------------------------
// Learned planner features.  All logs are natural logs of (1 + x).
namespace {
inline float L1p(double x) { return (float)std::log1p(std::max(x, 0.0)); }
}

void step_features(const PlanContext& c, const std::vector<uint8_t>& placed, uint32_t nplaced, uint32_t edge,
                   uint32_t child, double log_rows, float* f) {
  const HostQuery& q = *c.q;
  const auto& cand = *c.cand;
  const auto& cs = *c.cet_size;
  uint32_t parent = q.ef[edge] == child ? q.et[edge] : q.ef[edge];
  double lC_w = L1p(cand[child]), lC_x = L1p(cand[parent]), lE = L1p((double)cs[edge]);
  uint32_t nback = 0, unplaced_nb = 0, placed_nb = 0;
  double bmin = 0, bmax = 0, bsum = 0, bsel = 0;
  for (size_t i = 0; i < q.ef.size(); ++i) {
    uint32_t a = q.ef[i], b = q.et[i];
    if (a != child && b != child) continue;
    uint32_t other = a == child ? b : a;
    if (!placed[other]) {
      ++unplaced_nb;
      continue;
    }
    ++placed_nb;
    if (i == edge) continue;
    double le = L1p((double)cs[i]);
    bmin = nback ? std::min(bmin, le) : le;
    bmax = nback ? std::max(bmax, le) : le;
    bsum += le;
    bsel += le - L1p(cand[other]) - lC_w;
    ++nback;
  }
  const double n = c.s->n;
  double lsel = lE - lC_x - lC_w;
  std::memset(f, 0, sizeof(float) * GPSM_FEAT);
  f[0] = (float)log_rows;
  f[1] = (float)lC_w;
  f[2] = (float)lE;
  f[3] = (float)lC_x;
  f[4] = (float)(lE - lC_x);
  f[5] = (float)nback;
  f[6] = (float)bmin;
  f[7] = (float)bmax;
  f[8] = (float)bsel;
  f[9] = (float)lsel;
  f[10] = (float)q.deg(child);
  f[11] = (float)unplaced_nb;
  f[12] = (float)std::log((double)std::max<uint32_t>(c.s->count(q.lab[child]), 1) / n);
  f[13] = (float)std::log((double)std::max<uint32_t>(c.s->count(q.lab[parent]), 1) / n);
  f[14] = (float)q.deg(parent);
  f[15] = (float)nplaced / (float)q.n;
  f[16] = (float)q.n / 64.0f;
  f[17] = (float)q.ef.size() / (float)q.n;
  f[18] = (float)std::log(n);
  f[19] = (float)std::log1p((double)c.s->adj / n);
  f[20] = (float)std::min(q.deg(child), q.deg(parent));
  f[21] = (float)std::max(q.deg(child), q.deg(parent));
  f[22] = (float)bsum;
  f[23] = q.lab[child] == q.lab[parent] ? 1.0f : 0.0f;
  f[24] = (float)placed_nb / (float)std::max<uint32_t>(q.deg(child), 1);
  f[25] = (float)(lC_w - std::log(n));
  f[26] = (float)(log_rows + lE - lC_x);          // rows x mean fan-out
  f[27] = (float)(log_rows + lE - lC_x + bsel);   // ... x back-edge densities
  f[28] = (float)(q.n - nplaced) / (float)q.n;
  f[29] = 1.0f;
}

void first_features(const PlanContext& c, uint32_t edge, uint32_t from, float* f) {
  std::vector<uint8_t> placed(c.q->n, 0);
  placed[from] = 1;
  uint32_t to = c.q->ef[edge] == from ? c.q->et[edge] : c.q->ef[edge];
  step_features(c, placed, 1, edge, to, L1p(c.cand->at(from)), f);
}

// The greedy search as a state machine, so a batch of queries can share one
// network call per decision round: learned_emit appends the feature rows of
// the pending decision, learned_consume takes their predictions.
//   round 0: the first edge and its orientation, minimising |CET| (the first
//            table's rows are known exactly) + min predicted rows of the
//            level after it;
//   later:   among unplaced vertices adjacent to the prefix (each through its
//            smallest-|CET| edge), the one with the smallest predicted rows.
void learned_start(const PlanContext& c, LearnedState& s) {
  s.c = &c;
  s.done = false;
  s.round = 0;
  s.order.clear();
  s.placed.assign(c.q->n, 0);
}

size_t learned_emit(LearnedState& s, std::vector<float>& feats) {
  const PlanContext& c = *s.c;
  const HostQuery& q = *c.q;
  const auto& cs = *c.cet_size;
  const size_t E = q.ef.size();
  const size_t row0 = feats.size() / GPSM_FEAT;
  if (s.round == 0) {
    s.first.clear();
    s.span.clear();
    for (uint32_t e = 0; e < E; ++e)
      for (int side = 0; side < 2; ++side) {
        const uint32_t a = side ? q.et[e] : q.ef[e], b = side ? q.ef[e] : q.et[e];
        std::vector<uint8_t> placed(q.n, 0);
        placed[a] = placed[b] = 1;
        const size_t begin = feats.size() / GPSM_FEAT - row0;
        const double lr = std::log1p((double)cs[e]);
        for (uint32_t i = 0; i < E; ++i) {
          const uint32_t x = q.ef[i], y = q.et[i];
          if (placed[x] == placed[y]) continue;
          feats.resize(feats.size() + GPSM_FEAT);
          step_features(c, placed, 2, i, placed[x] ? y : x, lr, feats.data() + feats.size() - GPSM_FEAT);
        }
        s.first.push_back({e, a});
        s.span.push_back({begin, feats.size() / GPSM_FEAT - row0});
      }
  } else {
    s.child.clear();
    for (uint32_t w = 0; w < q.n; ++w) {
      if (s.placed[w]) continue;
      int be = -1;
      for (uint32_t i = 0; i < E; ++i) {
        const bool hit = (q.ef[i] == w && s.placed[q.et[i]]) || (q.et[i] == w && s.placed[q.ef[i]]);
        if (hit && (be < 0 || cs[i] < cs[be])) be = (int)i;
      }
      if (be < 0) continue;
      s.child.push_back(w);
      feats.resize(feats.size() + GPSM_FEAT);
      step_features(c, s.placed, (uint32_t)s.order.size(), (uint32_t)be, w, s.lr,
                    feats.data() + feats.size() - GPSM_FEAT);
    }
  }
  return feats.size() / GPSM_FEAT - row0;
}

void learned_consume(LearnedState& s, const float* y) {
  const HostQuery& q = *s.c->q;
  const auto& cs = *s.c->cet_size;
  if (s.round == 0) {
    double best = 1e300;
    size_t bi = 0;
    for (size_t i = 0; i < s.first.size(); ++i) {
      double r2 = 0;
      if (s.span[i].second > s.span[i].first) {
        r2 = 1e300;
        for (size_t j = s.span[i].first; j < s.span[i].second; ++j) r2 = std::min(r2, std::expm1((double)y[j]));
      }
      const double cost = (double)cs[s.first[i].first] + r2;
      if (cost < best) {
        best = cost;
        bi = i;
      }
    }
    const uint32_t e0 = s.first[bi].first, a0 = s.first[bi].second;
    s.order = {a0, q.ef[e0] == a0 ? q.et[e0] : q.ef[e0]};
    s.placed[s.order[0]] = s.placed[s.order[1]] = 1;
    s.lr = std::log1p((double)cs[e0]);
  } else {
    if (s.child.empty()) {
      s.done = true;  // disconnected: learned_finish reports failure
      return;
    }
    size_t j = 0;
    for (size_t i = 1; i < s.child.size(); ++i)
      if (y[i] < y[j]) j = i;
    s.order.push_back(s.child[j]);
    s.placed[s.child[j]] = 1;
    s.lr = std::max(0.0f, y[j]);
  }
  ++s.round;
  if (s.order.size() >= q.n) s.done = true;
}

bool learned_finish(const LearnedState& s, OrderPlan& out) {
  if (s.order.size() != s.c->q->n) return false;
  return order_from_vertices(*s.c->q, *s.c->cet_size, s.order.data(), out);
}

namespace {

// Greedy continuation of a two-vertex prefix.  Every unplaced vertex adjacent
// to the prefix is scored by (bound neighbours, smallest |CET| of an edge to
// the prefix); constrained = true takes the most bound neighbours first.
void continue_order(const HostQuery& q, const std::vector<uint64_t>& cs, bool constrained, std::vector<uint32_t>& od) {
  std::vector<uint8_t> placed(q.n, 0);
  for (uint32_t v : od) placed[v] = 1;
  const size_t E = q.ef.size();
  while (od.size() < q.n) {
    int best = -1;
    uint32_t bnb = 0;
    uint64_t bsz = 0;
    for (uint32_t w = 0; w < q.n; ++w) {
      if (placed[w]) continue;
      uint32_t nb = 0;
      uint64_t sz = ~0ull;
      for (size_t i = 0; i < E; ++i) {
        const bool hit = (q.ef[i] == w && placed[q.et[i]]) || (q.et[i] == w && placed[q.ef[i]]);
        if (!hit) continue;
        ++nb;
        sz = std::min(sz, cs[i]);
      }
      if (!nb) continue;
      const bool better = best < 0 || (constrained ? (nb > bnb || (nb == bnb && sz < bsz)) : sz < bsz);
      if (better) {
        best = (int)w;
        bnb = nb;
        bsz = sz;
      }
    }
    if (best < 0) return;  // disconnected: the caller drops the order
    od.push_back((uint32_t)best);
    placed[best] = 1;
  }
}

constexpr uint32_t kFirst = 4;

// Every vertex order whose prefixes are connected, in lexicographic order
// (depth-first, smallest vertex first); false (out incomplete) once there
// are more than cap.
bool all_connected_orders(const HostQuery& q, size_t cap, std::vector<std::vector<uint32_t>>& out) {
  std::vector<std::vector<uint8_t>> adj(q.n, std::vector<uint8_t>(q.n, 0));
  for (size_t i = 0; i < q.ef.size(); ++i) adj[q.ef[i]][q.et[i]] = adj[q.et[i]][q.ef[i]] = 1;
  std::vector<uint32_t> od;
  std::vector<uint8_t> placed(q.n, 0);
  bool ok = true;
  auto rec = [&](auto&& self) -> void {
    if (!ok) return;
    if (od.size() == q.n) {
      if (out.size() == cap) {
        ok = false;
        return;
      }
      out.push_back(od);
      return;
    }
    for (uint32_t w = 0; w < q.n && ok; ++w) {
      if (placed[w]) continue;
      bool touches = od.empty();
      for (uint32_t v : od) touches = touches || adj[w][v];
      if (!touches) continue;
      placed[w] = 1;
      od.push_back(w);
      self(self);
      od.pop_back();
      placed[w] = 0;
    }
  };
  rec(rec);
  return ok;
}

}  // namespace

void candidate_orders(const HostQuery& q, const std::vector<uint64_t>& cet_size, uint64_t big,
                      std::vector<std::vector<uint32_t>>& out) {
  out.clear();
  OrderPlan h;
  heuristic_order(q, cet_size, big, h);
  out.push_back(h.order);
  // a small query: every connected order is a candidate
  std::vector<std::vector<uint32_t>> all;
  if (all_connected_orders(q, kAllOrders, all)) {
    for (auto& od : all)
      if (od != h.order) out.push_back(std::move(od));
    return;
  }
  const size_t E = q.ef.size();
  std::vector<uint32_t> edges(E);
  for (uint32_t i = 0; i < E; ++i) edges[i] = i;
  std::stable_sort(edges.begin(), edges.end(), [&](uint32_t a, uint32_t b) { return cet_size[a] < cet_size[b]; });
  for (size_t r = 0; r < edges.size() && r < kFirst; ++r)
    for (int side = 0; side < 2; ++side)
      for (int constrained = 0; constrained < 2; ++constrained) {
        const uint32_t e = edges[r];
        std::vector<uint32_t> od = {side ? q.et[e] : q.ef[e], side ? q.ef[e] : q.et[e]};
        continue_order(q, cet_size, constrained != 0, od);
        if (od.size() != q.n) continue;
        if (std::find(out.begin(), out.end(), od) == out.end()) out.push_back(std::move(od));
      }
}

namespace {

// Candidates examined per row of level k-1 when position k is bound.
double examine_factor(const PlanCost& s, uint32_t k) {
  const HostQuery& q = *s.c->q;
  const uint32_t e = s.plan.par_edge[k], w = s.plan.order[k];
  const uint32_t parent = q.ef[e] == w ? q.et[e] : q.ef[e];
  const double fan = (double)(*s.c->cet_size)[e] / (double)std::max<uint32_t>((*s.c->cand)[parent], 1);
  return fan * (1.0 + kBackWeight * (double)s.plan.backs[k].size());
}

// The final level needs no prediction: add its examined candidates and stop.
void settle(PlanCost& s) {
  if (!s.done && s.k + 1 >= s.c->q->n) {
    s.cost += s.rows * examine_factor(s, s.k);
    s.done = true;
  }
}

}  // namespace

bool cost_start(const PlanContext& c, const uint32_t* order, PlanCost& s) {
  s.c = &c;
  if (!order_from_vertices(*c.q, *c.cet_size, order, s.plan)) return false;
  s.placed.assign(c.q->n, 0);
  s.placed[order[0]] = s.placed[order[1]] = 1;
  s.k = 2;
  s.rows = (double)(*c.cet_size)[s.plan.first_edge];
  s.lr = std::log1p(s.rows);
  s.cost = s.rows;
  s.done = c.q->n <= 2;
  settle(s);
  return true;
}

size_t cost_emit(PlanCost& s, std::vector<float>& feats) {
  if (s.done) return 0;
  feats.resize(feats.size() + GPSM_FEAT);
  step_features(*s.c, s.placed, s.k, s.plan.par_edge[s.k], s.plan.order[s.k], s.lr,
                feats.data() + feats.size() - GPSM_FEAT);
  return 1;
}

void cost_consume(PlanCost& s, const float* y) {
  if (s.done) return;
  s.cost += s.rows * examine_factor(s, s.k);
  const double ly = std::max(0.0f, y[0]);
  s.rows = std::expm1(ly);
  s.lr = ly;
  s.placed[s.plan.order[s.k]] = 1;
  ++s.k;
  settle(s);
}

double analytic_cost(const PlanContext& c, const OrderPlan& plan) {
  PlanCost s;
  if (!cost_start(c, plan.order.data(), s)) return 0.0;
  std::vector<float> f;
  while (!s.done) {
    f.clear();
    cost_emit(s, f);
    const float y = f[27];
    cost_consume(s, &y);
  }
  return s.cost;
}

int learned_plans(const std::vector<const PlanContext*>& qs, PredictFn predict, void* user, uint64_t big,
                  std::vector<OrderPlan>& out, std::vector<uint32_t>* chosen, std::vector<double>* cost,
                  const std::vector<uint8_t>* rank) {
  const size_t nq = qs.size();
  out.assign(nq, OrderPlan());
  if (chosen) chosen->assign(nq, 0);
  if (cost) cost->assign(nq, 0.0);
  std::vector<float> feats, y;
  auto ranked = [&](size_t b) { return !rank || (*rank)[b]; };
  // 1. greedy learned orders, one decision round at a time for the batch
  std::vector<LearnedState> ls(nq);
  std::vector<size_t> live, rows_of(nq);
  for (size_t b = 0; b < nq; ++b) {
    learned_start(*qs[b], ls[b]);
    if (ranked(b)) live.push_back(b);
  }
  while (!live.empty()) {
    feats.clear();
    for (size_t b : live) rows_of[b] = learned_emit(ls[b], feats);
    const uint32_t R = (uint32_t)(feats.size() / GPSM_FEAT);
    y.assign(R + 1, 0.0f);
    if (R) {
      const int rc = predict(user, feats.data(), R, y.data());
      if (rc) return rc;
    }
    size_t at = 0;
    std::vector<size_t> next;
    for (size_t b : live) {
      learned_consume(ls[b], y.data() + at);
      at += rows_of[b];
      if (!ls[b].done) next.push_back(b);
    }
    live.swap(next);
  }
  // 2. candidates: heuristic first, the greedy order, the greedy continuations
  std::vector<std::vector<std::vector<uint32_t>>> cands(nq);
  std::vector<std::vector<PlanCost>> pc(nq);
  for (size_t b = 0; b < nq; ++b) {
    const PlanContext& c = *qs[b];
    if (ranked(b)) {
      candidate_orders(*c.q, *c.cet_size, big, cands[b]);
    } else {
      OrderPlan h;
      heuristic_order(*c.q, *c.cet_size, big, h);
      cands[b].assign(1, h.order);
    }
    if (ls[b].order.size() == c.q->n &&
        std::find(cands[b].begin(), cands[b].end(), ls[b].order) == cands[b].end())
      cands[b].push_back(ls[b].order);
    pc[b].resize(cands[b].size());
    for (size_t i = 0; i < cands[b].size(); ++i)
      if (!cost_start(c, cands[b][i].data(), pc[b][i])) pc[b][i].done = true, pc[b][i].cost = 1e300;
  }
  // 3. chained costs, one level per round for every candidate of the batch
  for (;;) {
    feats.clear();
    for (size_t b = 0; b < nq; ++b)
      for (auto& s : pc[b]) cost_emit(s, feats);
    const uint32_t R = (uint32_t)(feats.size() / GPSM_FEAT);
    if (!R) break;
    y.assign(R, 0.0f);
    const int rc = predict(user, feats.data(), R, y.data());
    if (rc) return rc;
    size_t at = 0;
    for (size_t b = 0; b < nq; ++b)
      for (auto& s : pc[b])
        if (!s.done) cost_consume(s, y.data() + at++);
  }
  // 4. the cheapest candidate, the heuristic unless another is clearly cheaper
  for (size_t b = 0; b < nq; ++b) {
    size_t best = 0;
    for (size_t i = 1; i < pc[b].size(); ++i)
      if (pc[b][i].cost < pc[b][best].cost) best = i;
    if (best && !(pc[b][best].cost < kLearnedMargin * pc[b][0].cost)) best = 0;
    if (pc[b][best].cost >= 1e300) {
      heuristic_order(*qs[b]->q, *qs[b]->cet_size, big, out[b]);
      if (cost) (*cost)[b] = analytic_cost(*qs[b], out[b]);
      continue;
    }
    out[b] = pc[b][best].plan;
    if (chosen) (*chosen)[b] = (uint32_t)best;
    if (cost) (*cost)[b] = pc[b][best].cost;
  }
  return 0;
}

}  // namespace gpsm
