// CPU oracle of the GpSM filter-and-join path — TEST INFRASTRUCTURE ONLY
// (checker for tests/, the smoke check and bench.py's CPU baseline).
// Citations are bookug/GpSM paths relative to the reference root.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <queue>
#include <vector>

namespace {

uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// Hash of one tuple: the product mod 2^64 of one odd 64-bit mix per
// (query vertex, image) pair; the set hash is the sum of tuple hashes.
uint64_t term(uint32_t j, uint32_t v) { return fmix64((((uint64_t)j << 32) | v) ^ 0x9E3779B97F4A7C15ULL) | 1ULL; }

// ---------------------------------------------------------------------------
// Query in the reference's representation: per-vertex in/out lists in file
// order (Graph::addEdge, graph/Graph.cpp:18-22) and the undirected CSR with
// sorted rows (Graph::transformToCSR, graph/Graph.cpp:30-64).
struct RefQuery {
  uint32_t n = 0;
  std::vector<uint32_t> lab;
  std::vector<std::vector<uint32_t>> in, out;
  std::vector<uint32_t> ro, ci;
  uint32_t deg(uint32_t u) const { return ro[u + 1] - ro[u]; }
};

bool build_query(const orc_query* q, RefQuery& r) {
  r.n = q->n;
  r.lab.assign(q->labels, q->labels + q->n);
  r.in.assign(q->n, {});
  r.out.assign(q->n, {});
  for (uint32_t e = 0; e < q->m; ++e) {
    if (q->src[e] >= q->n || q->dst[e] >= q->n) return false;
    r.out[q->src[e]].push_back(q->dst[e]);
    r.in[q->dst[e]].push_back(q->src[e]);
  }
  r.ro.assign(q->n + 1, 0);
  for (uint32_t u = 0; u < q->n; ++u) r.ro[u + 1] = r.ro[u] + (uint32_t)(r.in[u].size() + r.out[u].size());
  r.ci.resize(r.ro[q->n]);
  for (uint32_t u = 0; u < q->n; ++u) {
    uint32_t p = r.ro[u];
    for (uint32_t x : r.in[u]) r.ci[p++] = x;
    for (uint32_t x : r.out[u]) r.ci[p++] = x;
    std::sort(r.ci.begin() + r.ro[u], r.ci.begin() + r.ro[u + 1]);
  }
  return true;
}

struct O2 {
  const orc_graph* g;
  RefQuery Q;
  uint32_t dn = 0;
  std::vector<uint32_t> ef, et;  // edge_from / edge_to (from < to)
  std::vector<uint32_t> filter_order;
  std::vector<uint8_t> cand;  // [q][n] bool, match/Match.cu:1764

  uint32_t ddeg(uint32_t v) const { return g->row_offset[v + 1] - g->row_offset[v]; }

  // Match::match query edge list: every in-edge of every vertex, (min, max),
  // match/Match.cu:1735-1754.
  void query_edges() {
    ef.clear();
    et.clear();
    for (uint32_t i = 0; i < Q.n; ++i)
      for (uint32_t src : Q.in[i]) {
        ef.push_back(std::min(src, i));
        et.push_back(std::max(src, i));
      }
  }

  // Match::generateSpanningTree, match/Match.cu:338-429.  score = deg/freq
  // computed as (double)deg / (double)(float)freq, stored as float; the edge
  // with the largest float sum score[from]+score[to] (strict >, from 0) seeds
  // a BFS over the sorted query CSR.
  bool spanning_order() {
    std::vector<uint32_t> freq_lab;  // inverse label list, graph/Graph.cpp:67-101
    std::vector<float> score(Q.n);
    for (uint32_t i = 0; i < Q.n; ++i) {
      uint64_t cnt = 0;
      for (uint32_t v = 0; v < dn; ++v) cnt += g->labels[v] == Q.lab[i];
      float num = (float)cnt;
      int degree = (int)Q.deg(i);
      score[i] = (float)((degree + 0.0) / num);
    }
    float maxv = 0.0f;
    int maxi = -1;
    for (size_t i = 0; i < ef.size(); ++i) {
      float tmp = score[ef[i]] + score[et[i]];
      if (tmp > maxv) {
        maxv = tmp;
        maxi = (int)i;
      }
    }
    if (maxi < 0) return false;
    uint32_t from = ef[maxi], to = et[maxi];
    uint32_t start = from;
    if (score[to] > score[from]) start = to;
    std::vector<uint8_t> visit(Q.n, 0);
    filter_order.clear();
    filter_order.push_back(start);
    visit[start] = 1;
    std::queue<uint32_t> tq;
    tq.push(start);
    while (!tq.empty()) {
      uint32_t t = tq.front();
      tq.pop();
      for (uint32_t i = Q.ro[t]; i < Q.ro[t + 1]; ++i) {
        uint32_t adj = Q.ci[i];
        if (visit[adj]) continue;
        tq.push(adj);
        filter_order.push_back(adj);
        visit[adj] = 1;
      }
    }
    return filter_order.size() == Q.n;
  }

  uint8_t* row(uint32_t u) { return cand.data() + (size_t)u * dn; }

  // kernel_collect, match/Match.cu:488-526: ascending ids of C(u).
  std::vector<uint32_t> collect(uint32_t u) {
    std::vector<uint32_t> a;
    const uint8_t* r = row(u);
    for (uint32_t v = 0; v < dn; ++v)
      if (r[v]) a.push_back(v);
    return a;
  }

  // check_kernel, match/Match.cu:431-446: label equal and degree >= query degree.
  void check(uint32_t u) {
    uint8_t* r = row(u);
    for (uint32_t v = 0; v < dn; ++v)
      if (g->labels[v] == Q.lab[u] && ddeg(v) >= Q.deg(u)) r[v] = 1;
  }

  // explore_kernel + kernel_explore, match/Match.cu:528-659.  For each
  // candidate mu and each uninitialised neighbour v in CSR order: every data
  // neighbour with v's label and enough degree enters C(v); if none does, mu
  // leaves C(u) and the remaining neighbours are skipped.
  void explore(uint32_t u, const std::vector<uint32_t>& arr, const std::vector<uint8_t>& initialized) {
    std::vector<uint32_t> adjs;
    for (uint32_t i = Q.ro[u]; i < Q.ro[u + 1]; ++i)
      if (!initialized[Q.ci[i]]) adjs.push_back(Q.ci[i]);
    if (adjs.empty()) return;
    for (uint32_t mu : arr) {
      for (uint32_t v : adjs) {
        bool found = false;
        uint32_t lv = Q.lab[v], dv = Q.deg(v);
        uint8_t* rv = row(v);
        for (uint32_t e = g->row_offset[mu]; e < g->row_offset[mu + 1]; ++e) {
          uint32_t mv = g->col[e];
          if (g->labels[mv] == lv && ddeg(mv) >= dv) {
            rv[mv] = 1;
            found = true;
          }
        }
        if (!found) {
          row(u)[mu] = 0;
          break;
        }
      }
    }
  }

  // InitializeCandidateVertices, match/Match.cu:661-719.
  int initialize() {
    std::vector<uint8_t> initialized(Q.n, 0);
    cand.assign((size_t)Q.n * dn, 0);
    for (uint32_t i = 0; i < Q.n; ++i) {
      uint32_t u = filter_order[i];
      if (!initialized[u]) {
        check(u);
        initialized[u] = 1;
      }
      std::vector<uint32_t> arr = collect(u);
      if (arr.empty()) return ORC_EMPTY_INIT;
      explore(u, arr, initialized);
      for (uint32_t j = Q.ro[u]; j < Q.ro[u + 1]; ++j) initialized[Q.ci[j]] = 1;
    }
    return ORC_OK;
  }

  // generateSimplifiedGraph + refine_kernel + RefineCandidateVertices,
  // match/Match.cu:721-890: vertices of query degree > 1; 5 rounds, the last
  // only checking non-emptiness; Gauss-Seidel in id order.
  int refine() {
    std::vector<uint32_t> simple;
    std::vector<uint8_t> valid(Q.n, 0);
    for (uint32_t i = 0; i < Q.n; ++i)
      if (Q.deg(i) > 1) {
        valid[i] = 1;
        simple.push_back(i);
      }
    std::vector<std::vector<uint32_t>> sedges(simple.size());
    for (size_t i = 0; i < simple.size(); ++i) {
      for (uint32_t id : Q.in[simple[i]])
        if (valid[id]) sedges[i].push_back(id);
      for (uint32_t id : Q.out[simple[i]])
        if (valid[id]) sedges[i].push_back(id);
    }
    const int round = 5;
    for (int k = 0; k < round; ++k) {
      for (size_t i = 0; i < simple.size(); ++i) {
        uint32_t u = simple[i];
        std::vector<uint32_t> arr = collect(u);
        if (arr.empty()) return ORC_EMPTY_REFINE;
        if (k == round - 1) continue;
        if (sedges[i].empty()) continue;
        for (uint32_t mu : arr) {
          for (uint32_t v : sedges[i]) {
            const uint8_t* rv = row(v);
            bool found = false;
            for (uint32_t e = g->row_offset[mu]; e < g->row_offset[mu + 1] && !found; ++e)
              found = rv[g->col[e]] != 0;
            if (!found) {
              row(u)[mu] = 0;
              break;
            }
          }
        }
      }
    }
    return ORC_OK;
  }

  // Candidate-edge table of (u -> v): count_kernel / examine_kernel,
  // match/Match.cu:892-1080.  ids = C(u) ascending, off exclusive scan of the
  // per-candidate counts, lists = adj(mu) ∩ C(v) in CSR (ascending) order.
  struct CET {
    std::vector<uint32_t> ids, off, list;
    std::vector<uint32_t> rank;  // data vertex -> index in ids (or ~0)
  };

  CET make_cet(uint32_t u, uint32_t v) {
    CET t;
    t.ids = collect(u);
    t.off.assign(t.ids.size() + 1, 0);
    const uint8_t* rv = row(v);
    for (size_t j = 0; j < t.ids.size(); ++j) {
      uint32_t mu = t.ids[j];
      for (uint32_t e = g->row_offset[mu]; e < g->row_offset[mu + 1]; ++e)
        if (rv[g->col[e]]) t.list.push_back(g->col[e]);
      t.off[j + 1] = (uint32_t)t.list.size();
    }
    return t;
  }

  void index(CET& t) {
    t.rank.assign(dn, ~0u);
    for (size_t j = 0; j < t.ids.size(); ++j) t.rank[t.ids[j]] = (uint32_t)j;
  }
};

// ---------------------------------------------------------------------------
// Join of the candidate-edge tables, JoinCandidateEdges + table/join/expand/
// link kernels, match/Match.cu:1141-1708, as a depth-first walk that visits the
// rows of every intermediate table in the reference's row order.
struct Step {
  uint32_t edge;
  bool join;
  uint32_t apos, bpos;  // join: positions of edge_from/edge_to; expand: parent position
  int cet;              // index into tables
};

struct Joiner {
  std::vector<O2::CET>* tabs;
  std::vector<Step> steps;
  uint32_t q;
  std::vector<uint32_t> pos2id;
  int mode;
  std::vector<uint64_t> step_rows;
  uint64_t examined = 0;           // list entries visited by expand steps
  std::vector<uint64_t> step_exam;  // ... per step
  uint64_t count = 0, hash = 0;
  std::vector<uint32_t>* tuples = nullptr;
  uint64_t cap = 0, n_tuples = 0;
  std::vector<uint32_t> rowv;

  void emit() {
    std::vector<uint32_t> t(q);
    for (uint32_t p = 0; p < q; ++p) t[pos2id[p]] = rowv[p];
    if (mode == 1) {  // IO::verify injectivity, io/IO.cpp:149-160
      std::vector<uint32_t> s(t);
      std::sort(s.begin(), s.end());
      if (std::adjacent_find(s.begin(), s.end()) != s.end()) return;
    }
    uint64_t h = 1;
    for (uint32_t j = 0; j < q; ++j) h *= term(j, t[j]);
    hash += h;
    ++count;
    if (tuples && n_tuples < cap) tuples->insert(tuples->end(), t.begin(), t.end());
    ++n_tuples;
  }

  void run(size_t s) {
    if (s == steps.size()) {
      emit();
      return;
    }
    const Step& st = steps[s];
    O2::CET& t = (*tabs)[st.cet];
    if (st.join) {
      // join_kernel: row[vpos] must be in the list of row[upos]
      uint32_t j = t.rank[rowv[st.apos]];
      if (j == ~0u) return;
      if (!std::binary_search(t.list.begin() + t.off[j], t.list.begin() + t.off[j + 1], rowv[st.bpos])) return;
      ++step_rows[s];
      run(s + 1);
    } else {
      // expand_kernel + link_kernel: append every list entry of row[pos]
      uint32_t j = t.rank[rowv[st.apos]];
      if (j == ~0u) return;
      examined += t.off[j + 1] - t.off[j];
      step_exam[s] += t.off[j + 1] - t.off[j];
      for (uint32_t e = t.off[j]; e < t.off[j + 1]; ++e) {
        uint32_t mv = t.list[e];
        if (mode == 0 && std::find(rowv.begin(), rowv.end(), mv) != rowv.end()) continue;
        rowv.push_back(mv);
        ++step_rows[s];
        run(s + 1);
        rowv.pop_back();
      }
    }
  }
};

}  // namespace

extern "C" {

uint64_t orc_tuple_hash(const uint32_t* t, uint32_t q) {
  uint64_t h = 1;
  for (uint32_t j = 0; j < q; ++j) h *= term(j, t[j]);
  return h;
}

int orc_o2_filter(const orc_graph* g, const orc_query* q, uint8_t* init, uint8_t* refined,
                  uint32_t* filter_order) {
  O2 o;
  o.g = g;
  o.dn = g->n;
  if (!build_query(q, o.Q)) return ORC_ERR;
  o.query_edges();
  if (!o.spanning_order()) return ORC_ERR;
  if (filter_order) std::copy(o.filter_order.begin(), o.filter_order.end(), filter_order);
  int st = o.initialize();
  if (init) std::memcpy(init, o.cand.data(), o.cand.size());
  if (st != ORC_OK) {
    if (refined) std::memcpy(refined, o.cand.data(), o.cand.size());
    return st;
  }
  st = o.refine();
  if (refined) std::memcpy(refined, o.cand.data(), o.cand.size());
  return st;
}

int64_t orc_o2_cet(const orc_graph* g, const orc_query* q, uint32_t* buf, uint64_t cap,
                   uint64_t* table_off, uint32_t* edge_from, uint32_t* edge_to, uint32_t* cand_num,
                   uint32_t* n_edges) {
  O2 o;
  o.g = g;
  o.dn = g->n;
  if (!build_query(q, o.Q)) return ORC_ERR;
  o.query_edges();
  if (!o.spanning_order()) return ORC_ERR;
  if (o.initialize() != ORC_OK || o.refine() != ORC_OK) return -2;
  uint64_t total = 0;
  *n_edges = (uint32_t)o.ef.size();
  for (size_t e = 0; e < o.ef.size(); ++e) {
    O2::CET t = o.make_cet(o.ef[e], o.et[e]);
    edge_from[e] = o.ef[e];
    edge_to[e] = o.et[e];
    cand_num[e] = (uint32_t)t.ids.size();
    table_off[e] = total;
    uint64_t words = t.ids.size() + t.off.size() + t.list.size();
    if (buf && total + words <= cap) {
      uint32_t* p = buf + total;
      std::copy(t.ids.begin(), t.ids.end(), p);
      std::copy(t.off.begin(), t.off.end(), p + t.ids.size());
      std::copy(t.list.begin(), t.list.end(), p + t.ids.size() + t.off.size());
    }
    total += words;
  }
  return (int64_t)total;
}

int orc_o2_match(const orc_graph* g, const orc_query* q, int mode, const uint32_t* order,
                 orc_o2_out* out) {
  uint32_t* tbuf = out->tuples;
  uint64_t tcap = out->tuple_cap;
  std::memset(out, 0, sizeof(*out));
  out->tuples = tbuf;
  out->tuple_cap = tcap;
  if (q->n > ORC_MAXQ || q->m > ORC_MAXE) return ORC_ERR;
  O2 o;
  o.g = g;
  o.dn = g->n;
  if (!build_query(q, o.Q)) return ORC_ERR;
  o.query_edges();
  if (o.ef.empty() || !o.spanning_order()) return ORC_ERR;
  std::copy(o.filter_order.begin(), o.filter_order.end(), out->filter_order);
  int st = o.initialize();
  if (st == ORC_OK) st = o.refine();
  out->status = st;
  const uint32_t E = (uint32_t)o.ef.size(), Qn = o.Q.n;
  out->n_edges = E;
  for (uint32_t e = 0; e < E; ++e) {
    out->edge_from[e] = o.ef[e];
    out->edge_to[e] = o.et[e];
  }
  for (uint32_t u = 0; u < Qn; ++u) out->cand[u] = (uint32_t)o.collect(u).size();
  if (st != ORC_OK) return ORC_OK;

  // FindCandidateEdges forward and reverse, match/Match.cu:1841-1842.
  std::vector<O2::CET> tabs;
  tabs.reserve(2 * E);
  for (uint32_t e = 0; e < E; ++e) tabs.push_back(o.make_cet(o.ef[e], o.et[e]));
  for (uint32_t e = 0; e < E; ++e) tabs.push_back(o.make_cet(o.et[e], o.ef[e]));
  for (auto& t : tabs) o.index(t);
  std::vector<uint32_t> size(E);
  for (uint32_t e = 0; e < E; ++e) size[e] = out->cet_size[e] = (uint32_t)tabs[e].list.size();

  // Join order, JoinCandidateEdges, match/Match.cu:1599-1700.
  std::vector<uint8_t> vis_v(Qn, 0), vis_e(E, 0);
  std::vector<int> id2pos(Qn, -1);
  std::vector<uint32_t> pos2id;
  std::vector<Step> steps;
  auto add_mapping = [&](uint32_t id) {  // match/Match.cu:84-90
    id2pos[id] = (int)pos2id.size();
    pos2id.push_back(id);
  };
  const uint32_t big = (uint32_t)(g->row_offset[g->n]) + 1;  // data->eSize()+1
  uint32_t first = 0;
  if (!order) {
    uint32_t minv = big;
    for (uint32_t i = 0; i < E; ++i)
      if (size[i] < minv) {
        minv = size[i];
        first = i;
      }
    add_mapping(o.ef[first]);
    add_mapping(o.et[first]);
    steps.push_back({first, false, 0, 0, (int)first});  // table_kernel on the forward table
  } else {
    uint32_t a = order[0], b = order[1];
    bool found = false;
    for (uint32_t i = 0; i < E && !found; ++i)
      if ((o.ef[i] == a && o.et[i] == b) || (o.ef[i] == b && o.et[i] == a)) {
        first = i;
        found = true;
      }
    if (!found) return ORC_ERR;
    add_mapping(a);
    add_mapping(b);
    steps.push_back({first, false, 0, 0, (int)(o.ef[first] == a ? first : E + first)});
  }
  vis_v[pos2id[0]] = vis_v[pos2id[1]] = 1;
  vis_e[first] = 1;
  uint32_t next_explicit = 2;
  for (uint32_t cnt = 1; cnt < E; ++cnt) {
    std::vector<uint32_t> next;
    for (uint32_t i = 0; i < E; ++i)
      if (!vis_e[i] && vis_v[o.ef[i]] && vis_v[o.et[i]]) next.push_back(i);
    bool expand = next.empty();
    if (expand) {
      // Expand mode over edges with exactly one bound endpoint: the
      // reference's loop (match/Match.cu:1639-1650) also admits edges with
      // none, whose position lookup reads id2pos = -1; that defect is fixed.
      if (!order) {
        for (uint32_t i = 0; i < E; ++i)
          if (!vis_e[i] && (vis_v[o.ef[i]] != vis_v[o.et[i]])) next.push_back(i);
      } else {
        if (next_explicit >= Qn) return ORC_ERR;
        uint32_t w = order[next_explicit];
        for (uint32_t i = 0; i < E; ++i)
          if (!vis_e[i] && ((o.ef[i] == w && vis_v[o.et[i]]) || (o.et[i] == w && vis_v[o.ef[i]])))
            next.push_back(i);
        if (next.empty()) return ORC_ERR;
      }
    }
    uint32_t nminv = big, nminx = 0;
    for (uint32_t id : next)
      if (size[id] < nminv) {
        nminv = size[id];
        nminx = id;
      }
    uint32_t u = o.ef[nminx], v = o.et[nminx];
    if (expand) {
      if (vis_v[u]) {
        steps.push_back({nminx, false, (uint32_t)id2pos[u], 0, (int)nminx});
        add_mapping(v);
      } else {
        steps.push_back({nminx, false, (uint32_t)id2pos[v], 0, (int)(E + nminx)});
        add_mapping(u);
      }
      ++next_explicit;
    } else {
      steps.push_back({nminx, true, (uint32_t)id2pos[u], (uint32_t)id2pos[v], (int)nminx});
    }
    vis_v[u] = vis_v[v] = 1;
    vis_e[nminx] = 1;
  }
  if (pos2id.size() != Qn) return ORC_ERR;
  for (uint32_t p = 0; p < Qn; ++p) out->order[p] = pos2id[p];

  Joiner J;
  J.tabs = &tabs;
  J.q = Qn;
  J.pos2id = pos2id;
  J.mode = mode;
  J.step_rows.assign(steps.size(), 0);
  std::vector<uint32_t> tl;
  J.tuples = out->tuples ? &tl : nullptr;
  J.cap = out->tuple_cap;
  // step 0: table_kernel rows (mu, mv) of the first table
  {
    O2::CET& t0 = tabs[steps[0].cet];
    J.steps.assign(steps.begin() + 1, steps.end());
    std::vector<uint64_t> rest(J.steps.size(), 0);
    J.step_rows = rest;
    J.step_exam = rest;
    uint64_t r0 = 0;
    for (size_t j = 0; j < t0.ids.size(); ++j)
      for (uint32_t e = t0.off[j]; e < t0.off[j + 1]; ++e) {
        J.rowv = {t0.ids[j], t0.list[e]};
        ++r0;
        J.run(0);
      }
    out->n_steps = (uint32_t)steps.size();
    out->step_rows[0] = r0;
    out->step_edge[0] = steps[0].edge;
    out->step_join[0] = 0;
    for (size_t s = 1; s < steps.size(); ++s) {
      out->step_rows[s] = J.step_rows[s - 1];
      out->step_edge[s] = steps[s].edge;
      out->step_join[s] = steps[s].join;
    }
  }
  // level rows: rows after the last step at each width
  uint32_t width = 2;
  out->level_rows[1] = out->step_rows[0];
  for (size_t s = 1; s < steps.size(); ++s) {
    if (!steps[s].join) ++width;
    out->level_rows[width - 1] = out->step_rows[s];
  }
  for (size_t s = 0; s < steps.size(); ++s)
    if (out->step_rows[s] == 0) {
      out->status = ORC_EMPTY_JOIN;
      break;
    }
  out->count = J.count;
  out->hash = J.hash;
  out->join_work = out->step_rows[0] + J.examined;
  out->level_work[1] = out->step_rows[0];
  for (size_t s = 1, w = 2; s < steps.size(); ++s)
    if (!steps[s].join) out->level_work[w++] = J.step_exam[s - 1];
  out->n_tuples = J.n_tuples;
  if (out->tuples) {
    uint64_t rows = std::min<uint64_t>(J.n_tuples, out->tuple_cap);
    std::vector<uint64_t> idx(rows);
    for (uint64_t i = 0; i < rows; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
      return std::lexicographical_compare(tl.begin() + a * Qn, tl.begin() + (a + 1) * Qn, tl.begin() + b * Qn,
                                          tl.begin() + (b + 1) * Qn);
    });
    for (uint64_t i = 0; i < rows; ++i) std::copy(tl.begin() + idx[i] * Qn, tl.begin() + (idx[i] + 1) * Qn, out->tuples + i * Qn);
  }
  return ORC_OK;
}

// ---------------------------------------------------------------------------
// O1: plain backtracking.  Query vertices in BFS order from vertex 0; each
// vertex takes the data neighbours of its BFS parent's image that carry its
// label, are unused, and are adjacent to the images of all its other earlier
// neighbours.
int orc_o2_match_many(const orc_graph* g, const orc_query* qs, uint32_t nq, int threads, uint64_t* counts,
                      uint64_t* hashes) {
  if (threads <= 0) threads = omp_get_max_threads();
  int rc = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)nq; ++i) {
    orc_o2_out o;
    o.tuples = nullptr;
    o.tuple_cap = 0;
    const int r = orc_o2_match(g, &qs[i], 0, nullptr, &o);
    counts[i] = o.count;
    hashes[i] = o.hash;
    if (r < 0) {
#pragma omp critical
      if (!rc) rc = r;
    }
  }
  return rc;
}

int orc_o1_enumerate(const orc_graph* g, const orc_query* q, int threads, uint64_t* count,
                     uint64_t* hash, uint32_t* tuples, uint64_t cap, uint64_t* n_tuples) {
  const uint32_t k = q->n;
  if (k == 0 || k > ORC_MAXQ) return ORC_ERR;
  std::vector<std::vector<uint32_t>> adj(k);
  for (uint32_t e = 0; e < q->m; ++e) {
    uint32_t a = q->src[e], b = q->dst[e];
    if (a >= k || b >= k || a == b) return ORC_ERR;
    adj[a].push_back(b);
    adj[b].push_back(a);
  }
  for (auto& l : adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  std::vector<uint32_t> ord{0}, posof(k, ~0u), parent(k, ~0u);
  posof[0] = 0;
  for (size_t h = 0; h < ord.size(); ++h)
    for (uint32_t w : adj[ord[h]])
      if (posof[w] == ~0u) {
        posof[w] = (uint32_t)ord.size();
        parent[w] = ord[h];
        ord.push_back(w);
      }
  if (ord.size() != k) return ORC_ERR;
  std::vector<std::vector<uint32_t>> backs(k);  // earlier neighbours by position, parent excluded
  for (uint32_t i = 1; i < k; ++i)
    for (uint32_t w : adj[ord[i]])
      if (posof[w] < i && w != parent[ord[i]]) backs[i].push_back(posof[w]);
  if (threads <= 0) threads = omp_get_max_threads();
  uint64_t tot = 0, hs = 0;
  std::vector<std::vector<uint32_t>> tl(threads);
  const uint32_t lab0 = q->labels[ord[0]];
#pragma omp parallel num_threads(threads) reduction(+ : tot, hs)
  {
    std::vector<uint32_t> img(k);
    std::vector<uint32_t> it(k), end(k);
    auto& mine = tl[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 64)
    for (int64_t r = 0; r < (int64_t)g->n; ++r) {
      if (g->labels[r] != lab0) continue;
      img[0] = (uint32_t)r;
      if (k == 1) {
        hs += term(ord[0], img[0]);
        ++tot;
        continue;
      }
      int d = 1;
      uint32_t pv = img[posof[parent[ord[1]]]];
      it[1] = g->row_offset[pv];
      end[1] = g->row_offset[pv + 1];
      while (d >= 1) {
        if (it[d] == end[d]) {
          --d;
          continue;
        }
        uint32_t c = g->col[it[d]++];
        uint32_t qv = ord[d];
        if (g->labels[c] != q->labels[qv]) continue;
        bool ok = true;
        for (int j = 0; j < d && ok; ++j) ok = img[j] != c;
        for (uint32_t bp : backs[d]) {
          if (!ok) break;
          const uint32_t* lo = g->col + g->row_offset[c];
          const uint32_t* hi = g->col + g->row_offset[c + 1];
          ok = std::binary_search(lo, hi, img[bp]);
        }
        if (!ok) continue;
        img[d] = c;
        if (d == (int)k - 1) {
          uint64_t h = 1;
          for (uint32_t p = 0; p < k; ++p) h *= term(ord[p], img[p]);
          hs += h;
          ++tot;
          if (tuples) {
            size_t base = mine.size();
            mine.resize(base + k);
            for (uint32_t p = 0; p < k; ++p) mine[base + ord[p]] = img[p];
          }
          continue;
        }
        ++d;
        uint32_t pv2 = img[posof[parent[ord[d]]]];
        it[d] = g->row_offset[pv2];
        end[d] = g->row_offset[pv2 + 1];
      }
    }
  }
  *count = tot;
  *hash = hs;
  if (n_tuples) *n_tuples = tot;
  if (tuples) {
    std::vector<uint32_t> all;
    for (auto& v : tl) all.insert(all.end(), v.begin(), v.end());
    uint64_t rows = all.size() / k;
    std::vector<uint64_t> idx(rows);
    for (uint64_t i = 0; i < rows; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
      return std::lexicographical_compare(all.begin() + a * k, all.begin() + (a + 1) * k, all.begin() + b * k,
                                          all.begin() + (b + 1) * k);
    });
    uint64_t w = std::min(rows, cap);
    for (uint64_t i = 0; i < w; ++i) std::copy(all.begin() + idx[i] * k, all.begin() + (idx[i] + 1) * k, tuples + i * k);
  }
  return ORC_OK;
}

uint64_t orc_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t x = in[i];
    out[i] = (uint32_t)acc;
    acc += x;
  }
  return acc;
}

uint64_t orc_compact_bitmap(const uint32_t* bits, uint64_t nbits, uint32_t* ids) {
  uint64_t c = 0;
  for (uint64_t i = 0; i < nbits; ++i)
    if ((bits[i >> 5] >> (i & 31)) & 1u) {
      if (ids) ids[c] = (uint32_t)i;
      ++c;
    }
  return c;
}

// ---------------------------------------------------------------------------
// MLP reference.
static inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

namespace {
constexpr int F = 32, H = 128;
struct Params {
  const float *W1, *b1, *W2, *b2, *W3, *b3;
  explicit Params(const float* p) {
    W1 = p;
    b1 = W1 + H * F;
    W2 = b1 + H;
    b2 = W2 + H * H;
    W3 = b2 + H;
    b3 = W3 + H;
  }
};
}  // namespace

int orc_mlp_forward(const float* params, const float* x, uint32_t B, int bf16, float* y) {
  Params P(params);
  auto R = [&](float v) { return bf16 ? bf16_round(v) : v; };
  std::vector<double> h1(H), h2(H);
  for (uint32_t b = 0; b < B; ++b) {
    const float* xb = x + (size_t)b * F;
    for (int i = 0; i < H; ++i) {
      double s = 0;
      for (int j = 0; j < F; ++j) s += (double)R(P.W1[i * F + j]) * (double)R(xb[j]);
      s += P.b1[i];
      h1[i] = R((float)std::max(s, 0.0));
    }
    for (int i = 0; i < H; ++i) {
      double s = 0;
      for (int j = 0; j < H; ++j) s += (double)R(P.W2[i * H + j]) * h1[j];
      s += P.b2[i];
      h2[i] = std::max(s, 0.0);
    }
    double s = 0;
    for (int j = 0; j < H; ++j) s += (double)R(P.W3[j]) * h2[j];
    y[b] = (float)(s + P.b3[0]);
  }
  return 0;
}

// Gradient of 0.5*mean((y - t)^2).  With bf16 != 0 the backward operands are
// rounded where the device rounds them: W2^T dz2 uses BF16 W2 and BF16 dz2,
// the weight gradients use BF16 activations and BF16 deltas.
int orc_mlp_grad(const float* params, const float* x, const float* t, uint32_t B, int bf16,
                 double* loss, double* grad) {
  Params P(params);
  auto R = [&](double v) { return bf16 ? (double)bf16_round((float)v) : v; };
  const size_t np = (size_t)H * F + H + (size_t)H * H + H + H + 1;
  std::fill(grad, grad + np, 0.0);
  double* gW1 = grad;
  double* gb1 = gW1 + H * F;
  double* gW2 = gb1 + H;
  double* gb2 = gW2 + H * H;
  double* gW3 = gb2 + H;
  double* gb3 = gW3 + H;
  double L = 0;
  std::vector<double> z1(H), h1(H), z2(H), h2(H), d2(H), d1(H);
  for (uint32_t b = 0; b < B; ++b) {
    const float* xb = x + (size_t)b * F;
    for (int i = 0; i < H; ++i) {
      double s = 0;
      for (int j = 0; j < F; ++j) s += R(P.W1[i * F + j]) * R(xb[j]);
      z1[i] = s + P.b1[i];
      h1[i] = R(std::max(z1[i], 0.0));
    }
    for (int i = 0; i < H; ++i) {
      double s = 0;
      for (int j = 0; j < H; ++j) s += R(P.W2[i * H + j]) * h1[j];
      z2[i] = s + P.b2[i];
      h2[i] = std::max(z2[i], 0.0);
    }
    double yv = 0;
    for (int j = 0; j < H; ++j) yv += R(P.W3[j]) * h2[j];
    yv += P.b3[0];
    double dy = (yv - t[b]) / B;
    L += 0.5 * (yv - t[b]) * (yv - t[b]) / B;
    gb3[0] += dy;
    for (int j = 0; j < H; ++j) {
      gW3[j] += dy * h2[j];
      d2[j] = z2[j] > 0 ? dy * R(P.W3[j]) : 0.0;
    }
    for (int i = 0; i < H; ++i) {
      gb2[i] += d2[i];
      double di = R(d2[i]);
      for (int j = 0; j < H; ++j) gW2[i * H + j] += di * h1[j];
    }
    for (int j = 0; j < H; ++j) {
      double s = 0;
      for (int i = 0; i < H; ++i) s += R(P.W2[i * H + j]) * R(d2[i]);
      d1[j] = z1[j] > 0 ? s : 0.0;
    }
    for (int i = 0; i < H; ++i) {
      gb1[i] += d1[i];
      double di = R(d1[i]);
      for (int j = 0; j < F; ++j) gW1[i * F + j] += di * R(xb[j]);
    }
  }
  *loss = L;
  return 0;
}

}  // extern "C"
