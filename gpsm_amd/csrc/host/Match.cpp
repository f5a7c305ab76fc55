#include "Match.h"

#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <set>
#include <vector>

namespace {

gpsm_ctx* g_ctx = nullptr;
std::map<const Graph*, gpsm_graph*> g_resident;  // data graphs uploaded once
std::set<const gpsm_graph*> g_directed;          // ... with their out-lists attached

// The query as the reference lists it: every in-edge of every vertex
// (io/IO.cpp:77-86 builds in/out lists in file order).
struct QueryArrays {
  std::vector<uint32_t> labels, src, dst;
  gpsm_query q{};
  explicit QueryArrays(const Graph* g) {
    for (const Vertex& v : g->vertices) labels.push_back((uint32_t)v.label);
    for (size_t i = 0; i < g->vertices.size(); ++i)
      for (const Neighbor& nb : g->vertices[i].out) {
        src.push_back((uint32_t)i);
        dst.push_back((uint32_t)nb.vid);
      }
    q.n = (uint32_t)labels.size();
    q.m = (uint32_t)src.size();
    q.labels = labels.data();
    q.src = src.data();
    q.dst = dst.data();
  }
};

}  // namespace

void Match::initGPU(int dev) {
  if (g_ctx) return;
  int rc = gpsm_init(dev, &g_ctx);
  if (rc != GPSM_OK) {
    std::fprintf(stderr, "initGPU: %s\n", gpsm_strerror(rc));
    g_ctx = nullptr;
    return;
  }
  std::printf("Using device %d: %s\n", dev, gpsm_device_name(g_ctx));
}

bool Match::gpuReady() { return g_ctx != nullptr; }

void Match::releaseGPU() {
  for (auto& kv : g_resident) gpsm_graph_free(kv.second);
  g_resident.clear();
  g_directed.clear();
  gpsm_shutdown(g_ctx);
  g_ctx = nullptr;
}

void Match::forget(const Graph* data) {
  auto it = g_resident.find(data);
  if (it == g_resident.end()) return;
  g_directed.erase(it->second);
  gpsm_graph_free(it->second);
  g_resident.erase(it);
}

Match::Match(Graph* q, Graph* d) : query(q), data(d) {}

Match::~Match() { delete[] id2pos; }

gpsm_graph* Match::resident() {
  auto it = g_resident.find(data);
  if (it != g_resident.end()) return it->second;
  gpsm_graph* g = nullptr;
  status = gpsm_graph_upload(g_ctx, data->vertex_num, data->undirected_row_offset,
                             data->undirected_column_index, data->vertex_value, &g);
  if (status != GPSM_OK) return nullptr;
  g_resident[data] = g;
  return g;
}

int Match::count(uint64_t& embeddings, uint64_t& hash) {
  embeddings = hash = 0;
  if (!g_ctx) return status = GPSM_ENODEV;
  gpsm_graph* g = resident();
  if (!g) return status;
  QueryArrays qa(query);
  gpsm_result r;
  status = gpsm_match(g_ctx, g, &qa.q, nullptr, &r);
  embeddings = r.count;
  hash = r.hash;
  return status;
}

void Match::match(IO&, unsigned*& final_result, unsigned& result_row_num, unsigned& result_col_num, int*& id_map) {
  final_result = nullptr;
  result_row_num = 0;
  result_col_num = query->vertex_num;
  delete[] id2pos;
  id2pos = new int[query->vertex_num];
  for (unsigned i = 0; i < query->vertex_num; ++i) id2pos[i] = (int)i;
  id_map = id2pos;
  if (!g_ctx) {
    status = GPSM_ENODEV;
    return;
  }
  gpsm_graph* g = resident();
  if (!g) return;
  QueryArrays qa(query);
  gpsm_opts o{};
  o.want_tuples = 1;
  o.tuple_cap = 1ull << 28;
  gpsm_result r;
  status = gpsm_match(g_ctx, g, &qa.q, &o, &r);
  if (status != GPSM_OK) {
    std::fprintf(stderr, "match: %s\n", gpsm_strerror(status));
    gpsm_result_free(&r);
    return;
  }
  result_row_num = (unsigned)r.n_tuples;
  if (result_row_num) {
    final_result = new unsigned[(size_t)result_row_num * result_col_num];
    std::memcpy(final_result, r.tuples, sizeof(unsigned) * (size_t)result_row_num * result_col_num);
  }
  gpsm_result_free(&r);
}

int Match::verify_directed(unsigned* final_result, unsigned& result_row_num) {
  if (!g_ctx) return status = GPSM_ENODEV;
  gpsm_graph* g = resident();
  if (!g) return status;
  if (!g_directed.count(g)) {
    const unsigned n = (unsigned)data->vertices.size();
    if (n != data->vertex_num) return status = GPSM_EINVAL;  // no in/out lists (a binary cache)
    std::vector<uint32_t> row(n + 1, 0), col, lab;
    std::vector<std::pair<uint32_t, uint32_t>> arcs;
    for (unsigned v = 0; v < n; ++v) {
      arcs.clear();
      for (const Neighbor& nb : data->vertices[v].out) arcs.emplace_back((uint32_t)nb.vid, (uint32_t)nb.elb);
      std::sort(arcs.begin(), arcs.end());
      for (const auto& a : arcs) {
        col.push_back(a.first);
        lab.push_back(a.second);
      }
      row[v + 1] = (uint32_t)col.size();
    }
    if ((status = gpsm_graph_set_directed(g, row.data(), col.data(), lab.data())) != GPSM_OK) return status;
    g_directed.insert(g);
  }
  QueryArrays qa(query);
  std::vector<uint32_t> elabel;
  for (const Vertex& v : query->vertices)
    for (const Neighbor& nb : v.out) elabel.push_back((uint32_t)nb.elb);
  const unsigned q = query->vertex_num;
  std::vector<uint8_t> keep(result_row_num);
  status = gpsm_verify_directed(g_ctx, g, q, qa.q.m, qa.src.data(), qa.dst.data(), elabel.data(), final_result,
                                result_row_num, keep.data(), nullptr);
  if (status != GPSM_OK) return status;
  unsigned w = 0;
  for (unsigned r = 0; r < result_row_num; ++r)
    if (keep[r]) {
      if (w != r) std::memmove(final_result + (size_t)w * q, final_result + (size_t)r * q, sizeof(unsigned) * q);
      ++w;
    }
  result_row_num = w;
  return status;
}

int Match::write_rows(int fd, const unsigned* rows, unsigned n, unsigned cols) {
  if (!g_ctx) return GPSM_ENODEV;
  const int rc = gpsm_write_tuples(g_ctx, fd, rows, n, cols, nullptr, nullptr);
  if (rc != GPSM_OK) std::fprintf(stderr, "output: %s\n", gpsm_strerror(rc));
  return rc;
}
