// C ABI of the matcher (include/gpsm.h): device context, resident graph and
// the batched filter -> collect -> candidate-edge -> join pipeline.
//
// Phases of one batch (host synchronisations marked *):
//   A  zero bitmaps, label/degree check of every start vertex, one superstep
//      per filter step index (explore / refine ops of all queries), row
//      popcounts                                                   * |C(u)|
//   B  compaction of every C(u) + rank index, forward |CET| counts * |CET|
//   -- host: join order per query (heuristic, learned or explicit)
//   C  candidate-edge tables of the directed edges the plans use
//   D  persistent join over tasks (slices of first tables)        * results
// Phases A and C may be captured in hipGraphs (opts->use_graph); the join
// (D) is launched directly, its two launches side by side on two streams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "api_internal.h"
#include "gpsm.h"
#include "kernels/launch.h"
#include "plan.h"

using namespace gpsm;

#define HIPCHK(x)                                       \
  do {                                                  \
    hipError_t e_ = (x);                                \
    if (e_ != hipSuccess) {                             \
      set_error(#x, e_);                                \
      return e_ == hipErrorOutOfMemory ? GPSM_ENOMEM : GPSM_EDEVICE; \
    }                                                   \
  } while (0)

// Scratch of the allocation-free scan (gpsm_scan_u32_noalloc): the scan's
// total (2 words) and its block sums.  In host code the array's name is not
// a device address: gpsm_init takes the address of this device's copy with
// hipGetSymbolAddress and every use goes through that pointer.
static __device__ __align__(256) uint32_t scratch[GPSM_SCAN_NOALLOC_SCRATCH_BYTES / sizeof(uint32_t)];

namespace {

thread_local std::string g_last_error;

void set_error(const char* what, hipError_t e) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
}

}  // namespace

namespace gpsm {

int DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes) return GPSM_OK;
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  size_t want = std::max<size_t>(nbytes + nbytes / 4, 4096);
  if (hipMalloc(&ptr, want) != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return GPSM_ENOMEM;
  }
  bytes = want;
  return GPSM_OK;
}

void DevBuf::release() {
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int HostBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes) return GPSM_OK;
  if (ptr) hipHostFree(ptr);
  ptr = nullptr;
  bytes = 0;
  size_t want = std::max<size_t>(nbytes + nbytes / 4, 4096);
  if (hipHostMalloc(&ptr, want, hipHostMallocDefault) != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return GPSM_ENOMEM;
  }
  bytes = want;
  return GPSM_OK;
}

void HostBuf::release() {
  if (ptr) hipHostFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

void* PinnedArena::take(size_t nbytes) {
  nbytes = (nbytes + 255) & ~(size_t)255;
  while (cur < blocks.size() && used + nbytes > blocks[cur].bytes) {
    ++cur;
    used = 0;
  }
  if (cur == blocks.size()) {
    blocks.emplace_back();
    if (blocks.back().ensure(std::max<size_t>(nbytes, 1u << 20))) {
      blocks.pop_back();
      return nullptr;
    }
    used = 0;
  }
  void* p = static_cast<char*>(blocks[cur].ptr) + used;
  used += nbytes;
  return p;
}

const char* const kProfNames[K_KINDS_] = {"filter_step", "join",         "cet_count",   "cet_write",
                                           "compact_rows", "init",         "other",       "wall:filter",
                                           "wall:cet_count", "wall:plan", "wall:join"};

hipEvent_t Prof::get() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

}  // namespace gpsm

namespace {

// Hub adjacency bitmaps of a freshly uploaded graph (see DevGraph::hub_of).
// Without room for them the graph has no hubs and every edge test searches.
int build_hubs(gpsm_ctx* c, const uint32_t* row_offset, gpsm_graph* g) {
  DevGraph& d = g->dg;
  const uint32_t n = d.n;
  auto deg = [&](uint32_t v) { return row_offset[v + 1] - row_offset[v]; };
  auto count_over = [&](uint64_t D) {
    uint64_t k = 0;
    for (uint32_t v = 0; v < n; ++v) k += deg(v) > D;
    return k;
  };
  uint64_t D = std::max<uint32_t>(c->hub_min_degree, 1);
  uint64_t nh = c->hub_max_bytes ? count_over(D) : 0;
  while (nh && nh * d.W * 4ull > c->hub_max_bytes) nh = count_over(D *= 2);
  std::vector<uint32_t> hub_of(n, ~0u), hubs, block_base(1, 0);
  if (nh) {
    for (uint32_t v = 0; v < n; ++v)
      if (deg(v) > D) {
        hub_of[v] = (uint32_t)hubs.size();
        hubs.push_back(v);
        block_base.push_back(block_base.back() + (deg(v) + 63) / 64);  // 64-entry blocks, one wave each
      }
  }
  if (hipMalloc(&d.hub_of, sizeof(uint32_t) * n) != hipSuccess) {
    (void)hipGetLastError();
    return GPSM_ENOMEM;
  }
  uint32_t* tmp = nullptr;
  if (!hubs.empty()) {
    if (hipMalloc(&d.hub_bits, sizeof(uint32_t) * (size_t)hubs.size() * d.W) != hipSuccess ||
        hipMalloc(&tmp, sizeof(uint32_t) * (hubs.size() + block_base.size())) != hipSuccess) {
      // no room: no hubs
      (void)hipGetLastError();
      hipFree(d.hub_bits);
      d.hub_bits = nullptr;
      hubs.clear();
      std::fill(hub_of.begin(), hub_of.end(), ~0u);
    }
  }
  d.nhub = (uint32_t)hubs.size();
  d.hub_degree = d.nhub ? (uint32_t)D : 0;
  hipStream_t s = c->stream;
  if (hipMemcpyAsync(d.hub_of, hub_of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s) != hipSuccess)
    return GPSM_EDEVICE;
  if (d.nhub) {
    launch_fill_u32(d.hub_bits, 0u, (uint64_t)d.nhub * d.W, s);
    if (hipMemcpyAsync(tmp, hubs.data(), sizeof(uint32_t) * hubs.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(tmp + hubs.size(), block_base.data(), sizeof(uint32_t) * block_base.size(),
                       hipMemcpyHostToDevice, s) != hipSuccess)
      return GPSM_EDEVICE;
    launch_hub_bits(d, tmp, tmp + hubs.size(), block_base.back(), s);
  }
  const bool ok = hipStreamSynchronize(s) == hipSuccess && hipGetLastError() == hipSuccess;
  hipFree(tmp);
  if (!ok) return GPSM_EDEVICE;
  g->bytes += sizeof(uint32_t) * ((uint64_t)n + (uint64_t)d.nhub * d.W);
  return GPSM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" {

int gpsm_profile_enable(gpsm_ctx* c, int on) {
  if (!c) return GPSM_EINVAL;
  if (on < 0 || on > 3) return GPSM_EINVAL;
  c->prof.mode = on;
  for (auto& a : c->prof.acc) a = Prof::Acc();
  for (auto& p : c->prof.pending) {
    c->prof.pool.push_back(p.a);
    c->prof.pool.push_back(p.b);
  }
  c->prof.pending.clear();
  return GPSM_OK;
}

int gpsm_profile_read(gpsm_ctx* c, gpsm_kstat* out, uint32_t max, uint32_t* n) {
  if (!c || !n || (max && !out)) return GPSM_EINVAL;
  uint32_t k = 0;
  for (int i = 0; i < K_KINDS_ && k < max; ++i) {
    const Prof::Acc& a = c->prof.acc[i];
    if (!a.launches) continue;
    std::memset(&out[k], 0, sizeof(gpsm_kstat));
    std::snprintf(out[k].name, sizeof(out[k].name), "%s", kProfNames[i]);
    out[k].launches = a.launches;
    out[k].total_ms = a.ms;
    out[k].bytes = a.bytes;
    ++k;
  }
  *n = k;
  return GPSM_OK;
}

int gpsm_graph_cache_stats(const gpsm_ctx* c, uint64_t* instantiations, uint64_t* updates, uint64_t* replays) {
  if (!c) return GPSM_EINVAL;
  if (instantiations) *instantiations = c->graph_instantiations;
  if (updates) *updates = c->graph_updates;
  if (replays) *replays = c->graph_replays;
  return GPSM_OK;
}

const char* gpsm_strerror(int code) {
  switch (code) {
    case GPSM_OK: return "ok";
    case GPSM_EINVAL: return "invalid argument";
    case GPSM_ENOMEM: return "out of memory";
    case GPSM_EDEVICE: return "HIP runtime error";
    case GPSM_ETOOBIG: return "query larger than GPSM_MAX_QV";
    case GPSM_EOVERFLOW: return "tuple capacity exceeded";
    case GPSM_ENODEV: return "no HIP device";
    case GPSM_EDISCONN: return "query not connected or without edges";
    case GPSM_EFORMAT: return "malformed or unterminated input";
    case GPSM_EBUDGET: return "join work budget exceeded";
    case GPSM_ESCRATCH: return "input larger than the fixed scratch of an allocation-free entry point";
    default: return "unknown error";
  }
}

int gpsm_init(int device, gpsm_ctx** out) {
  if (!out) return GPSM_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return GPSM_ENODEV;
  }
  if (device < 0 || device >= count) return GPSM_EINVAL;
  HIPCHK(hipSetDevice(device));
  gpsm_ctx* c = new gpsm_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    if (prop.name[0])
      std::snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
    else
      std::snprintf(c->name, sizeof(c->name), "%s", prop.gcnArchName);
    c->ncu = prop.multiProcessorCount;
    c->hub_max_bytes = std::min<uint64_t>(48ull << 30, (uint64_t)prop.totalGlobalMem / 4);
  } else {
    c->ncu = 256;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return GPSM_EDEVICE;
  }
  // warm-up: first allocation and launch (match/Match.cu:37-45)
  if (c->bufs[B_SCRATCH].ensure(1 << 20) != GPSM_OK) {
    delete c;
    return GPSM_ENOMEM;
  }
  // this device's copy of the allocation-free scan's scratch array
  if (hipGetSymbolAddress(reinterpret_cast<void**>(&c->noalloc_scratch), HIP_SYMBOL(scratch)) != hipSuccess ||
      !c->noalloc_scratch) {
    set_error("hipGetSymbolAddress(scratch)", hipGetLastError());
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipEventDestroy(c->ev_fork);
    hipEventDestroy(c->ev_join);
    hipStreamDestroy(c->stream);
    hipStreamDestroy(c->stream2);
    c->bufs[B_SCRATCH].release();
    delete c;
    return GPSM_EDEVICE;
  }
  *out = c;
  return GPSM_OK;
}

void gpsm_shutdown(gpsm_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  hipStreamSynchronize(c->stream2);
  for (auto& e : c->gexec)
    if (e) hipGraphExecDestroy(e);
  for (auto& b : c->bufs) b.release();
  for (auto& b : c->h_stage) b.release();
  c->arena.release();
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  hipEventDestroy(c->ev_fork);
  hipEventDestroy(c->ev_join);
  hipStreamDestroy(c->stream);
  hipStreamDestroy(c->stream2);
  delete c;
}

const char* gpsm_device_name(const gpsm_ctx* c) { return c ? c->name : ""; }

int gpsm_scan_u32_noalloc(gpsm_ctx* c, const uint32_t* d_in, uint32_t* d_out, uint64_t n, uint64_t* total) {
  if (!c || (n && (!d_in || !d_out))) return GPSM_EINVAL;
  if (total) *total = 0;
  if (n == 0) return GPSM_OK;
  if (n > GPSM_SCAN_NOALLOC_MAX) return GPSM_ESCRATCH;
  if (hipSetDevice(c->device) != hipSuccess) return GPSM_EDEVICE;
  hipStream_t s = c->stream;
  // words 0-1: the total; from word 2: two words per block sum
  uint32_t* scr = c->noalloc_scratch;
  unsigned long long* dtot = reinterpret_cast<unsigned long long*>(scr);
  const uint64_t words = GPSM_SCAN_NOALLOC_SCRATCH_BYTES / sizeof(uint32_t);
  launch_exclusive_scan(d_in, d_out, n, dtot, scr + 2, words - 2, s);
  unsigned long long t = 0;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&t, dtot, sizeof(t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPSM_EDEVICE;
  if (total) *total = t;
  return GPSM_OK;
}

}  // extern "C"

namespace {

void free_graph_arrays(DevGraph& d) {
  hipFree(d.row);
  hipFree(d.col);
  hipFree(d.key);
  hipFree(d.ckey);
  hipFree(d.fchunk);
  hipFree(d.wrange);
  hipFree(d.hub_of);
  hipFree(d.hub_bits);
  d = DevGraph();
}

// The rest of a resident graph once row/col are on the device: label
// statistics, the label/degree keys, the filter's vertex ranges and the hub
// bitmaps.  row_offset and labels are host copies.  On failure g is freed.
int finish_graph(gpsm_ctx* c, gpsm_graph* g, const uint32_t* row_offset, const uint32_t* labels) {
  DevGraph& d = g->dg;
  const uint32_t n = d.n;
  auto fail = [&](int code) {
    free_graph_arrays(d);
    delete g;
    return code;
  };
  GraphStats& s = g->stats;
  s.n = n;
  s.adj = d.adj;
  std::vector<uint32_t> lv(labels, labels + n);
  std::sort(lv.begin(), lv.end());
  if (lv[0] == 0) return fail(GPSM_EINVAL);
  for (uint32_t i = 0; i < n; ++i) {
    if (i == 0 || lv[i] != lv[i - 1]) {
      s.label_value.push_back(lv[i]);
      s.label_count.push_back(0);
    }
    s.label_count.back()++;
  }
  if (s.label_value.size() > MAX_LABELS) return fail(GPSM_EINVAL);
  std::vector<uint32_t> key(n);
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t deg = row_offset[v + 1] - row_offset[v];
    key[v] = (s.dense(labels[v]) << KEY_DEG_BITS) | std::min(deg, KEY_DEG_MASK);
  }
  d.W = words_of(n);
  d.NB = nblocks_of(n);
  std::vector<uint32_t> fch;
  filter_chunks(n, row_offset, fch);
  d.nfchunk = (uint32_t)fch.size() - 1;
  std::vector<uint32_t> wr(d.W);
  for (uint32_t w = 0, r = 0; w < d.W; ++w) {
    while (fch[r + 1] <= 32 * w) ++r;
    wr[w] = r;
  }
  if (hipMalloc(&d.key, sizeof(uint32_t) * n) != hipSuccess ||
      hipMalloc(&d.ckey, sizeof(uint32_t) * std::max<uint64_t>(d.adj, 1)) != hipSuccess ||
      hipMalloc(&d.fchunk, sizeof(uint32_t) * fch.size()) != hipSuccess ||
      hipMalloc(&d.wrange, sizeof(uint32_t) * d.W) != hipSuccess) {
    (void)hipGetLastError();
    return fail(GPSM_ENOMEM);
  }
  if (hipMemcpyAsync(d.key, key.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(d.fchunk, fch.data(), sizeof(uint32_t) * fch.size(), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipMemcpyAsync(d.wrange, wr.data(), sizeof(uint32_t) * d.W, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    (void)hipGetLastError();
    return fail(GPSM_EDEVICE);
  }
  launch_neighbour_keys(d, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess || hipGetLastError() != hipSuccess) return fail(GPSM_EDEVICE);
  g->bytes = sizeof(uint32_t) * ((uint64_t)n + 1 + 2 * d.adj + n + fch.size() + d.W);
  if (int rc = build_hubs(c, row_offset, g)) return fail(rc);
  return GPSM_OK;
}

}  // namespace

extern "C" {

int gpsm_graph_upload(gpsm_ctx* c, uint32_t n, const uint32_t* row_offset, const uint32_t* col,
                      const uint32_t* labels, gpsm_graph** out) {
  if (!c || !out || !row_offset || !labels || n == 0) return GPSM_EINVAL;
  *out = nullptr;
  const uint64_t A = row_offset[n];
  if (A && !col) return GPSM_EINVAL;
  // validate the CSR: monotone offsets, sorted rows, no self-loops/duplicates
  if (row_offset[0] != 0) return GPSM_EINVAL;
  for (uint32_t v = 0; v < n; ++v) {
    if (row_offset[v + 1] < row_offset[v]) return GPSM_EINVAL;
    for (uint32_t e = row_offset[v]; e < row_offset[v + 1]; ++e) {
      if (col[e] >= n || col[e] == v) return GPSM_EINVAL;
      if (e > row_offset[v] && col[e] <= col[e - 1]) return GPSM_EINVAL;
    }
    if (labels[v] == 0) return GPSM_EINVAL;
  }
  hipSetDevice(c->device);
  gpsm_graph* g = new gpsm_graph();
  g->ctx = c;
  DevGraph& d = g->dg;
  d.n = n;
  d.adj = A;
  if (hipMalloc(&d.row, sizeof(uint32_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&d.col, sizeof(uint32_t) * std::max<uint64_t>(A, 1)) != hipSuccess) {
    (void)hipGetLastError();
    free_graph_arrays(d);
    delete g;
    return GPSM_ENOMEM;
  }
  if (hipMemcpyAsync(d.row, row_offset, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      (A && hipMemcpyAsync(d.col, col, sizeof(uint32_t) * A, hipMemcpyHostToDevice, c->stream) != hipSuccess)) {
    (void)hipGetLastError();
    free_graph_arrays(d);
    delete g;
    return GPSM_EDEVICE;
  }
  if (int rc = finish_graph(c, g, row_offset, labels)) return rc;
  *out = g;
  return GPSM_OK;
}

int gpsm_graph_upload_edges(gpsm_ctx* c, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                            const uint32_t* labels, gpsm_graph** out) {
  if (!c || !out || !labels || n == 0 || (m && (!src || !dst)) || 2 * m >= (1ull << 32)) return GPSM_EINVAL;
  *out = nullptr;
  hipSetDevice(c->device);
  gpsm_graph* g = new gpsm_graph();
  g->ctx = c;
  DevGraph& d = g->dg;
  d.n = n;
  uint64_t adj = 0;
  int rc = build_csr_from_edges(n, m, src, dst, c->stream, &d.row, &d.col, &adj);
  if (rc) {
    free_graph_arrays(d);
    delete g;
    return rc;
  }
  d.adj = adj;
  std::vector<uint32_t> ro(n + 1);
  if (hipMemcpyAsync(ro.data(), d.row, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    (void)hipGetLastError();
    free_graph_arrays(d);
    delete g;
    return GPSM_EDEVICE;
  }
  if ((rc = finish_graph(c, g, ro.data(), labels))) return rc;
  *out = g;
  return GPSM_OK;
}

int gpsm_graph_info(const gpsm_graph* g, uint32_t* n, uint64_t* adj) {
  if (!g) return GPSM_EINVAL;
  if (n) *n = g->dg.n;
  if (adj) *adj = g->dg.adj;
  return GPSM_OK;
}

int gpsm_graph_csr(const gpsm_graph* g, uint32_t* row_offset, uint32_t* col) {
  if (!g || !row_offset || (g->dg.adj && !col)) return GPSM_EINVAL;
  hipSetDevice(g->ctx->device);
  hipStream_t s = g->ctx->stream;
  if (hipMemcpyAsync(row_offset, g->dg.row, sizeof(uint32_t) * ((size_t)g->dg.n + 1), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      (g->dg.adj &&
       hipMemcpyAsync(col, g->dg.col, sizeof(uint32_t) * g->dg.adj, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipGetLastError();
    return GPSM_EDEVICE;
  }
  return GPSM_OK;
}

int gpsm_graph_free(gpsm_graph* g) {
  if (!g) return GPSM_EINVAL;
  hipSetDevice(g->ctx->device);
  hipStreamSynchronize(g->ctx->stream);
  free_graph_arrays(g->dg);
  g->directed.release();
  delete g;
  return GPSM_OK;
}

uint64_t gpsm_graph_device_bytes(const gpsm_graph* g) {
  if (!g) return 0;
  const uint64_t directed = g->directed.row ? 4 * ((uint64_t)g->dg.n + 1 + 2 * g->directed.arcs) : 0;
  return g->bytes + directed;
}

int gpsm_set_hub_policy(gpsm_ctx* c, uint32_t min_degree, uint64_t max_bytes) {
  if (!c || !min_degree) return GPSM_EINVAL;
  c->hub_min_degree = min_degree;
  c->hub_max_bytes = max_bytes;
  return GPSM_OK;
}

int gpsm_graph_hubs(const gpsm_graph* g, uint32_t* nhub, uint32_t* hub_degree) {
  if (!g) return GPSM_EINVAL;
  if (nhub) *nhub = g->dg.nhub;
  if (hub_degree) *hub_degree = g->dg.hub_degree;
  return GPSM_OK;
}

void gpsm_result_free(gpsm_result* r) {
  if (!r) return;
  std::free(r->tuples);
  r->tuples = nullptr;
  r->n_tuples = 0;
}

uint64_t gpsm_tuple_hash(const uint32_t* t, uint32_t q);  // gen.cpp

}  // extern "C"

// ---------------------------------------------------------------------------
// The batch pipeline.

namespace gpsm {

namespace {

// Device memory the join's level buffers may take (all workgroups together).
// A batch of 64-vertex queries at 4096 rows per level needs 8 MB per
// workgroup, so the bound matters only for very wide queries.
constexpr double kLevelBytes = 16.0 * (1ull << 30);

template <typename T>
T* dptr(gpsm_ctx* c, int b) {
  return reinterpret_cast<T*>(c->bufs[b].ptr);
}

template <typename T>
int upload(gpsm_ctx* c, int b, const std::vector<T>& v) {
  int rc = c->bufs[b].ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (rc) return rc;
  if (v.empty()) return GPSM_OK;
  // staged in the batch's pinned arena: the copy is queued, not waited for
  void* h = c->arena.take(sizeof(T) * v.size());
  if (!h) return GPSM_ENOMEM;
  std::memcpy(h, v.data(), sizeof(T) * v.size());
  if (hipMemcpyAsync(c->bufs[b].ptr, h, sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    return GPSM_EDEVICE;
  return GPSM_OK;
}

struct QState {
  int status = GPSM_OK;
  bool active = false;  // still matching (valid query, no empty candidate set)
  HostQuery hq;
  FilterSchedule fs;
  uint32_t vbase = 0;
  OrderPlan plan;
  std::vector<uint64_t> cet_size;  // forward |CET| per query edge
  uint32_t first_desc = 0;
};

// Ends a stream capture begun by the caller and launches it through the
// context's cached executable graph of that phase: updated in place when the
// captured topology matches, else (re)instantiated.  The captured graph is
// destroyed on every path.
int launch_captured(gpsm_ctx* c, int phase, hipStream_t s, const std::vector<uint64_t>& sig) {
  hipGraph_t graph = nullptr;
  c->gsig[phase].clear();
  if (hipStreamEndCapture(s, &graph) != hipSuccess || !graph) {
    (void)hipGetLastError();
    if (graph) hipGraphDestroy(graph);
    return GPSM_EDEVICE;
  }
  hipGraphExec_t& exec = c->gexec[phase];
  bool updated = false;
  if (exec) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult why;
    updated = hipGraphExecUpdate(exec, graph, &err_node, &why) == hipSuccess && why == hipGraphExecUpdateSuccess;
    if (!updated) {
      (void)hipGetLastError();
      hipGraphExecDestroy(exec);
      exec = nullptr;
    }
  }
  if (!exec) {
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGetLastError();
      exec = nullptr;
      hipGraphDestroy(graph);
      return GPSM_EDEVICE;
    }
    ++c->graph_instantiations;
  } else {
    ++c->graph_updates;
  }
  hipGraphDestroy(graph);
  if (hipGraphLaunch(exec, s) != hipSuccess) {
    (void)hipGetLastError();
    return GPSM_EDEVICE;
  }
  c->gsig[phase] = sig;
  return GPSM_OK;
}

// The argument list of a phase's launches, word by word: every value a
// kernel argument or a launch shape of the phase is computed from (buffer
// addresses, the data graph's fields, sizes, counts, tuning).  A batch whose
// list equals the one the cached executable graph was captured with
// relaunches that graph as it is: its kernels read the batch's own data from
// the same buffers, uploaded before the phase.
struct PhaseSig {
  std::vector<uint64_t> w;
  void add(uint64_t v) { w.push_back(v); }
  void add(const void* p) { w.push_back((uint64_t)(uintptr_t)p); }
  void add_graph(const DevGraph& g) {
    add(g.n), add(g.adj), add(g.W), add(g.NB), add(g.row), add(g.col), add(g.key), add(g.ckey), add(g.fchunk),
        add(g.nfchunk), add(g.wrange), add(g.hub_of), add(g.hub_bits), add(g.nhub), add(g.hub_degree);
  }
};

// Relaunches phase's cached graph when sig matches its capture; false when
// the phase has to be captured (or launched) afresh.
bool replay_cached(gpsm_ctx* c, int phase, const std::vector<uint64_t>& sig, hipStream_t s, int* rc) {
  *rc = GPSM_OK;
  if (!c->gexec[phase] || c->gsig[phase].empty() || c->gsig[phase] != sig) return false;
  if (hipGraphLaunch(c->gexec[phase], s) != hipSuccess) {
    (void)hipGetLastError();
    *rc = GPSM_EDEVICE;
  } else {
    ++c->graph_replays;
  }
  return true;
}

int predict_with_ctx(void* user, const float* x, uint32_t B, float* y) {
  return gpsm_mlp_forward(reinterpret_cast<gpsm_mlp*>(user), x, B, y);
}

}  // namespace

int run_batch(gpsm_ctx* c, const gpsm_graph* g, uint32_t nq, const gpsm_query* qs, const gpsm_opts* opts,
              gpsm_result* out, const BatchHooks* hooks) {
  const DevGraph& dg = g->dg;
  const GraphStats& st = g->stats;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  // the upload arena is reused from its start: no earlier copy may still read it
  HIPCHK(hipStreamSynchronize(s));
  c->arena.reset();
  gpsm_opts o{};
  if (opts) o = *opts;
  const uint32_t fblock = o.block ? o.block : c->tune.filter_block;
  // a level buffer must hold at least two expansion steps (2 x 4 x 256 rows)
  const uint32_t cap = std::max<uint32_t>(o.join_cap ? o.join_cap : c->tune.join_cap, 2048);
  const uint32_t wg_per_cu = c->tune.join_wg_per_cu;

  std::vector<QState> Q(nq);
  uint32_t nslots = 0, maxsteps = 0, qmax = 2;
  for (uint32_t b = 0; b < nq; ++b) {
    std::memset(&out[b], 0, sizeof(gpsm_result));
    QState& S = Q[b];
    S.status = S.hq.build(&qs[b]);
    out[b].q = qs[b].n;
    if (S.status != GPSM_OK) continue;
    make_filter_schedule(S.hq, st, S.fs);
    S.active = true;
    for (uint32_t u = 0; u < S.hq.n; ++u)
      if (st.dense(S.hq.lab[u]) == 0) S.active = false;  // a label absent from the data graph
    S.vbase = nslots;
    nslots += S.hq.n;
    maxsteps = std::max<uint32_t>(maxsteps, (uint32_t)S.fs.steps.size());
    qmax = std::max(qmax, S.hq.n);
  }
  const uint32_t W = dg.W, NB = dg.NB;
  int rc;
  // per-launch timing (not inside captured graphs)
  const bool prof = c->prof.mode > 0 && !o.use_graph;
  const bool counters = prof && c->prof.mode == 2;
  for (auto& p : c->prof.pending) {
    c->prof.pool.push_back(p.a);
    c->prof.pool.push_back(p.b);
  }
  c->prof.pending.clear();
  // mode 3 times the join alone: no events between the other launches
  const bool join_only = c->prof.mode == 3;
  auto pbegin = [&](int kind) -> hipEvent_t {
    if (!prof || (join_only && kind != K_JOIN)) return nullptr;
    hipEvent_t e = c->prof.get();
    hipEventRecord(e, s);
    return e;
  };
  auto pend = [&](int kind, hipEvent_t a, double bytes, int slot) {
    if (!prof || !a) return;
    hipEvent_t b = c->prof.get();
    hipEventRecord(b, s);
    c->prof.pending.push_back({kind, a, b, bytes, slot});
  };
  auto t0 = std::chrono::steady_clock::now();
  hipEventRecord(c->ev0, s);
  auto tphase = t0;
  auto wall = [&](int kind) {
    if (!prof) return;
    const auto now = std::chrono::steady_clock::now();
    Prof::Acc& acc = c->prof.acc[kind];
    acc.launches += 1;
    acc.ms += std::chrono::duration<double, std::milli>(now - tphase).count();
    tphase = now;
  };

  // ---------------- phase A: filter ----------------
  std::vector<InitOp> init;
  std::vector<FilterOp> fops;
  std::vector<uint32_t> step_off(maxsteps + 1, 0), nbr_slot, nbr_lo, nbr_hi;
  for (uint32_t b = 0; b < nq; ++b)
    if (Q[b].active) init.push_back({Q[b].vbase + Q[b].fs.start, Q[b].fs.start_lo, Q[b].fs.start_hi});
  for (uint32_t sidx = 0; sidx < maxsteps; ++sidx) {
    step_off[sidx] = (uint32_t)fops.size();
    for (uint32_t b = 0; b < nq; ++b) {
      const QState& S = Q[b];
      if (!S.active || sidx >= S.fs.steps.size()) continue;
      const FilterStep& fsx = S.fs.steps[sidx];
      FilterOp op;
      op.type = fsx.type;
      op.u_slot = S.vbase + fsx.u;
      op.nv = (uint32_t)fsx.v.size();
      op.nbr_off = (uint32_t)nbr_slot.size();
      op.mark = op.need = ~0u;
      if (fsx.type == OP_REFINE) {
        op.mark = b * 4 + fsx.round;
        if (fsx.round > 0) op.need = b * 4 + fsx.round - 1;
      }
      for (size_t i = 0; i < fsx.v.size(); ++i) {
        nbr_slot.push_back(S.vbase + fsx.v[i]);
        nbr_lo.push_back(fsx.type == OP_EXPLORE ? fsx.lo[i] : 0);
        nbr_hi.push_back(fsx.type == OP_EXPLORE ? fsx.hi[i] : 0);
      }
      fops.push_back(op);
    }
  }
  step_off[maxsteps] = (uint32_t)fops.size();
  if ((rc = c->bufs[B_BITMAPS].ensure(sizeof(uint32_t) * (size_t)std::max<uint32_t>(nslots, 1) * W))) return rc;
  if ((rc = upload(c, B_INIT, init)) || (rc = upload(c, B_FOPS, fops)) || (rc = upload(c, B_NBR_SLOT, nbr_slot)) ||
      (rc = upload(c, B_NBR_LO, nbr_lo)) || (rc = upload(c, B_NBR_HI, nbr_hi)))
    return rc;
  if ((rc = c->bufs[B_COUNTS].ensure(sizeof(uint32_t) * std::max<uint32_t>(nslots, 1)))) return rc;
  if ((rc = c->bufs[B_SEGC].ensure(sizeof(uint32_t) * std::max<uint64_t>((uint64_t)nslots * segs_of(W), 1))))
    return rc;
  uint32_t max_ops = 0;  // ops of the largest superstep
  for (uint32_t sidx = 0; sidx < maxsteps; ++sidx) max_ops = std::max(max_ops, step_off[sidx + 1] - step_off[sidx]);
  if ((rc = c->bufs[B_FITEMS].ensure(sizeof(uint2) * std::max<uint64_t>((uint64_t)max_ops * dg.nfchunk, 1))) ||
      (rc = c->bufs[B_FNITEMS].ensure(sizeof(uint32_t) * std::max<uint32_t>(maxsteps, 1))) ||
      (rc = c->bufs[B_RFLAGS].ensure(sizeof(uint32_t) * 4 * (uint64_t)std::max<uint32_t>(nq, 1))))
    return rc;
  // work counters (profile mode 2): one slot per filter superstep, then the
  // candidate-edge count of phase B, the count of phase C and the write
  const size_t fstat_step = (size_t)STAT_COPIES * STAT_STRIDE;
  const uint32_t nstat_slots = maxsteps + 4;
  const int slot_cet_b = (int)maxsteps + 1, slot_cet_c = (int)maxsteps + 2, slot_cet_w = (int)maxsteps + 3;
  if ((rc = c->bufs[B_FSTATS].ensure(sizeof(unsigned long long) * fstat_step * nstat_slots))) return rc;
  if ((rc = c->h_stage[0].ensure(sizeof(uint32_t) * std::max<uint32_t>(nslots, 1)))) return rc;
  uint32_t* bitmaps = dptr<uint32_t>(c, B_BITMAPS);
  unsigned long long* fstats = dptr<unsigned long long>(c, B_FSTATS);
  PhaseSig sig_a;
  if (o.use_graph) {
    sig_a.add(uint64_t{0});
    sig_a.add_graph(dg);
    for (int i : {B_BITMAPS, B_FNITEMS, B_RFLAGS, B_INIT, B_FOPS, B_NBR_SLOT, B_NBR_LO, B_NBR_HI, B_FITEMS, B_SEGC,
                  B_COUNTS})
      sig_a.add(c->bufs[i].ptr);
    sig_a.add(nslots), sig_a.add(maxsteps), sig_a.add(nq), sig_a.add(init.size()), sig_a.add(fblock);
    for (uint32_t v : step_off) sig_a.add(v);
  }
  bool replayed = o.use_graph && replay_cached(c, 0, sig_a.w, s, &rc);
  if (replayed && rc) return rc;
  if (!replayed) {
    bool captured = false;
    if (o.use_graph && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) captured = true;
    launch_fill_u32(bitmaps, 0u, (uint64_t)nslots * W, s);
    if (maxsteps) launch_fill_u32(dptr<uint32_t>(c, B_FNITEMS), 0u, maxsteps, s);
    launch_fill_u32(dptr<uint32_t>(c, B_RFLAGS), 0u, 4ull * std::max<uint32_t>(nq, 1), s);
    if (counters) launch_fill_u32(reinterpret_cast<uint32_t*>(fstats), 0u, 2ull * fstat_step * nstat_slots, s);
    {
      hipEvent_t a = pbegin(K_INIT);
      launch_init(dg, dptr<InitOp>(c, B_INIT), (uint32_t)init.size(), bitmaps, s);
      // the keys are read once for all ops, each op's bitmap written once
      pend(K_INIT, a, 4.0 * dg.n + (double)init.size() * 4.0 * W, -1);
    }
    for (uint32_t sidx = 0; sidx < maxsteps; ++sidx) {
      const uint32_t nops_s = step_off[sidx + 1] - step_off[sidx];
      hipEvent_t a = pbegin(K_FILTER);
      launch_filter_step(dg, dptr<FilterOp>(c, B_FOPS) + step_off[sidx], nops_s, dptr<uint32_t>(c, B_NBR_SLOT),
                         dptr<uint32_t>(c, B_NBR_LO), dptr<uint32_t>(c, B_NBR_HI), bitmaps, fblock,
                         65536, dptr<uint2>(c, B_FITEMS), dptr<uint32_t>(c, B_FNITEMS) + sidx,
                         dptr<uint32_t>(c, B_RFLAGS),
                         counters ? fstats + fstat_step * sidx : nullptr, s);
      pend(K_FILTER, a, 4.0 * W * nops_s, counters ? (int)sidx : -1);
    }
    launch_row_popcount(dg, bitmaps, nslots, dptr<uint32_t>(c, B_SEGC), dptr<uint32_t>(c, B_COUNTS), s);
    if (captured && (rc = launch_captured(c, 0, s, sig_a.w))) return rc;
  }
  {
    HIPCHK(hipGetLastError());
    std::vector<uint32_t>& counts = c->h_counts;
    counts.assign(std::max<uint32_t>(nslots, 1), 0);
    HIPCHK(hipMemcpyAsync(c->h_stage[0].ptr, dptr<uint32_t>(c, B_COUNTS), sizeof(uint32_t) * nslots,
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memcpy(counts.data(), c->h_stage[0].ptr, sizeof(uint32_t) * nslots);
  }
  wall(K_WALL_FILTER);
  const std::vector<uint32_t>& counts = c->h_counts;
  if (hooks && hooks->after_filter) {
    rc = hooks->after_filter(hooks->user, c, bitmaps, counts.data(), nslots, W);
    if (rc) return rc;
  }
  for (uint32_t b = 0; b < nq; ++b) {
    QState& S = Q[b];
    if (S.status != GPSM_OK) continue;
    for (uint32_t u = 0; u < S.hq.n && u < GPSM_MAX_QV; ++u) out[b].cand[u] = S.active ? counts[S.vbase + u] : 0;
    for (uint32_t u = 0; u < S.hq.n && S.active; ++u)
      if (counts[S.vbase + u] == 0) S.active = false;
  }

  // ---------------- phase B: compaction + forward |CET| ----------------
  std::vector<uint64_t> coff(std::max<uint32_t>(nslots, 1) + 1, 0);
  for (uint32_t r = 0; r < nslots; ++r) coff[r + 1] = coff[r] + counts[r];
  if ((rc = upload(c, B_COFF, coff))) return rc;
  if ((rc = c->bufs[B_CLIST].ensure(sizeof(uint32_t) * std::max<uint64_t>(coff[nslots], 1)))) return rc;
  if ((rc = c->bufs[B_BPREFIX].ensure(sizeof(uint32_t) * (size_t)std::max<uint32_t>(nslots, 1) * NB))) return rc;
  {
    hipEvent_t a = pbegin(K_COMPACT);
    launch_compact_rows(dg, bitmaps, nslots, dptr<uint32_t>(c, B_SEGC), dptr<uint64_t>(c, B_COFF),
                        dptr<uint32_t>(c, B_CLIST), dptr<uint32_t>(c, B_BPREFIX), s);
    pend(K_COMPACT, a, (double)nslots * 4.0 * (W + NB) + 4.0 * coff[nslots], -1);
  }
  std::vector<CetDesc> fdesc;
  std::vector<uint32_t> fdesc_of_q(nq, 0);
  for (uint32_t b = 0; b < nq; ++b) {
    QState& S = Q[b];
    fdesc_of_q[b] = (uint32_t)fdesc.size();
    if (!S.active) continue;
    for (size_t e = 0; e < S.hq.ef.size(); ++e) {
      CetDesc d{};
      d.x_slot = S.vbase + S.hq.ef[e];
      d.w_slot = S.vbase + S.hq.et[e];
      d.ncand = counts[d.x_slot];
      d.clist_off = coff[d.x_slot];
      key_range(S.hq, st, S.hq.et[e], d.w_lo, d.w_hi);
      fdesc.push_back(d);
    }
  }
  // The forward counts are kept per candidate (offset regions of every
  // query edge, back to back): a plan edge taken in the forward direction
  // reuses them in phase C, and only reversed plan edges are counted again.
  // The buffer is sized here for those too (at most one region of |C(to)|+1
  // words per query edge), so that phase C does not reallocate it.
  uint64_t fwd_words = 0, rev_bound = 0;
  for (auto& d : fdesc) {
    d.off_base = fwd_words;
    fwd_words += d.ncand + 1;
    rev_bound += counts[d.w_slot] + 1;
  }
  const uint32_t cgrid = std::max<uint32_t>(c->ncu * 8, 64);
  if ((rc = upload(c, B_FDESC, fdesc))) return rc;
  if ((rc = c->bufs[B_CHUNKS].ensure(sizeof(uint32_t) * (fdesc.size() + 2)))) return rc;
  if ((rc = c->bufs[B_OFFSETS].ensure(sizeof(uint32_t) * std::max<uint64_t>(fwd_words + rev_bound, 1)))) return rc;
  const uint64_t off_capacity = c->bufs[B_OFFSETS].bytes / sizeof(uint32_t);
  {
    hipEvent_t a = pbegin(K_CET_COUNT);
    launch_cet_count(dg, bitmaps, dptr<uint32_t>(c, B_CLIST), dptr<CetDesc>(c, B_FDESC), (uint32_t)fdesc.size(),
                     dptr<uint32_t>(c, B_OFFSETS), 1, dptr<uint32_t>(c, B_CHUNKS), cgrid, s,
                     counters ? fstats + fstat_step * slot_cet_b : nullptr);
    pend(K_CET_COUNT, a, 0.0, counters ? slot_cet_b : -1);
  }
  HIPCHK(hipGetLastError());
  if ((rc = c->h_stage[1].ensure(sizeof(CetDesc) * std::max<size_t>(fdesc.size(), 1)))) return rc;
  if (!fdesc.empty())
    HIPCHK(hipMemcpyAsync(c->h_stage[1].ptr, dptr<CetDesc>(c, B_FDESC), sizeof(CetDesc) * fdesc.size(),
                          hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (!fdesc.empty()) std::memcpy(fdesc.data(), c->h_stage[1].ptr, sizeof(CetDesc) * fdesc.size());
  wall(K_WALL_CET);
  if (hooks && hooks->after_cet_count) {
    rc = hooks->after_cet_count(hooks->user, c, fdesc.data(), (uint32_t)fdesc.size());
    if (rc) return rc;
  }

  // ---------------- host: join orders ----------------
  std::vector<CetDesc> jdesc;
  std::vector<JoinPlan> plans;
  std::vector<JoinTask> tasks, tasks_cyc;
  std::vector<uint32_t> plan_of_q(nq, ~0u);
  uint64_t off_words = fwd_words, list_words = 0, term_words = 0;  // reversed tables' regions follow the forward ones
  std::vector<uint32_t> recount;  // tables of reversed plan edges (count again)
  std::vector<uint32_t> term_desc;   // tables of last join levels
  std::vector<uint64_t> term_start;  // their term ranges, back to back (+ total)
  for (uint32_t b = 0; b < nq; ++b) {
    QState& S = Q[b];
    if (!S.active) continue;
    const size_t E = S.hq.ef.size();
    S.cet_size.resize(E);
    for (size_t e = 0; e < E; ++e) S.cet_size[e] = fdesc[fdesc_of_q[b] + e].total;
  }
  const bool learned = o.order_mode == GPSM_ORDER_LEARNED && c->planner;
  // planning statistics of every query; est: predicted join work of its plan
  std::vector<std::vector<uint32_t>> cand(nq);
  std::vector<PlanContext> pcs(nq);
  std::vector<double> est(nq, 0.0);
  std::vector<const PlanContext*> live;
  std::vector<uint32_t> live_b;
  for (uint32_t b = 0; b < nq; ++b) {
    QState& S = Q[b];
    if (!S.active) continue;
    cand[b].resize(S.hq.n);
    for (uint32_t u = 0; u < S.hq.n; ++u) cand[b][u] = counts[S.vbase + u];
    pcs[b] = PlanContext{&S.hq, &st, &cand[b], &S.cet_size};
    live.push_back(&pcs[b]);
    live_b.push_back(b);
  }
  if (learned) {
    // a query whose heuristic plan predicts little join work keeps it (its
    // join takes less time than ranking its candidate orders would); the
    // network still prices it, so that every query's task share (below)
    // comes from the same cost model
    std::vector<uint8_t> rank(live_b.size(), 0);
    for (size_t i = 0; i < live_b.size(); ++i) {
      QState& S = Q[live_b[i]];
      OrderPlan h;
      heuristic_order(S.hq, S.cet_size, st.adj + 1, h);
      rank[i] = analytic_cost(pcs[live_b[i]], h) >= kLearnedMinWork;
    }
    // every planning round of the whole batch is one network call
    std::vector<OrderPlan> chosen_plan;
    std::vector<double> cost;
    const int prc = learned_plans(
        live,
        [](void* u, const float* x, uint32_t B, float* y) {
          return gpsm_mlp_forward(static_cast<gpsm_mlp*>(u), x, B, y);
        },
        c->planner, st.adj + 1, chosen_plan, nullptr, &cost, &rank);
    if (prc) return prc;
    for (size_t i = 0; i < live_b.size(); ++i) {
      Q[live_b[i]].plan = std::move(chosen_plan[i]);
      est[live_b[i]] = cost[i];
    }
  }
  for (uint32_t b = 0; b < nq; ++b) {
    QState& S = Q[b];
    if (!S.active) continue;
    const size_t E = S.hq.ef.size();
    bool ok = true;
    if (o.order_mode == GPSM_ORDER_EXPLICIT) {
      ok = o.explicit_order &&
           order_from_vertices(S.hq, S.cet_size, o.explicit_order + (size_t)b * GPSM_MAX_QV, S.plan);
      if (!ok) S.status = GPSM_EINVAL;
    } else if (learned) {
      // planned above
    } else {
      heuristic_order(S.hq, S.cet_size, st.adj + 1, S.plan);
    }
    if (!ok) {
      S.active = false;
      continue;
    }
    for (uint32_t p = 0; p < S.hq.n && p < GPSM_MAX_QV; ++p) out[b].order[p] = S.plan.order[p];
    if (!learned) est[b] = analytic_cost(pcs[b], S.plan);
    // any empty forward table means no embedding
    bool empty = false;
    for (size_t e = 0; e < E; ++e) empty |= S.cet_size[e] == 0;
    if (empty) {
      S.active = false;
      continue;
    }
    JoinPlan jp{};
    jp.q = S.hq.n;
    jp.vbase = S.vbase;
    auto add_desc = [&](uint32_t xq, uint32_t wq, uint32_t edge) {
      CetDesc d{};
      d.x_slot = S.vbase + xq;
      d.w_slot = S.vbase + wq;
      d.ncand = counts[d.x_slot];
      d.clist_off = coff[d.x_slot];
      if (xq == S.hq.ef[edge]) {
        d.off_base = fdesc[fdesc_of_q[b] + edge].off_base;  // its counts are there since phase B
      } else {
        d.off_base = off_words;
        off_words += d.ncand + 1;
        recount.push_back((uint32_t)jdesc.size());
      }
      d.list_base = list_words;
      d.list_cap = (uint32_t)S.cet_size[edge];
      d.term_base = ~0ull;
      key_range(S.hq, st, wq, d.w_lo, d.w_hi);
      list_words += S.cet_size[edge];
      jdesc.push_back(d);
      return (uint32_t)jdesc.size() - 1;
    };
    const OrderPlan& pl = S.plan;
    jp.first_cet = add_desc(pl.order[0], pl.order[1], pl.first_edge);
    S.first_desc = jp.first_cet;
    uint32_t nb = 0;
    for (uint32_t k = 0; k < S.hq.n; ++k) {
      jp.order[k] = pl.order[k];
      jp.back_off[k] = nb;
      if (k >= 2) {
        jp.par[k] = pl.par_pos[k];
        jp.cet[k] = add_desc(pl.order[pl.par_pos[k]], pl.order[k], pl.par_edge[k]);
        if (k + 1 == S.hq.n) {
          CetDesc& d = jdesc.back();
          d.term_base = term_words;
          d.term_qid = pl.order[k];
          term_desc.push_back(jp.cet[k]);
          term_start.push_back(term_words);
          term_words += S.cet_size[pl.par_edge[k]];
        }
        for (uint32_t x : pl.backs[k]) jp.backs[nb++] = (uint8_t)x;
        uint64_t inj = 0;
        for (uint32_t c = 0; c < k; ++c)
          if (S.hq.lab[pl.order[c]] == S.hq.lab[pl.order[k]]) inj |= 1ull << c;
        inj &= ~(1ull << pl.par_pos[k]);
        for (uint32_t x : pl.backs[k]) inj &= ~(1ull << x);
        jp.inj[k] = inj;
      }
    }
    jp.back_off[S.hq.n] = nb;
    jp.nback_total = nb;
    plan_of_q[b] = (uint32_t)plans.size();
    plans.push_back(jp);
  }
  // tasks: interleaved row sets of the first tables.  A query gets tasks in
  // proportion to its predicted join work (tasks_per_wg tasks per workgroup
  // over the batch, 32 by default), and at least one per tsize first-table
  // rows, so a query whose few first rows fan out to most of the batch's
  // work does not leave a handful of long tasks running after the rest of
  // the grid is idle.  The persistent grid holds at most as many workgroups
  // as are resident at once: a workgroup beyond that would start only once
  // the task counter has run out.
  const uint32_t nwg = std::max<uint32_t>(1, c->ncu * std::min(wg_per_cu, join_resident_per_cu(qmax, false)));
  const uint32_t nwg_cyc = std::max<uint32_t>(1, c->ncu * std::min(wg_per_cu, join_resident_per_cu(qmax, true)));
  const uint64_t tpw = std::max<uint32_t>(1, c->tune.tasks_per_wg);
  uint64_t rows_total = 0;
  double est_total = 0;
  for (uint32_t b = 0; b < nq; ++b)
    if (plan_of_q[b] != ~0u) {
      rows_total += Q[b].cet_size[Q[b].plan.first_edge];
      est_total += est[b];
    }
  const uint64_t tsize =
      std::min<uint64_t>(std::max<uint64_t>(rows_total / ((uint64_t)nwg * tpw) + 1, c->tune.task_min), 1u << 20);
  const double target = est_total / ((double)nwg * tpw);
  for (uint32_t b = 0; b < nq; ++b) {
    if (plan_of_q[b] == ~0u) continue;
    const uint64_t R1 = Q[b].cet_size[Q[b].plan.first_edge];
    uint64_t T = (R1 + tsize - 1) / tsize;
    if (target > 0 && est[b] > target) T = std::max<uint64_t>(T, (uint64_t)std::min(est[b] / target + 1.0, 1e9));
    T = std::min<uint64_t>(std::max<uint64_t>(T, 1), R1);
    // tree plans' tasks first, the cyclic plans' after them (two launches)
    std::vector<JoinTask>& dst = plans[plan_of_q[b]].nback_total ? tasks_cyc : tasks;
    for (uint64_t t = 0; t < T; ++t) dst.push_back({plan_of_q[b], (uint32_t)T, t, (R1 - t + T - 1) / T});
  }
  const uint32_t ntree = (uint32_t)tasks.size(), ncyc = (uint32_t)tasks_cyc.size();
  tasks.insert(tasks.end(), tasks_cyc.begin(), tasks_cyc.end());

  // ---------------- phase C: candidate-edge tables ----------------
  wall(K_WALL_PLAN);
  term_start.push_back(term_words);
  if ((rc = upload(c, B_JDESC, jdesc)) || (rc = upload(c, B_PLANS, plans)) || (rc = upload(c, B_TASKS, tasks)) ||
      (rc = upload(c, B_TERM_DESC, term_desc)) || (rc = upload(c, B_TERM_START, term_start)))
    return rc;
  if ((rc = c->bufs[B_TERMS].ensure(sizeof(uint64_t) * std::max<uint64_t>(term_words, 1)))) return rc;
  std::vector<CetDesc> cdesc;
  for (uint32_t i : recount) cdesc.push_back(jdesc[i]);
  if (off_words > off_capacity) return GPSM_EINVAL;  // cannot happen: rev_bound covers every reversed table
  if ((rc = upload(c, B_CDESC, cdesc)) ||
      (rc = c->bufs[B_LISTS].ensure(sizeof(uint32_t) * std::max<uint64_t>(list_words, 1))) ||
      (rc = c->bufs[B_CHUNKS].ensure(sizeof(uint32_t) * (jdesc.size() + 2))))
    return rc;
  const uint32_t nplans = (uint32_t)plans.size();
  if ((rc = c->bufs[B_JOUT].ensure(sizeof(JoinOut) * std::max<uint32_t>(nplans, 1))) ||
      (rc = c->h_stage[2].ensure(sizeof(JoinOut) * std::max<uint32_t>(nplans, 1))))
    return rc;
  if ((rc = c->bufs[B_CTR].ensure(256))) return rc;
  // Level buffers of the workgroups launched (the tree plans' launch, then
  // the cyclic plans', which run at the same time), within kLevelBytes: rows
  // per level halve (to 2048) and then workgroups are dropped, from both
  // launches in proportion, until they fit.
  uint32_t gt = ntree ? std::min<uint32_t>(nwg, ntree) : 0, gc = ncyc ? std::min<uint32_t>(nwg_cyc, ncyc) : 0;
  uint32_t lcap = cap;
  const uint64_t gsum = std::max<uint64_t>(1, (uint64_t)gt + gc);
  while (lcap > 2048 && (double)join_level_words(qmax, lcap) * 4.0 * gsum > kLevelBytes) lcap /= 2;
  const uint64_t lw = join_level_words(qmax, lcap);
  if (lw && (double)lw * 4.0 * gsum > kLevelBytes) {
    const uint64_t fit = std::max<uint64_t>(2, (uint64_t)(kLevelBytes / (4.0 * lw)));
    if (gt) gt = (uint32_t)std::max<uint64_t>(1, (uint64_t)gt * fit / gsum);
    if (gc) gc = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)gc * fit / gsum + 1, fit - gt));
  }
  if ((rc = c->bufs[B_LEVELS].ensure(sizeof(uint32_t) * std::max<uint64_t>(lw * ((uint64_t)gt + gc), 1)))) return rc;
  uint64_t emit_cap = 0;
  if (o.want_tuples && nq == 1) {
    emit_cap = o.tuple_cap;
    if ((rc = c->bufs[B_EMIT].ensure(sizeof(uint32_t) * std::max<uint64_t>(emit_cap * qmax, 1)))) return rc;
  }
  PhaseSig sig_c;
  if (o.use_graph) {
    sig_c.add(uint64_t{1});
    sig_c.add_graph(dg);
    for (int i : {B_BITMAPS, B_CLIST, B_CDESC, B_OFFSETS, B_CHUNKS, B_JDESC, B_LISTS, B_TERM_DESC, B_TERM_START,
                  B_TERMS, B_JOUT, B_CTR})
      sig_c.add(c->bufs[i].ptr);
    for (uint64_t v : {(uint64_t)cdesc.size(), (uint64_t)cgrid, (uint64_t)jdesc.size(), (uint64_t)term_desc.size(),
                       (uint64_t)term_words, (uint64_t)nplans})
      sig_c.add(v);
  }
  bool replayed_c = o.use_graph && replay_cached(c, 1, sig_c.w, s, &rc);
  if (replayed_c && rc) return rc;
  if (!replayed_c) {
    bool captured = false;
    if (o.use_graph && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) captured = true;
    hipEvent_t pa = pbegin(K_CET_COUNT);
    if (!cdesc.empty())
      launch_cet_count(dg, bitmaps, dptr<uint32_t>(c, B_CLIST), dptr<CetDesc>(c, B_CDESC), (uint32_t)cdesc.size(),
                       dptr<uint32_t>(c, B_OFFSETS), 1, dptr<uint32_t>(c, B_CHUNKS), cgrid, s,
                       counters ? fstats + fstat_step * slot_cet_c : nullptr);
    pend(K_CET_COUNT, pa, 0.0, counters ? slot_cet_c : -1);
    launch_cet_offsets(dptr<CetDesc>(c, B_JDESC), (uint32_t)jdesc.size(), dptr<uint32_t>(c, B_OFFSETS), s);
    pa = pbegin(K_CET_WRITE);
    launch_cet_write(dg, bitmaps, dptr<uint32_t>(c, B_CLIST), dptr<CetDesc>(c, B_JDESC), (uint32_t)jdesc.size(),
                     dptr<uint32_t>(c, B_OFFSETS), dptr<uint32_t>(c, B_LISTS), dptr<uint32_t>(c, B_CHUNKS), cgrid, s,
                     counters ? fstats + fstat_step * slot_cet_w : nullptr);
    launch_cet_terms(dptr<uint32_t>(c, B_LISTS), dptr<CetDesc>(c, B_JDESC), dptr<uint32_t>(c, B_TERM_DESC),
                     dptr<uint64_t>(c, B_TERM_START), (uint32_t)term_desc.size(), term_words,
                     dptr<unsigned long long>(c, B_TERMS), s);
    pend(K_CET_WRITE, pa, 0.0, counters ? slot_cet_w : -1);
    launch_fill_u32(dptr<uint32_t>(c, B_JOUT), 0u, sizeof(JoinOut) / 4 * (uint64_t)std::max<uint32_t>(nplans, 1), s);
    launch_fill_u32(dptr<uint32_t>(c, B_CTR), 0u, 64, s);
    if (captured && (rc = launch_captured(c, 1, s, sig_c.w))) return rc;
  }
  // ---------------- phase D: join ----------------
  // launched directly, also with opts.use_graph: the two launches run side
  // by side on two streams, which a captured graph's replay does not keep
  {
    JoinArgs ja{};
    ja.row = dg.row;
    ja.col = dg.col;
    ja.hub_of = dg.hub_of;
    ja.hub_bits = dg.hub_bits;
    ja.bitmaps = bitmaps;
    ja.bprefix = dptr<uint32_t>(c, B_BPREFIX);
    ja.clist = dptr<uint32_t>(c, B_CLIST);
    ja.offsets = dptr<uint32_t>(c, B_OFFSETS);
    ja.lists = dptr<uint32_t>(c, B_LISTS);
    ja.terms = dptr<unsigned long long>(c, B_TERMS);
    ja.cets = dptr<CetDesc>(c, B_JDESC);
    ja.plans = dptr<JoinPlan>(c, B_PLANS);
    ja.tasks = dptr<JoinTask>(c, B_TASKS);
    ja.ntasks = (uint32_t)tasks.size();
    ja.W = W;
    ja.NB = NB;
    ja.cap = lcap;
    ja.qmax = qmax;
    ja.levels = dptr<uint32_t>(c, B_LEVELS);
    ja.level_stride = lw;
    ja.task_counter = dptr<unsigned int>(c, B_CTR);
    ja.out = dptr<JoinOut>(c, B_JOUT);
    if (emit_cap) {
      ja.emit = dptr<uint32_t>(c, B_EMIT);
      ja.emit_ctr = reinterpret_cast<unsigned long long*>(dptr<uint32_t>(c, B_CTR) + 16);
      ja.emit_cap = emit_cap;
    }
    ja.budget = o.work_budget;
    ja.count_bytes = counters ? 1u : 0u;
    // tree plans and cyclic plans: one launch each, its own task counter;
    // timed together as one join
    hipEvent_t pa = pbegin(K_JOIN);
    // the cyclic plans on stream2 (own task counter and level buffers),
    // forked after everything before on stream and joined back before what
    // follows; the tree plans on stream
    if (ncyc) {
      JoinArgs jc = ja;
      jc.tasks = dptr<JoinTask>(c, B_TASKS) + ntree;
      jc.ntasks = ncyc;
      jc.task_counter = dptr<unsigned int>(c, B_CTR) + 1;
      jc.levels = dptr<uint32_t>(c, B_LEVELS) + lw * gt;
      hipStream_t s2 = ntree ? c->stream2 : s;
      if (ntree) {
        HIPCHK(hipEventRecord(c->ev_fork, s));
        HIPCHK(hipStreamWaitEvent(s2, c->ev_fork, 0));
      }
      launch_join(jc, gc, true, s2);
      if (ntree) HIPCHK(hipEventRecord(c->ev_join, s2));
    }
    if (ntree) {
      ja.ntasks = ntree;
      launch_join(ja, gt, false, s);
    }
    if (ncyc && ntree) HIPCHK(hipStreamWaitEvent(s, c->ev_join, 0));
    pend(K_JOIN, pa, 0.0, -1);
    HIPCHK(hipGetLastError());
    std::vector<JoinOut> jo(std::max<uint32_t>(nplans, 1));
    HIPCHK(hipMemcpyAsync(c->h_stage[2].ptr, dptr<JoinOut>(c, B_JOUT), sizeof(JoinOut) * nplans,
                          hipMemcpyDeviceToHost, s));
    unsigned long long emitted = 0;
    if (emit_cap)
      HIPCHK(hipMemcpyAsync(&emitted, dptr<uint32_t>(c, B_CTR) + 16, sizeof(emitted), hipMemcpyDeviceToHost, s));
    hipEventRecord(c->ev1, s);
    HIPCHK(hipStreamSynchronize(s));
    if (nplans) std::memcpy(jo.data(), c->h_stage[2].ptr, sizeof(JoinOut) * nplans);
    wall(K_WALL_JOIN);
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (prof) {
      std::vector<unsigned long long> fs(fstat_step * nstat_slots, 0);
      if (counters)
        HIPCHK(hipMemcpy(fs.data(), fstats, sizeof(unsigned long long) * fs.size(), hipMemcpyDeviceToHost));
      double join_bytes = 0;
      for (uint32_t i = 0; i < nplans; ++i) join_bytes += (double)jo[i].bytes;
      for (auto& pnd : c->prof.pending)
        if (pnd.kind == K_JOIN) pnd.bytes += join_bytes;
      for (auto& pnd : c->prof.pending) {
        float kms = 0;
        hipEventElapsedTime(&kms, pnd.a, pnd.b);
        double bytes = pnd.bytes;
        if (pnd.stats_slot >= 0) {
          unsigned long long f[5] = {0, 0, 0, 0, 0};
          for (uint32_t cp = 0; cp < STAT_COPIES; ++cp)
            for (int i = 0; i < 5; ++i) f[i] += fs[fstat_step * pnd.stats_slot + cp * STAT_STRIDE + i];
          if (pnd.kind == K_FILTER) {
            // candidates: two row offsets; entries: column index or neighbour
            // key; neighbour keys read beside the column index (explore's
            // second pass); bitmap probes and bit writes count one bit each
            bytes += 8.0 * f[0] + 4.0 * f[1] + 4.0 * f[2] + (f[3] + f[4]) / 8.0;
          } else {
            // candidate-edge count / write: per candidate its id and two row
            // offsets, plus its count written (count) or its list offset read
            // (write); per adjacency entry the column index and the
            // neighbour's key; one bit per probe of C(w); 4 per list entry
            // written
            bytes += 16.0 * f[0] + 8.0 * f[1] + f[2] / 8.0 + 4.0 * f[3];
          }
        }
        Prof::Acc& acc = c->prof.acc[pnd.kind];
        acc.launches += 1;
        acc.ms += kms;
        acc.bytes += bytes;
        c->prof.pool.push_back(pnd.a);
        c->prof.pool.push_back(pnd.b);
      }
      c->prof.pending.clear();
    }
    for (uint32_t b = 0; b < nq; ++b) {
      out[b].status = Q[b].status;
      out[b].device_ms = ms;
      if (plan_of_q[b] == ~0u) continue;
      const JoinOut& r = jo[plan_of_q[b]];
      out[b].count = r.count;
      out[b].hash = r.hash;
      for (uint32_t k = 1; k < Q[b].hq.n && k < GPSM_MAX_QV; ++k) out[b].level_rows[k] = r.level_rows[k];
      out[b].join_work = r.work;
      for (uint32_t k = 1; k < Q[b].hq.n && k < GPSM_MAX_QV; ++k) out[b].level_work[k] = r.level_work[k];
      if (r.aborted) {
        // abandoned part-way: no partial count or hash
        out[b].status = GPSM_EBUDGET;
        out[b].count = 0;
        out[b].hash = 0;
      }
    }
    if (emit_cap && nq == 1 && plan_of_q[0] != ~0u) {
      const uint32_t qn = Q[0].hq.n;
      if (emitted > emit_cap) {
        out[0].status = GPSM_EOVERFLOW;
      } else {
        std::vector<uint32_t> tup(emitted * qn);
        if (emitted)
          HIPCHK(hipMemcpy(tup.data(), dptr<uint32_t>(c, B_EMIT), sizeof(uint32_t) * tup.size(),
                           hipMemcpyDeviceToHost));
        std::vector<uint64_t> idx(emitted);
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](uint64_t x, uint64_t y) {
          return std::lexicographical_compare(tup.begin() + x * qn, tup.begin() + (x + 1) * qn, tup.begin() + y * qn,
                                              tup.begin() + (y + 1) * qn);
        });
        out[0].tuples = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<uint64_t>(emitted * qn, 1));
        if (!out[0].tuples) return GPSM_ENOMEM;
        for (uint64_t i = 0; i < emitted; ++i)
          std::memcpy(out[0].tuples + i * qn, tup.data() + idx[i] * qn, sizeof(uint32_t) * qn);
        out[0].n_tuples = emitted;
      }
    }
  }
  if (hooks && hooks->after_join) hooks->after_join(hooks->user, c, &Q[0].first_desc, nq);
  (void)t0;
  return GPSM_OK;
}

}  // namespace gpsm

extern "C" {

int gpsm_match(gpsm_ctx* c, const gpsm_graph* g, const gpsm_query* q, const gpsm_opts* opts, gpsm_result* out) {
  if (!c || !g || !q || !out) return GPSM_EINVAL;
  int rc = run_batch(c, g, 1, q, opts, out, nullptr);
  if (rc) return rc;
  return out->status;
}

int gpsm_match_batch(gpsm_ctx* c, const gpsm_graph* g, uint32_t nq, const gpsm_query* qs, const gpsm_opts* opts,
                     gpsm_result* out) {
  if (!c || !g || (nq && (!qs || !out))) return GPSM_EINVAL;
  if (nq == 0) return GPSM_OK;
  gpsm_opts o{};
  if (opts) o = *opts;
  o.want_tuples = 0;
  return run_batch(c, g, nq, qs, &o, out, nullptr);
}

}  // extern "C"

extern "C" {

int gpsm_tuning(const gpsm_ctx* c, uint32_t* fb, uint32_t* cap, uint32_t* wg, uint32_t* tpw) {
  if (!c) return GPSM_EINVAL;
  if (fb) *fb = c->tune.filter_block;
  if (cap) *cap = c->tune.join_cap;
  if (wg) *wg = c->tune.join_wg_per_cu;
  if (tpw) *tpw = c->tune.tasks_per_wg;
  return GPSM_OK;
}

int gpsm_set_tuning(gpsm_ctx* c, uint32_t fb, uint32_t cap, uint32_t wg, uint32_t tpw) {
  if (!c || (fb != 64 && fb != 128 && fb != 256) || cap < 2048 || cap > (1u << 20) || wg < 1 || wg > 16 ||
      tpw < 1 || tpw > 256)
    return GPSM_EINVAL;
  c->tune.filter_block = fb;
  c->tune.join_cap = cap;
  c->tune.join_wg_per_cu = wg;
  c->tune.tasks_per_wg = tpw;
  return GPSM_OK;
}

int gpsm_autotune(gpsm_ctx* c, const gpsm_graph* g, uint32_t nq, const gpsm_query* qs, const gpsm_opts* opts,
                  double* best_ms) {
  if (!c || !g || !nq || !qs) return GPSM_EINVAL;
  gpsm_opts o{};
  if (opts) o = *opts;
  o.want_tuples = 0;
  o.block = 0;
  o.join_cap = 0;
  o.use_graph = 0;  // the trials time launches with HIP events, which captured graphs do not record
  std::vector<gpsm_result> ref(nq), out(nq);
  // The join's settings are compared by the join kernel's own time (HIP
  // events, profile mode 3) and the filter's by the batch's wall time: the
  // other phases' run-to-run variation is as large as the difference
  // between neighbouring join settings.  The caller's profiling state is
  // put back afterwards.
  const int saved_mode = c->prof.mode;
  Prof::Acc saved_acc[sizeof(c->prof.acc) / sizeof(c->prof.acc[0])];
  std::copy(std::begin(c->prof.acc), std::end(c->prof.acc), saved_acc);
  auto restore = [&]() {
    c->prof.mode = saved_mode;
    std::copy(std::begin(saved_acc), std::end(saved_acc), std::begin(c->prof.acc));
  };
  bool join_param = false;
  auto trial = [&](double& ms, bool check) -> int {
    // best of three runs (the first also warms the buffers up)
    double best = 1e300;
    for (int rep = 0; rep < 3; ++rep) {
      c->prof.mode = 3;
      for (auto& a : c->prof.acc) a = Prof::Acc();
      auto t0 = std::chrono::steady_clock::now();
      int rc = run_batch(c, g, nq, qs, &o, check ? out.data() : ref.data(), nullptr);
      if (rc) return rc;
      const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      best = std::min(best, join_param ? c->prof.acc[K_JOIN].ms : wall);
    }
    ms = best;
    if (check)
      for (uint32_t i = 0; i < nq; ++i)
        if (out[i].count != ref[i].count || out[i].hash != ref[i].hash) return GPSM_EDEVICE;
    return GPSM_OK;
  };
  double cur;
  int rc = trial(cur, false);
  if (rc) {
    restore();
    return rc;
  }
  // join workgroups per CU above the number resident at once change nothing
  // (run_batch caps them), so the search stops at 8
  const uint32_t values[4][4] = {{64, 128, 256, 0}, {2048, 4096, 8192, 0}, {2, 4, 6, 8}, {8, 16, 32, 64}};
  for (int param = 0; param < 4; ++param) {
    if (param == 1) {
      // from here on the join's time decides; the current setting first
      join_param = true;
      if ((rc = trial(cur, true))) {
        restore();
        return rc;
      }
    }
    Tuning best = c->tune;
    for (int i = 0; i < 4 && values[param][i]; ++i) {
      Tuning t = best;
      if (param == 0) t.filter_block = values[param][i];
      if (param == 1) t.join_cap = values[param][i];
      if (param == 2) t.join_wg_per_cu = values[param][i];
      if (param == 3) t.tasks_per_wg = values[param][i];
      Tuning keep = c->tune;
      c->tune = t;
      double ms;
      rc = trial(ms, true);
      if (rc) {
        c->tune = keep;
        restore();
        return rc;
      }
      if (ms < cur) {
        cur = ms;
        best = t;
      }
      c->tune = keep;
    }
    c->tune = best;
  }
  restore();
  if (best_ms) *best_ms = cur;
  return GPSM_OK;
}

}  // extern "C"
