// The reference's final-set check (IO::verify, io/IO.cpp:146-180) on the
// device: injectivity, then every directed, edge-labelled query edge
// (from -> to, label) must be a data edge t[from] -> t[to] with that label.
// The matcher's own semantics are undirected and unlabelled (its result set
// is the superset these checks filter); this entry point recovers the
// reference's semantics for directed, edge-labelled inputs.
//
// Data-graph side: the out-edges of every vertex, sorted by (target, label),
// attached to a resident graph.  One thread per tuple; an edge test is a
// binary search for the target in the source's out-row, then a scan of the
// (usually one) entries with that target for the label.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "api_internal.h"
#include "gpsm.h"

using namespace gpsm;

namespace {

__device__ __forceinline__ bool has_labelled_arc(const uint32_t* __restrict__ orow, const uint32_t* __restrict__ ocol,
                                                 const uint32_t* __restrict__ olab, uint32_t a, uint32_t b,
                                                 uint32_t label) {
  uint32_t lo = orow[a], hi = orow[a + 1];
  while (lo < hi) {  // first entry with target >= b
    const uint32_t mid = (lo + hi) >> 1;
    if (ocol[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  for (uint32_t e = lo; e < orow[a + 1] && ocol[e] == b; ++e)
    if (olab[e] == label) return true;
  return false;
}

__global__ void verify_kernel(const uint32_t* __restrict__ orow, const uint32_t* __restrict__ ocol,
                              const uint32_t* __restrict__ olab, uint32_t n, const uint32_t* __restrict__ qsrc,
                              const uint32_t* __restrict__ qdst, const uint32_t* __restrict__ qlab, uint32_t m,
                              uint32_t q, const uint32_t* __restrict__ tuples, uint64_t nt,
                              uint8_t* __restrict__ keep, unsigned long long* __restrict__ kept) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long mine = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += stride) {
    const uint32_t* t = tuples + i * q;
    bool ok = true;
    for (uint32_t a = 0; a < q && ok; ++a) ok = t[a] < n;
    for (uint32_t a = 0; a < q && ok; ++a)
      for (uint32_t b = a + 1; b < q && ok; ++b) ok = t[a] != t[b];  // IO::verify, io/IO.cpp:149-160
    for (uint32_t e = 0; e < m && ok; ++e)                              // io/IO.cpp:162-178
      ok = has_labelled_arc(orow, ocol, olab, t[qsrc[e]], t[qdst[e]], qlab[e]);
    keep[i] = ok ? 1 : 0;
    mine += ok;
  }
  if (mine) atomicAdd(kept, mine);
}

template <typename T>
int to_device(const T* h, size_t n, T** d, hipStream_t s) {
  if (hipMalloc(d, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
    (void)hipGetLastError();
    *d = nullptr;
    return GPSM_ENOMEM;
  }
  if (n && hipMemcpyAsync(*d, h, sizeof(T) * n, hipMemcpyHostToDevice, s) != hipSuccess) return GPSM_EDEVICE;
  return GPSM_OK;
}

}  // namespace

extern "C" {

int gpsm_graph_set_directed(gpsm_graph* g, const uint32_t* out_row, const uint32_t* out_col,
                            const uint32_t* out_label) {
  if (!g || !out_row) return GPSM_EINVAL;
  const uint32_t n = g->dg.n;
  const uint64_t A = out_row[n];
  if (out_row[0] != 0 || (A && (!out_col || !out_label)) || A >= (1ull << 32)) return GPSM_EINVAL;
  for (uint32_t v = 0; v < n; ++v) {
    if (out_row[v + 1] < out_row[v]) return GPSM_EINVAL;
    for (uint64_t e = out_row[v]; e < out_row[v + 1]; ++e) {
      if (out_col[e] >= n) return GPSM_EINVAL;
      if (e > out_row[v] && (out_col[e] < out_col[e - 1] ||
                             (out_col[e] == out_col[e - 1] && out_label[e] < out_label[e - 1])))
        return GPSM_EINVAL;  // rows sorted by (target, label)
    }
  }
  hipSetDevice(g->ctx->device);
  hipStream_t s = g->ctx->stream;
  DirectedGraph& d = g->directed;
  d.release();
  int rc;
  if ((rc = to_device(out_row, (size_t)n + 1, &d.row, s)) || (rc = to_device(out_col, (size_t)A, &d.col, s)) ||
      (rc = to_device(out_label, (size_t)A, &d.label, s)) || hipStreamSynchronize(s) != hipSuccess) {
    d.release();
    return rc ? rc : GPSM_EDEVICE;
  }
  d.arcs = A;
  return GPSM_OK;
}

int gpsm_verify_directed(gpsm_ctx* c, const gpsm_graph* g, uint32_t q, uint32_t m, const uint32_t* src,
                         const uint32_t* dst, const uint32_t* elabel, const uint32_t* tuples, uint64_t n_tuples,
                         uint8_t* keep, uint64_t* kept) {
  if (!c || !g || !q || q > GPSM_MAX_QV || (m && (!src || !dst || !elabel)) || (n_tuples && (!tuples || !keep)))
    return GPSM_EINVAL;
  if (!g->directed.row) return GPSM_EINVAL;  // gpsm_graph_set_directed first
  for (uint32_t e = 0; e < m; ++e)
    if (src[e] >= q || dst[e] >= q) return GPSM_EINVAL;
  if (kept) *kept = 0;
  if (!n_tuples) return GPSM_OK;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  uint32_t *qs = nullptr, *qd = nullptr, *ql = nullptr, *dt = nullptr;
  uint8_t* dk = nullptr;
  unsigned long long* dcount = nullptr;
  auto cleanup = [&]() {
    hipFree(qs);
    hipFree(qd);
    hipFree(ql);
    hipFree(dt);
    hipFree(dk);
    hipFree(dcount);
  };
  // tuples go through the device in chunks of at most 2^24
  const uint64_t chunk = std::min<uint64_t>(n_tuples, 1ull << 24);
  int rc;
  if ((rc = to_device(src, m, &qs, s)) || (rc = to_device(dst, m, &qd, s)) || (rc = to_device(elabel, m, &ql, s)) ||
      hipMalloc(&dt, sizeof(uint32_t) * chunk * q) != hipSuccess || hipMalloc(&dk, chunk) != hipSuccess ||
      hipMalloc(&dcount, sizeof(unsigned long long)) != hipSuccess) {
    (void)hipGetLastError();
    cleanup();
    return rc ? rc : GPSM_ENOMEM;
  }
  unsigned long long total = 0;
  for (uint64_t base = 0; base < n_tuples && rc == GPSM_OK; base += chunk) {
    const uint64_t nt = std::min(chunk, n_tuples - base);
    const unsigned long long zero = 0;
    uint64_t grid = (nt + 255) / 256;
    if (grid > 16384) grid = 16384;
    unsigned long long got = 0;
    if (hipMemcpyAsync(dt, tuples + base * q, sizeof(uint32_t) * nt * q, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dcount, &zero, sizeof(zero), hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = GPSM_EDEVICE;
      break;
    }
    hipLaunchKernelGGL(verify_kernel, dim3((uint32_t)grid), dim3(256), 0, s, g->directed.row, g->directed.col,
                       g->directed.label, g->dg.n, qs, qd, ql, m, q, dt, nt, dk, dcount);
    if (hipMemcpyAsync(keep + base, dk, nt, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&got, dcount, sizeof(got), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = GPSM_EDEVICE;
      break;
    }
    total += got;
  }
  cleanup();
  if (rc == GPSM_OK && kept) *kept = total;
  return rc;
}

}  // extern "C"
