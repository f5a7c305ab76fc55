// Undirected CSR built on the device from an edge list (the reference
// builds it on the host from per-vertex in/out vectors and sorts each row,
// graph/Graph.cpp:30-104; its own note at graph/Graph.cpp:26-28 asks for a
// GPU build).  Every edge {a, b} becomes the two keys a<<32|b and b<<32|a;
// one radix sort of the 64-bit keys orders them by (row, column), one unique
// pass drops duplicates, and the row offsets are lower bounds of v<<32.
// Self-loops become the key n<<32, which sorts after every real row and is
// left out of the offsets.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "gpsm.h"
#include "launch.h"

namespace gpsm {

namespace {

__global__ void edge_keys_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t m,
                                 uint32_t n, unsigned long long* __restrict__ keys, uint32_t* __restrict__ bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const uint32_t a = src[i], b = dst[i];
    unsigned long long k0 = (unsigned long long)n << 32, k1 = k0;
    if (a >= n || b >= n) {
      atomicOr(bad, 1u);
    } else if (a != b) {
      k0 = ((unsigned long long)a << 32) | b;
      k1 = ((unsigned long long)b << 32) | a;
    }
    keys[2 * i] = k0;
    keys[2 * i + 1] = k1;
  }
}

// row[v] = first key position with row >= v, for v = 0..n (a lower bound).
__global__ void row_offsets_kernel(const unsigned long long* __restrict__ keys, uint64_t count, uint32_t n,
                                   uint32_t* __restrict__ row) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += stride) {
    const unsigned long long target = v << 32;
    uint64_t lo = 0, hi = count;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (keys[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    row[v] = (uint32_t)lo;
  }
}

__global__ void columns_kernel(const unsigned long long* __restrict__ keys, uint64_t adj, uint32_t* __restrict__ col) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < adj; i += stride)
    col[i] = (uint32_t)keys[i];
}

uint32_t grid_for(uint64_t items) {
  const uint64_t g = (items + 255) / 256;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 16384));
}

}  // namespace

int build_csr_from_edges(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, hipStream_t s,
                         uint32_t** row, uint32_t** col, uint64_t* adj_out) {
  *row = *col = nullptr;
  *adj_out = 0;
  const uint64_t M2 = 2 * m;
  uint32_t *d_src = nullptr, *d_dst = nullptr, *d_bad = nullptr;
  unsigned long long *keys = nullptr, *sorted = nullptr, *uniq = nullptr, *d_count = nullptr;
  void* temp = nullptr;
  int rc = GPSM_OK;
  auto cleanup = [&]() {
    hipFree(d_src);
    hipFree(d_dst);
    hipFree(d_bad);
    hipFree(keys);
    hipFree(sorted);
    hipFree(uniq);
    hipFree(d_count);
    hipFree(temp);
  };
  // bits that can differ: the column (32) and the row, up to the sentinel row n
  unsigned rbits = 1;
  while (rbits < 32 && ((uint64_t)n >> rbits)) ++rbits;
  const unsigned end_bit = 32 + rbits;
  size_t sort_bytes = 0, uniq_bytes = 0;
  if (M2) {
    rocprim::radix_sort_keys(nullptr, sort_bytes, keys, sorted, (size_t)M2, 0, end_bit, s);
    rocprim::unique(nullptr, uniq_bytes, sorted, uniq, d_count, (size_t)M2, rocprim::equal_to<unsigned long long>(), s);
  }
  const size_t temp_bytes = std::max<size_t>(std::max(sort_bytes, uniq_bytes), 16);
  if (hipMalloc(&d_bad, sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&d_count, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(row, sizeof(uint32_t) * ((size_t)n + 1)) != hipSuccess ||
      (M2 && (hipMalloc(&d_src, sizeof(uint32_t) * m) != hipSuccess ||
              hipMalloc(&d_dst, sizeof(uint32_t) * m) != hipSuccess ||
              hipMalloc(&keys, sizeof(unsigned long long) * M2) != hipSuccess ||
              hipMalloc(&sorted, sizeof(unsigned long long) * M2) != hipSuccess ||
              hipMalloc(&uniq, sizeof(unsigned long long) * M2) != hipSuccess ||
              hipMalloc(&temp, temp_bytes) != hipSuccess))) {
    (void)hipGetLastError();
    cleanup();
    return GPSM_ENOMEM;
  }
  uint32_t bad = 0;
  unsigned long long count = 0;
  if (hipMemsetAsync(d_bad, 0, sizeof(uint32_t), s) != hipSuccess ||
      hipMemsetAsync(d_count, 0, sizeof(unsigned long long), s) != hipSuccess)
    rc = GPSM_EDEVICE;
  if (!rc && M2) {
    if (hipMemcpyAsync(d_src, src, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_dst, dst, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = GPSM_EDEVICE;
    } else {
      hipLaunchKernelGGL(edge_keys_kernel, dim3(grid_for(m)), dim3(256), 0, s, d_src, d_dst, m, n, keys, d_bad);
      size_t b1 = temp_bytes, b2 = temp_bytes;
      if (rocprim::radix_sort_keys(temp, b1, keys, sorted, (size_t)M2, 0, end_bit, s) != hipSuccess ||
          rocprim::unique(temp, b2, sorted, uniq, d_count, (size_t)M2, rocprim::equal_to<unsigned long long>(), s) !=
              hipSuccess)
        rc = GPSM_EDEVICE;
    }
  }
  if (!rc && (hipMemcpyAsync(&bad, d_bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipMemcpyAsync(&count, d_count, sizeof(count), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess))
    rc = GPSM_EDEVICE;
  if (!rc && bad) rc = GPSM_EINVAL;  // an endpoint >= n
  if (!rc) {
    hipLaunchKernelGGL(row_offsets_kernel, dim3(grid_for((uint64_t)n + 1)), dim3(256), 0, s, uniq, count, n, *row);
    uint32_t adj32 = 0;
    if (hipMemcpyAsync(&adj32, *row + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = GPSM_EDEVICE;
    } else {
      *adj_out = adj32;
      if (hipMalloc(col, sizeof(uint32_t) * std::max<uint64_t>(adj32, 1)) != hipSuccess) {
        (void)hipGetLastError();
        rc = GPSM_ENOMEM;
      } else {
        if (adj32)
          hipLaunchKernelGGL(columns_kernel, dim3(grid_for(adj32)), dim3(256), 0, s, uniq, (uint64_t)adj32, *col);
        if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) rc = GPSM_EDEVICE;
      }
    }
  }
  cleanup();
  if (rc) {
    hipFree(*row);
    hipFree(*col);
    *row = *col = nullptr;
  }
  return rc;
}

}  // namespace gpsm
