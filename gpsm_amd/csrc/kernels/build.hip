// Undirected CSR built on the device from an edge list (the reference
// builds it on the host from per-vertex in/out vectors and sorts each row,
// graph/Graph.cpp:30-104; its own note at graph/Graph.cpp:26-28 asks for a
// GPU build).  Every edge {a, b} becomes the two keys a<<32|b and b<<32|a;
// an LSD radix sort of the 64-bit keys (8-bit digits, only the bits that
// can differ) orders them by (row, column), a scan of "first of its run"
// flags drops duplicates, and the row offsets are lower bounds of v<<32.
// Self-loops become the key n<<32, which sorts after every real row and is
// left out of the offsets.
//
// One radix pass: each workgroup counts the digits of its tile of
// RADIX_TILE keys (LDS histogram), an exclusive scan of the counts in
// digit-major order gives every (digit, tile) its output start, and each
// workgroup scatters its tile stably: 256 keys at a time, in index order,
// a key's place among the equal digits of its wave comes from eight
// ballots (one per digit bit), the waves before it from per-wave digit
// counts in LDS, the chunks before it from a running count per digit.
#include <algorithm>

#include "gpsm.h"
#include "launch.h"
#include "wave.h"

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

constexpr uint32_t RADIX_ITEMS = 16;                  // keys per thread and tile
constexpr uint32_t RADIX_TILE = 256 * RADIX_ITEMS;     // keys per workgroup

__global__ void __launch_bounds__(256) radix_hist_kernel(const unsigned long long* __restrict__ keys, uint64_t n,
                                                         uint32_t shift, uint32_t ntiles,
                                                         uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RADIX_TILE;
  for (uint32_t j = 0; j < RADIX_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];  // digit-major
}

__global__ void __launch_bounds__(256) radix_scatter_kernel(const unsigned long long* __restrict__ in,
                                                            unsigned long long* __restrict__ out, uint64_t n,
                                                            uint32_t shift, uint32_t ntiles,
                                                            const uint32_t* __restrict__ start) {
  __shared__ uint32_t run[256];     // output position of the next key of each digit
  __shared__ uint32_t wc[4][256];   // this chunk's keys per (wave, digit)
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE;
  const unsigned long long below = lane ? (~0ULL >> (WAVE - lane)) : 0ULL;
  run[threadIdx.x] = start[(size_t)threadIdx.x * ntiles + blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * RADIX_TILE;
  for (uint32_t j = 0; j < RADIX_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
    const bool valid = i < n;
    const unsigned long long key = valid ? in[i] : 0ULL;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    // the lanes of this wave holding the same digit
    unsigned long long same = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
      const unsigned long long ones = __ballot(valid && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? ones : ~ones;
    }
    const uint32_t rank = (uint32_t)__popcll(same & below);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) wc[w][lane + 64 * k] = 0;
    if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(same);
    __syncthreads();  // also publishes run (first chunk) and the previous chunk's counts
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (uint32_t x = 0; x < w; ++x) pos += wc[x][d];
      out[pos] = key;
    }
    __syncthreads();
    run[threadIdx.x] += wc[0][threadIdx.x] + wc[1][threadIdx.x] + wc[2][threadIdx.x] + wc[3][threadIdx.x];
    __syncthreads();
  }
}

// flags[i] = 1 where key i starts a run of equal keys (the keys sorted)
__global__ void run_heads_kernel(const unsigned long long* __restrict__ keys, uint64_t n,
                                 uint32_t* __restrict__ flags) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void keep_heads_kernel(const unsigned long long* __restrict__ keys, uint64_t n,
                                  const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos,
                                  unsigned long long* __restrict__ uniq) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (flags[i]) uniq[pos[i]] = keys[i];
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
  if (M2 > 0xFFFFFFFFull) return GPSM_EINVAL;  // positions are 32-bit
  uint32_t *d_src = nullptr, *d_dst = nullptr, *d_bad = nullptr, *hist = nullptr, *hstart = nullptr;
  uint32_t *flags = nullptr, *fpos = nullptr, *scratch = nullptr;
  unsigned long long *keys = nullptr, *sorted = nullptr, *d_count = nullptr;
  int rc = GPSM_OK;
  auto cleanup = [&]() {
    hipFree(d_src);
    hipFree(d_dst);
    hipFree(d_bad);
    hipFree(hist);
    hipFree(hstart);
    hipFree(flags);
    hipFree(fpos);
    hipFree(scratch);
    hipFree(keys);
    hipFree(sorted);
    hipFree(d_count);
  };
  // bits that can differ: the column (32) and the row, up to the sentinel row n
  unsigned rbits = 1;
  while (rbits < 32 && ((uint64_t)n >> rbits)) ++rbits;
  const unsigned end_bit = 32 + rbits;
  const uint32_t ntiles = (uint32_t)std::max<uint64_t>(1, (M2 + RADIX_TILE - 1) / RADIX_TILE);
  const uint64_t nhist = 256ull * ntiles;
  const uint64_t sw = scan_scratch_words(std::max<uint64_t>(nhist, M2)) + 2;
  if (hipMalloc(&d_bad, sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&d_count, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(row, sizeof(uint32_t) * ((size_t)n + 1)) != hipSuccess ||
      (M2 && (hipMalloc(&d_src, sizeof(uint32_t) * m) != hipSuccess ||
              hipMalloc(&d_dst, sizeof(uint32_t) * m) != hipSuccess ||
              hipMalloc(&keys, sizeof(unsigned long long) * M2) != hipSuccess ||
              hipMalloc(&sorted, sizeof(unsigned long long) * M2) != hipSuccess ||
              hipMalloc(&hist, sizeof(uint32_t) * nhist) != hipSuccess ||
              hipMalloc(&hstart, sizeof(uint32_t) * nhist) != hipSuccess ||
              hipMalloc(&flags, sizeof(uint32_t) * M2) != hipSuccess ||
              hipMalloc(&fpos, sizeof(uint32_t) * M2) != hipSuccess ||
              hipMalloc(&scratch, sizeof(uint32_t) * sw) != hipSuccess))) {
    (void)hipGetLastError();
    cleanup();
    return GPSM_ENOMEM;
  }
  uint32_t bad = 0;
  unsigned long long count = 0;
  if (hipMemsetAsync(d_bad, 0, sizeof(uint32_t), s) != hipSuccess ||
      hipMemsetAsync(d_count, 0, sizeof(unsigned long long), s) != hipSuccess)
    rc = GPSM_EDEVICE;
  // the sorted, deduplicated keys end up in `keys` (the passes ping-pong)
  unsigned long long* uniq = keys;
  if (!rc && M2) {
    if (hipMemcpyAsync(d_src, src, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_dst, dst, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = GPSM_EDEVICE;
    } else {
      hipLaunchKernelGGL(edge_keys_kernel, dim3(grid_for(m)), dim3(256), 0, s, d_src, d_dst, m, n, keys, d_bad);
      unsigned long long *a = keys, *b = sorted;
      for (uint32_t shift = 0; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(radix_hist_kernel, dim3(ntiles), dim3(256), 0, s, a, M2, shift, ntiles, hist);
        launch_exclusive_scan(hist, hstart, nhist, d_count, scratch + 2, sw - 2, s);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(ntiles), dim3(256), 0, s, a, b, M2, shift, ntiles, hstart);
        std::swap(a, b);
      }
      // a holds the sorted keys; the first key of each run of equal keys is kept, in order, in b
      hipLaunchKernelGGL(run_heads_kernel, dim3(grid_for(M2)), dim3(256), 0, s, a, M2, flags);
      launch_exclusive_scan(flags, fpos, M2, d_count, scratch + 2, sw - 2, s);
      hipLaunchKernelGGL(keep_heads_kernel, dim3(grid_for(M2)), dim3(256), 0, s, a, M2, flags, fpos, b);
      uniq = b;
      if (hipGetLastError() != hipSuccess) rc = GPSM_EDEVICE;
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
