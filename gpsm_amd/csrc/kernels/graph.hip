// Graph-resident structures built once at upload.
//
// Neighbour keys: ckey[e] = key[col[e]], the label/degree key of every
// adjacency entry's target, stored beside col (4 bytes per entry).
//
// Hub adjacency bitmaps: one n-bit bitmap per hub (a vertex above the hub
// degree), so that the join's edge-membership test of a pair touching a hub
// is one bit instead of a binary search in an adjacency list of up to 10^5
// entries (the R-MAT graph's largest), which costs ~17 dependent loads.
#include "launch.h"
#include "wave.h"

namespace gpsm {

namespace {

// One wave per (hub, 64-entry block of its adjacency): grid-stride over the
// blocks of all hubs, found with a binary search over their prefix sums.
__global__ void __launch_bounds__(256) hub_bits_kernel(const uint32_t* __restrict__ row,
                                                       const uint32_t* __restrict__ col,
                                                       const uint32_t* __restrict__ hubs,
                                                       const uint32_t* __restrict__ block_base, uint32_t nhub,
                                                       uint32_t W, uint32_t* __restrict__ bits) {
  const uint32_t lane = lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  const uint32_t nblocks = block_base[nhub];
  for (uint64_t it = wave; it < nblocks; it += nwaves) {
    const uint32_t h = uniform(seg_search(block_base, nhub + 1, (uint32_t)it));
    const uint32_t v = hubs[h];
    const uint32_t e = row[v] + ((uint32_t)it - block_base[h]) * WAVE + lane;
    if (e < row[v + 1]) {
      const uint32_t w = col[e];
      atomicOr(bits + (size_t)h * W + (w >> 5), 1u << (w & 31));
    }
  }
}

__global__ void neighbour_keys_kernel(const uint32_t* __restrict__ col, const uint32_t* __restrict__ key,
                                     uint64_t adj, uint32_t* __restrict__ ckey) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < adj; e += stride) ckey[e] = key[col[e]];
}

}  // namespace

void launch_neighbour_keys(const DevGraph& g, hipStream_t s) {
  if (!g.adj) return;
  uint64_t grid = (g.adj + 255) / 256;
  if (grid > 16384) grid = 16384;
  hipLaunchKernelGGL(neighbour_keys_kernel, dim3((uint32_t)grid), dim3(256), 0, s, g.col, g.key, g.adj, g.ckey);
}

void launch_hub_bits(const DevGraph& g, const uint32_t* hubs, const uint32_t* block_base, uint32_t nblocks,
                     hipStream_t s) {
  if (!g.nhub || !nblocks) return;
  const uint32_t block = 256;
  uint64_t grid = ((uint64_t)nblocks * WAVE + block - 1) / block;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(hub_bits_kernel, dim3((uint32_t)grid), dim3(block), 0, s, g.row, g.col, hubs, block_base,
                     g.nhub, g.W, g.hub_bits);
}

}  // namespace gpsm
