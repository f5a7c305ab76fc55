// Wave64 primitives for gfx950 (CDNA4): ballots, lane scans and the
// wave-local LDS barrier.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gpsm {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Makes this wave's LDS writes visible to its other lanes (no workgroup
// barrier: waves of a block run independent loops).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t y = __shfl_up(x, d, WAVE);
    if (l >= (uint32_t)d) x += y;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_incl_scan64(unsigned long long x) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    unsigned long long y = __shfl_up(x, d, WAVE);
    if (l >= (uint32_t)d) x += y;
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, WAVE);
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, WAVE);
  return x;
}

__device__ __forceinline__ uint32_t bit_of(const uint32_t* bm, uint32_t v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

// Sets bit v, reading first so that already-set bits cost no atomic.
__device__ __forceinline__ void set_bit(uint32_t* bm, uint32_t v) {
  uint32_t m = 1u << (v & 31);
  if (!(__hip_atomic_load(bm + (v >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m)) atomicOr(bm + (v >> 5), m);
}

__device__ __forceinline__ void clear_bit(uint32_t* bm, uint32_t v) { atomicAnd(bm + (v >> 5), ~(1u << (v & 31))); }

// Index of the last entry of pre[0..n) (non-decreasing) that is <= g.
__device__ __forceinline__ uint32_t seg_search(const uint32_t* pre, uint32_t n, uint32_t g) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// For K keys at once (key[k] in, index out): the largest i in [0, n) with
// pre[i] <= key[k] (pre non-decreasing, pre[0] <= every key).  ceil(log2 n)
// halvings with no early exit: n must be the same in every lane, so every
// lane takes the same steps, and each step's K LDS reads are in flight
// together (a per-key loop waits for one read at a time).
template <int K>
__device__ __forceinline__ void halving_search(const uint32_t* __restrict__ pre, uint32_t n, uint32_t* key) {
  uint32_t b[K];
#pragma unroll
  for (int k = 0; k < K; ++k) b[k] = 0;
  while (n > 1) {
    const uint32_t half = n >> 1;
#pragma unroll
    for (int k = 0; k < K; ++k) b[k] = pre[b[k] + half] <= key[k] ? b[k] + half : b[k];
    n -= half;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) key[k] = b[k];
}

__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// One factor of a tuple's hash (odd): the tuple hash is the product of the
// factors of its (query vertex, image) pairs mod 2^64, so it does not depend
// on the order in which the join binds the vertices.
__device__ __forceinline__ uint64_t tuple_term(uint32_t qid, uint32_t v) {
  return fmix64((((uint64_t)qid << 32) | v) ^ 0x9E3779B97F4A7C15ULL) | 1ULL;
}

}  // namespace gpsm

namespace gpsm {

// Exclusive scan over a workgroup of up to 1024 threads (multiple of 64);
// tmp holds one u32 per wave plus one.  Returns the exclusive prefix and
// the group total in *total.  Contains __syncthreads.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* tmp, uint32_t* total) {
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const uint32_t incl = wave_incl_scan(x);
  if (lane == WAVE - 1) tmp[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t i = 0; i < nw; ++i) {
      uint32_t t = tmp[i];
      tmp[i] = acc;
      acc += t;
    }
    tmp[nw] = acc;
  }
  __syncthreads();
  const uint32_t r = tmp[w] + incl - x;
  *total = tmp[nw];
  __syncthreads();
  return r;
}

}  // namespace gpsm

namespace gpsm {

__device__ __forceinline__ float wave_sum_f(float x) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, WAVE);
  return x;
}

}  // namespace gpsm
