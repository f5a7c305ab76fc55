// Scan and compaction on gfx950: the reference's kernel_collect
// (transform_kernel + thrust::exclusive_scan + scatter_kernel,
// match/Match.cu:463-526) and its unbuilt Blelloch prefix scan / stream
// compaction (util/prefix_scan.cu:158-427, util/compact_scan.cu:7-70).
//
// Candidate sets are bitmaps here, so "collect" is a popcount scan.  Each
// bitmap row is cut into segments of SEG_WORDS words, so that a batch of a
// few hundred rows still fills 256 CUs: a first pass popcounts every
// segment (one workgroup per (row, segment)), a row's segment counts are
// scanned in place into row-relative offsets (one workgroup per row; the
// total is the row's candidate count), and the second pass (one workgroup
// per four consecutive segments of a row) takes the group's base from them
// and compacts the four one after the other (each
// thread owns eight consecutive words, one block scan places them), ids
// staged in LDS and written back coalesced.  The same pass writes the rank
// index (set bits before each 256-bit block, row-relative) the join uses to
// find a vertex's candidate-edge list.
#include "launch.h"
#include "wave.h"

namespace gpsm {

namespace {

__global__ void __launch_bounds__(256) seg_popcount_kernel(const uint32_t* __restrict__ bitmaps, uint32_t W,
                                                           uint32_t nseg, uint32_t* __restrict__ segc) {
  __shared__ uint32_t part[4];
  const uint32_t r = blockIdx.x / nseg, sg = blockIdx.x % nseg;
  const uint32_t* bm = bitmaps + (size_t)r * W;
  const uint32_t base = sg * SEG_WORDS;
  uint32_t v[SEG_WORDS / 256];
  if (base + SEG_WORDS <= W) {  // a whole segment: the loads unconditional, all in flight together
#pragma unroll
    for (uint32_t j = 0; j < SEG_WORDS / 256; ++j) v[j] = bm[base + j * 256 + threadIdx.x];
  } else {
#pragma unroll
    for (uint32_t j = 0; j < SEG_WORDS / 256; ++j) {
      const uint32_t i = base + j * 256 + threadIdx.x;
      v[j] = i < W ? bm[i] : 0u;
    }
  }
  uint32_t c = 0;
#pragma unroll
  for (uint32_t j = 0; j < SEG_WORDS / 256; ++j) c += __popc(v[j]);
  c = wave_sum(c);
  if (lane_id() == 0) part[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) segc[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of p[0..n) in place by one workgroup of NT threads, in
// chunks of NT * ITEMS values in striped order (thread t holds p[c + j * NT
// + t]): a wave scan per row j, one wave scan of the (row, wave) totals in
// element order, a running carry across chunks.  Returns the total (in
// every thread).  Contains __syncthreads.
template <typename T, uint32_t NT, uint32_t ITEMS>
__device__ __forceinline__ T block_scan_in_place(T* __restrict__ p, uint32_t n) {
  constexpr uint32_t NW = NT / WAVE;
  static_assert(NT % WAVE == 0 && ITEMS * NW == WAVE, "one wave scans the (row, wave) totals");
  __shared__ T wt[ITEMS * NW + 1];  // (row j, wave w) at j * NW + w, then the chunk total
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE;
  T carry = 0;
  for (uint32_t base = 0; base < n; base += NT * ITEMS) {
    T v[ITEMS];
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) {
      const uint32_t i = base + j * NT + threadIdx.x;
      v[j] = i < n ? p[i] : T(0);
    }
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) {
      T incl;
      if constexpr (sizeof(T) == 8) incl = wave_incl_scan64(v[j]);
      else incl = wave_incl_scan(v[j]);
      if (lane == WAVE - 1) wt[j * NW + w] = incl;
      v[j] = incl - v[j];  // exclusive within the wave's 64 values of row j
    }
    __syncthreads();
    if (w == 0) {
      const T x = wt[lane];
      T incl;
      if constexpr (sizeof(T) == 8) incl = wave_incl_scan64(x);
      else incl = wave_incl_scan(x);
      wt[lane] = incl - x;
      if (lane == WAVE - 1) wt[ITEMS * NW] = incl;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) {
      const uint32_t i = base + j * NT + threadIdx.x;
      if (i < n) p[i] = carry + wt[j * NW + w] + v[j];
    }
    carry += wt[ITEMS * NW];
    __syncthreads();  // the next chunk rewrites wt
  }
  return carry;
}

// Row r's segment popcounts become their exclusive prefix sums (in place),
// and counts[r] their total: one workgroup per row, 4096 segments per step.
// The compaction then reads a group's row-relative base with one load.
// (One thread per row adding its counts took 0.5 ms for the 8192 segments
// of a 2^29-bit row, and each compaction workgroup re-added the counts
// before its group.)
__global__ void __launch_bounds__(256) seg_offsets_kernel(uint32_t* __restrict__ segc, uint32_t nseg,
                                                          uint32_t* __restrict__ counts) {
  const uint32_t t = block_scan_in_place<uint32_t, 256, 16>(segc + (size_t)blockIdx.x * nseg, nseg);
  if (threadIdx.x == 0) counts[blockIdx.x] = t;
}

// A workgroup compacts COMPACT_GROUP consecutive segments of one row, one
// after the other: thread t holds the words j * 256 + t of a segment (eight
// loads in flight, and the next segment's eight issued before the current
// one is scanned); a wave scan per row j of its words' popcounts and one
// scan of the (row, wave) totals give every word its output position, the
// first word of each 256-bit block writes its rank-index entry, and the ids
// go through LDS (in rounds of STAGE_IDS output positions when the segment
// holds more) and are written back coalesced.  The row's set bits before
// the group come from the segment popcounts; within the group the
// workgroup carries them.  Measured per batch on the 1M Chung-Lu mix: 0.29
// ms for one segment per workgroup, eight consecutive words per thread and
// dense segments written by the threads; 0.27 ms for four segments; 0.23 ms
// with dense segments staged in rounds (stages of 4096 and 2048 ids 0.25
// and 0.41 ms; sixteen segments per workgroup 0.31 ms); 0.19 ms with the
// words in striped order.
constexpr uint32_t CW = 8;             // words per thread
constexpr uint32_t STAGE_IDS = 8192;   // ids staged in LDS per round
constexpr uint32_t COMPACT_GROUP = 4;  // segments per workgroup
static_assert(SEG_WORDS == 256 * CW && CW * 4 < WAVE, "a segment is CW words per thread; one wave scans the totals");

// Words of a segment in striped order: thread t holds words j * 256 + t,
// j < CW (every load instruction reads 256 contiguous bytes per wave).
__device__ __forceinline__ void load_words(const uint32_t* __restrict__ bm, uint32_t W, uint32_t s0,
                                           uint32_t (&w)[CW]) {
  if (s0 + SEG_WORDS <= W) {
#pragma unroll
    for (uint32_t j = 0; j < CW; ++j) w[j] = bm[s0 + j * 256 + threadIdx.x];
  } else {
#pragma unroll
    for (uint32_t j = 0; j < CW; ++j) {
      const uint32_t i = s0 + j * 256 + threadIdx.x;
      w[j] = i < W ? bm[i] : 0u;
    }
  }
}

__global__ void __launch_bounds__(256) compact_rows_kernel(const uint32_t* __restrict__ bitmaps, uint32_t W,
                                                           uint32_t NB, uint32_t nseg,
                                                           const uint32_t* __restrict__ segc,
                                                           const uint64_t* __restrict__ coff,
                                                           uint32_t* __restrict__ clist,
                                                           uint32_t* __restrict__ bprefix) {
  __shared__ uint32_t stage[STAGE_IDS];
  __shared__ uint32_t wt[CW * 4 + 1];  // (row j, wave) popcount totals, then their offsets; [CW * 4] = segment total
  __shared__ uint32_t rel;
  const uint32_t lane = lane_id(), wv = threadIdx.x / WAVE;
  const uint32_t ngrp = (nseg + COMPACT_GROUP - 1) / COMPACT_GROUP;
  const uint32_t r = blockIdx.x / ngrp, sg0 = (blockIdx.x % ngrp) * COMPACT_GROUP;
  const uint32_t sg1 = min(nseg, sg0 + COMPACT_GROUP);
  const uint32_t* bm = bitmaps + (size_t)r * W;
  uint32_t* bp = bprefix + (size_t)r * NB;
  uint32_t* out = clist + coff[r];
  if (threadIdx.x == 0) rel = segc[(size_t)r * nseg + sg0];  // set bits of the row's segments before the group
  uint32_t w[CW];
  load_words(bm, W, sg0 * SEG_WORDS, w);
  for (uint32_t sg = sg0; sg < sg1; ++sg) {
    const uint32_t s0 = sg * SEG_WORDS;
    uint32_t nx[CW] = {};
    if (sg + 1 < sg1) load_words(bm, W, s0 + SEG_WORDS, nx);
    // positions: per row j a wave scan of the 64 words' popcounts, then one
    // scan of the (row, wave) totals in word order
    uint32_t pos[CW];
#pragma unroll
    for (uint32_t j = 0; j < CW; ++j) {
      const uint32_t c = __popc(w[j]);
      const uint32_t incl = wave_incl_scan(c);
      if (lane == WAVE - 1) wt[j * 4 + wv] = incl;
      pos[j] = incl - c;
    }
    __syncthreads();  // also publishes rel
    if (wv == 0) {
      const uint32_t x = lane < CW * 4 ? wt[lane] : 0u;
      const uint32_t incl = wave_incl_scan(x);
      if (lane < CW * 4) wt[lane] = incl - x;
      if (lane == CW * 4 - 1) wt[CW * 4] = incl;
    }
    __syncthreads();
    const uint32_t carry = rel, total = wt[CW * 4];
#pragma unroll
    for (uint32_t j = 0; j < CW; ++j) {
      pos[j] += wt[j * 4 + wv];
      const uint32_t wi = s0 + j * 256 + threadIdx.x;
      if ((threadIdx.x & 7) == 0 && wi < W) bp[wi >> 3] = carry + pos[j];  // first word of its 256-bit block
    }
    // ids through LDS, in rounds of STAGE_IDS output positions when the
    // segment holds more, written back coalesced
    for (uint32_t base = 0; base < total; base += STAGE_IDS) {
      const uint32_t lim = base + STAGE_IDS;
#pragma unroll
      for (uint32_t j = 0; j < CW; ++j) {
        const uint32_t pc = __popc(w[j]);
        if (pos[j] + pc > base && pos[j] < lim) {
          const uint32_t id0 = (s0 + j * 256 + threadIdx.x) * 32;
          uint32_t p = pos[j];
          for (uint32_t x = w[j]; x; x &= x - 1, ++p)
            if (p >= base && p < lim) stage[p - base] = id0 + (uint32_t)(__ffs(x) - 1);
        }
      }
      __syncthreads();
      const uint32_t n = min(STAGE_IDS, total - base);
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) out[carry + base + i] = stage[i];
      __syncthreads();
    }
    // every thread has read wt (and stage): the last round ends with a
    // barrier, and a segment without set bits has none
    if (total == 0) __syncthreads();
    if (sg == nseg - 1 && threadIdx.x == 0) bp[NB - 1] = carry + total;
    if (threadIdx.x == 0) rel = carry + total;  // read after the next segment's first barrier
#pragma unroll
    for (uint32_t j = 0; j < CW; ++j) w[j] = nx[j];
  }
}

// Reduce-then-scan for generic u32 arrays (wrapping u32 values, u64 total).
constexpr uint32_t SCAN_ITEMS = 16;
constexpr uint32_t SCAN_TILE = 256 * SCAN_ITEMS;

__global__ void __launch_bounds__(256) scan_reduce_kernel(const uint32_t* __restrict__ in, uint64_t n,
                                                          unsigned long long* __restrict__ sums) {
  __shared__ unsigned long long part[4];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint32_t v[SCAN_ITEMS];
  if (base + SCAN_TILE <= n) {  // a whole tile: every load unconditional, all in flight together
#pragma unroll
    for (uint32_t j = 0; j < SCAN_ITEMS; ++j) v[j] = in[base + (uint64_t)j * 256 + threadIdx.x];
  } else {
#pragma unroll
    for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
      v[j] = i < n ? in[i] : 0u;
    }
  }
  unsigned long long s = 0;
#pragma unroll
  for (uint32_t j = 0; j < SCAN_ITEMS; ++j) s += v[j];
  s = wave_sum64(s);
  if (lane_id() == 0) part[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// The tile sums' exclusive scan, in place, and their total: one workgroup
// of 1024 threads, 4096 sums per step.  (A single thread's serial loop over
// them took 3.8 ms for 32,768 sums; 256 threads with 16 sums each, 45 us.)
__global__ void __launch_bounds__(1024) scan_sums_kernel(unsigned long long* __restrict__ sums, uint32_t nb,
                                                         unsigned long long* __restrict__ total) {
  const unsigned long long t = block_scan_in_place<unsigned long long, 1024, 4>(sums, nb);
  if (threadIdx.x == 0 && total) *total = t;
}

// A tile in striped order: thread t holds values j * 256 + t (every load
// and store instruction moves 256 contiguous bytes per wave).  Each wave
// scans its 64 values of every row j with shuffles, one wave scans the 64
// (row, wave) totals in element order, and each value gets the tile's base,
// its (row, wave) offset and its place in its wave's scan: two barriers per
// tile.
__global__ void __launch_bounds__(256) scan_tile_kernel(const uint32_t* in, uint32_t* out, uint64_t n,
                                                        const unsigned long long* __restrict__ sums) {
  __shared__ uint32_t wt[SCAN_ITEMS * 4];  // (row j, wave w) totals at j * 4 + w, then their offsets
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  const bool whole = base + SCAN_TILE <= n;
  uint32_t v[SCAN_ITEMS];
  if (whole) {
#pragma unroll
    for (uint32_t j = 0; j < SCAN_ITEMS; ++j) v[j] = in[base + (uint64_t)j * 256 + threadIdx.x];
  } else {
#pragma unroll
    for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
      v[j] = i < n ? in[i] : 0u;
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
    const uint32_t incl = wave_incl_scan(v[j]);
    if (lane == WAVE - 1) wt[j * 4 + w] = incl;
    v[j] = incl - v[j];  // exclusive within the wave's 64 values of row j
  }
  __syncthreads();
  static_assert(SCAN_ITEMS * 4 == WAVE, "one wave scans the (row, wave) totals");
  if (w == 0) {
    const uint32_t x = wt[lane];
    wt[lane] = wave_incl_scan(x) - x;
  }
  __syncthreads();
  const uint32_t off = (uint32_t)sums[blockIdx.x];
  if (whole) {
#pragma unroll
    for (uint32_t j = 0; j < SCAN_ITEMS; ++j) out[base + (uint64_t)j * 256 + threadIdx.x] = off + wt[j * 4 + w] + v[j];
  } else {
#pragma unroll
    for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
      if (i < n) out[i] = off + wt[j * 4 + w] + v[j];
    }
  }
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

}  // namespace

// Fill kernel instead of hipMemsetAsync so that captured graphs hold only
// kernel nodes.
void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  uint64_t grid = (n + 255) / 256;
  if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)grid), dim3(256), 0, s, p, v, n);
}

void launch_row_popcount(const DevGraph& g, const uint32_t* bitmaps, uint32_t nrows, uint32_t* segc,
                         uint32_t* counts, hipStream_t s) {
  if (!nrows) return;
  const uint32_t nseg = segs_of(g.W);
  hipLaunchKernelGGL(seg_popcount_kernel, dim3(nrows * nseg), dim3(256), 0, s, bitmaps, g.W, nseg, segc);
  hipLaunchKernelGGL(seg_offsets_kernel, dim3(nrows), dim3(256), 0, s, segc, nseg, counts);
}

void launch_compact_rows(const DevGraph& g, const uint32_t* bitmaps, uint32_t nrows, const uint32_t* segc,
                         const uint64_t* coff, uint32_t* clist, uint32_t* bprefix, hipStream_t s) {
  if (!nrows) return;
  const uint32_t nseg = segs_of(g.W);
  const uint32_t ngrp = (nseg + COMPACT_GROUP - 1) / COMPACT_GROUP;
  hipLaunchKernelGGL(compact_rows_kernel, dim3(nrows * ngrp), dim3(256), 0, s, bitmaps, g.W, g.NB, nseg, segc, coff,
                     clist, bprefix);
}

uint64_t scan_scratch_words(uint64_t n) { return 2 * ((n + SCAN_TILE - 1) / SCAN_TILE) + 2; }

void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, unsigned long long* total,
                           uint32_t* scratch, uint64_t scratch_words, hipStream_t s) {
  const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 0 || scratch_words < 2 * nb) return;
  unsigned long long* sums = reinterpret_cast<unsigned long long*>(scratch);
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((uint32_t)nb), dim3(256), 0, s, in, n, sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, s, sums, (uint32_t)nb, total);
  hipLaunchKernelGGL(scan_tile_kernel, dim3((uint32_t)nb), dim3(256), 0, s, in, out, n, sums);
}

uint64_t compact_scratch_words(uint64_t nbits) {
  const uint64_t W = (nbits + 31) / 32;
  return 2 + ((W + 7) / 8 + 1) + segs_of((uint32_t)W) + 1;
}

void launch_compact_bitmap(const uint32_t* bits, uint64_t nbits, uint32_t* ids, unsigned long long* count,
                           uint32_t* scratch, hipStream_t s) {
  // one row of ceil(nbits/32) words through the row compaction; scratch
  // holds the row offset (2 words), the rank index, the segment counts and
  // the row count
  DevGraph g;
  g.W = (uint32_t)((nbits + 31) / 32);
  g.NB = (g.W + 7) / 8 + 1;
  uint64_t* coff = reinterpret_cast<uint64_t*>(scratch);
  uint32_t* bp = scratch + 2;
  uint32_t* segc = bp + g.NB;
  uint32_t* rowc = segc + segs_of(g.W);
  hipMemsetAsync(coff, 0, sizeof(uint64_t), s);
  launch_row_popcount(g, bits, 1, segc, rowc, s);
  launch_compact_rows(g, bits, 1, segc, coff, ids, bp, s);
  hipMemsetAsync(count, 0, sizeof(unsigned long long), s);
  hipMemcpyAsync(count, bp + g.NB - 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
}

}  // namespace gpsm
