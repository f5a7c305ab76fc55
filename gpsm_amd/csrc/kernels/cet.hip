// Candidate-edge tables on gfx950 (count_kernel + thrust::exclusive_scan +
// examine_kernel, match/Match.cu:892-1139).  CET(x -> w) lists, for each
// mu in C(x) ascending, the data neighbours of mu that are in C(w), in
// adjacency (ascending) order.
//
// One wave per 64 candidates of C(x), taken with a stride: of the N waves
// of a table, wave j takes candidates j, j + N, j + 2N, ..., so the hubs of
// a power-law graph (its lowest ids, the first candidates) land on
// different waves.  The wave walks the concatenated adjacency of its
// candidates 64 entries per step (balanced across hub and leaf candidates)
// and tests C(w) membership in the bitmap, probing it only for entries
// whose neighbour key (ckey, read beside col) is in w's label/degree range:
// C(w) holds no vertex outside it.  Outputs are per candidate (its
// count, its list at its offset), so the assignment does not change them.
// The write pass ranks each hit inside its candidate's run with a ballot,
// so lists stay in adjacency order without a second scan.
#include "launch.h"
#include "wave.h"

namespace gpsm {

namespace {

constexpr int CET_UNROLL = 4;  // adjacency entries in flight per lane

struct CetLds {
  uint32_t pre[WAVE + 1];
  uint32_t off[WAVE];
  uint32_t cnt[WAVE];
};

// Optional work counters (profiling): per lane, candidates loaded, adjacency
// entries examined, bitmap probes and list entries written; one atomic per
// counter per workgroup into copy blockIdx % STAT_COPIES of stats.
struct CetTally {
  uint32_t cand = 0, ent = 0, probe = 0, wrote = 0;
};

__device__ __forceinline__ void flush_tally(const CetTally& t, unsigned long long* stats) {
  __shared__ unsigned long long red[4][4];
  const unsigned long long v[4] = {wave_sum64(t.cand), wave_sum64(t.ent), wave_sum64(t.probe), wave_sum64(t.wrote)};
  const uint32_t w = threadIdx.x / WAVE;
  if (lane_id() == 0)
    for (int i = 0; i < 4; ++i) red[w][i] = v[i];
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long tot = 0;
    for (uint32_t x = 0; x < blockDim.x / WAVE; ++x) tot += red[x][threadIdx.x];
    if (tot) atomicAdd(stats + (blockIdx.x % STAT_COPIES) * STAT_STRIDE + threadIdx.x, tot);
  }
}

// Lane l of chunk j of a table with nch chunks holds candidate j + l * nch.
__device__ __forceinline__ uint32_t cand_index(uint32_t j, uint32_t nch, uint32_t lane) { return j + lane * nch; }

__device__ __forceinline__ void load_chunk(const uint32_t* __restrict__ row, const uint32_t* __restrict__ clist,
                                           const CetDesc& d, uint32_t j, uint32_t nch, CetLds& L, bool& mine,
                                           uint32_t& D) {
  const uint32_t lane = lane_id();
  const uint32_t ci = cand_index(j, nch, lane);
  mine = ci < d.ncand;
  const uint32_t id = mine ? clist[d.clist_off + ci] : 0u;
  const uint32_t off = mine ? row[id] : 0u;
  const uint32_t deg = mine ? row[id + 1] - off : 0u;
  const uint32_t incl = wave_incl_scan(deg);
  D = uniform(__shfl(incl, WAVE - 1, WAVE));
  L.pre[lane] = incl - deg;
  L.off[lane] = off;
  L.cnt[lane] = 0;
  wave_sync();
}

// Eight waves per SIMD (64 VGPRs, no spill; seven at the compiler's own
// choice): 9-11% faster on the 1M Chung-Lu mix and the R-MAT 22 shapes.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) cet_count_kernel(const uint32_t* __restrict__ row,
                                                        const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ ckey,
                                                        const uint32_t* __restrict__ bitmaps, uint32_t W,
                                                        const uint32_t* __restrict__ clist, CetDesc* descs,
                                                        const uint32_t* __restrict__ chunk_base, uint32_t ndesc,
                                                        uint32_t* __restrict__ offsets, int per_cand,
                                                        unsigned long long* __restrict__ stats) {
  __shared__ CetLds lds[4];
  CetTally tl;
  CetLds& L = lds[threadIdx.x / WAVE];
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / WAVE;
  const uint32_t nitems = chunk_base[ndesc];
  for (uint32_t it = wave; it < nitems; it += nwaves) {
    const uint32_t di = uniform(seg_search(chunk_base, ndesc + 1, it));
    const CetDesc d = descs[di];
    const uint32_t j = it - chunk_base[di], nch = chunk_base[di + 1] - chunk_base[di];
    bool mine;
    uint32_t D;
    load_chunk(row, clist, d, j, nch, L, mine, D);
    tl.cand += mine;
    const uint32_t* bm_w = bitmaps + (size_t)d.w_slot * W;
    const uint32_t wlo = d.w_hi ? d.w_lo : 0u, whi = d.w_hi ? d.w_hi : ~0u;
    uint32_t tot = 0;
    for (uint32_t f = 0; f < D; f += WAVE * CET_UNROLL) {
      // CET_UNROLL entries per lane: all adjacency and key loads, then the
      // probes of the entries whose key is in range
      uint32_t cis[CET_UNROLL], mvs[CET_UNROLL], ks[CET_UNROLL], es[CET_UNROLL];
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) cis[u] = f + u * WAVE + lane;
      halving_search<CET_UNROLL>(L.pre, WAVE, cis);
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) es[u] = L.off[cis[u]] + (f + u * WAVE + lane - L.pre[cis[u]]);
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) {
        if (f + u * WAVE + lane < D) {
          mvs[u] = col[es[u]];
          ks[u] = ckey[es[u]];
        } else {
          cis[u] = WAVE;
        }
      }
      // the probes' bitmap words loaded together, then tested
      uint32_t pw[CET_UNROLL];
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u)
        pw[u] = cis[u] != WAVE && ks[u] >= wlo && ks[u] <= whi ? bm_w[mvs[u] >> 5] >> (mvs[u] & 31) : 0u;
      if (stats) {
#pragma unroll
        for (int u = 0; u < CET_UNROLL; ++u) {
          tl.ent += cis[u] != WAVE;
          tl.probe += cis[u] != WAVE && ks[u] >= wlo && ks[u] <= whi;
        }
      }
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u)
        if (pw[u] & 1u) {
          ++tot;
          if (per_cand) atomicAdd(&L.cnt[cis[u]], 1u);
        }
    }
    wave_sync();
    if (per_cand && mine) offsets[d.off_base + cand_index(j, nch, lane)] = L.cnt[lane];
    tot = wave_sum(tot);
    if (lane == 0 && tot) atomicAdd(reinterpret_cast<unsigned long long*>(&descs[di].total), (unsigned long long)tot);
    wave_sync();
  }
  if (stats) flush_tally(tl, stats);
}

// One workgroup per table scans its per-candidate counts in place, in tiles
// of 256 x OFF_ITEMS: each thread sums OFF_ITEMS consecutive counts (loads
// in flight together), one block scan places the thread sums, and each
// thread writes its run.  A table of 10^5 candidates takes 49 tiles instead
// of 391 block scans of 256.
constexpr uint32_t OFF_ITEMS = 8;

__global__ void __launch_bounds__(256) cet_offsets_kernel(const CetDesc* __restrict__ descs,
                                                          uint32_t* __restrict__ offsets) {
  __shared__ uint32_t tmp[8];
  const CetDesc d = descs[blockIdx.x];
  uint32_t* o = offsets + d.off_base;
  uint32_t carry = 0;
  for (uint32_t t0 = 0; t0 < d.ncand; t0 += blockDim.x * OFF_ITEMS) {
    const uint32_t i0 = t0 + threadIdx.x * OFF_ITEMS;
    uint32_t v[OFF_ITEMS], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < OFF_ITEMS; ++j) {
      v[j] = i0 + j < d.ncand ? o[i0 + j] : 0u;
      sum += v[j];
    }
    uint32_t total;
    uint32_t run = carry + block_excl_scan(sum, tmp, &total);
#pragma unroll
    for (uint32_t j = 0; j < OFF_ITEMS; ++j) {
      if (i0 + j < d.ncand) o[i0 + j] = min(run, d.list_cap);
      run += v[j];
    }
    carry += total;
  }
  if (threadIdx.x == 0) o[d.ncand] = min(carry, d.list_cap);
}

__global__ void __launch_bounds__(256) cet_write_kernel(const uint32_t* __restrict__ row,
                                                        const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ ckey,
                                                        const uint32_t* __restrict__ bitmaps, uint32_t W,
                                                        const uint32_t* __restrict__ clist,
                                                        const CetDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ chunk_base, uint32_t ndesc,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ lists,
                                                        unsigned long long* __restrict__ stats) {
  __shared__ CetLds lds[4];
  CetTally tl;
  CetLds& L = lds[threadIdx.x / WAVE];
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / WAVE;
  const uint32_t nitems = chunk_base[ndesc];
  for (uint32_t it = wave; it < nitems; it += nwaves) {
    const uint32_t di = uniform(seg_search(chunk_base, ndesc + 1, it));
    const CetDesc d = descs[di];
    const uint32_t j = it - chunk_base[di], nch = chunk_base[di + 1] - chunk_base[di];
    bool mine;
    uint32_t D;
    load_chunk(row, clist, d, j, nch, L, mine, D);
    tl.cand += mine;
    // each lane's list starts at its candidate's offset; cnt advances it
    L.cnt[lane_id()] = mine ? offsets[d.off_base + cand_index(j, nch, lane_id())] : 0u;
    wave_sync();
    const uint32_t* bm_w = bitmaps + (size_t)d.w_slot * W;
    const uint32_t wlo = d.w_hi ? d.w_lo : 0u, whi = d.w_hi ? d.w_hi : ~0u;
    uint32_t* out = lists + d.list_base;
    for (uint32_t f = 0; f < D; f += WAVE * CET_UNROLL) {
      // CET_UNROLL sub-steps of 64 entries: all adjacency and key loads, the
      // probes of in-range entries, then the writes sub-step by sub-step
      // (lists stay in adjacency order)
      uint32_t cis[CET_UNROLL], mvs[CET_UNROLL], ks[CET_UNROLL];
      bool hits[CET_UNROLL];
      uint32_t es[CET_UNROLL];
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) cis[u] = f + u * WAVE + lane;
      halving_search<CET_UNROLL>(L.pre, WAVE, cis);
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) es[u] = L.off[cis[u]] + (f + u * WAVE + lane - L.pre[cis[u]]);
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) {
        mvs[u] = 0;
        ks[u] = 0;
        hits[u] = false;
        if (f + u * WAVE + lane < D) {
          mvs[u] = col[es[u]];
          ks[u] = ckey[es[u]];
        }
      }
      // the probes' bitmap words loaded together, then tested
      uint32_t pw[CET_UNROLL];
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u)
        pw[u] = f + u * WAVE + lane < D && ks[u] >= wlo && ks[u] <= whi ? bm_w[mvs[u] >> 5] >> (mvs[u] & 31) : 0u;
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) hits[u] = (pw[u] & 1u) != 0;
      if (stats) {
#pragma unroll
        for (int u = 0; u < CET_UNROLL; ++u) {
          tl.ent += f + u * WAVE + lane < D;
          tl.probe += f + u * WAVE + lane < D && ks[u] >= wlo && ks[u] <= whi;
        }
      }
#pragma unroll
      for (int u = 0; u < CET_UNROLL; ++u) {
        const uint32_t fu = f + u * WAVE, ci = cis[u];
        const unsigned long long hm = __ballot(hits[u]);
        if (hits[u]) {
          // lanes of this sub-step that belong to candidate ci start at seg_lo
          const uint32_t seg_lo = L.pre[ci] > fu ? L.pre[ci] - fu : 0u;
          const unsigned long long below = (lane ? (~0ULL >> (WAVE - lane)) : 0ULL) & (~0ULL << seg_lo);
          const uint32_t rank = __popcll(hm & below);
          const uint32_t pos = L.cnt[ci] + rank;
          if (pos < d.list_cap) {
            out[pos] = mvs[u];
            ++tl.wrote;
          }
        }
        wave_sync();
        if (hits[u]) atomicAdd(&L.cnt[ci], 1u);
        wave_sync();
      }
    }
  }
  if (stats) flush_tally(tl, stats);
}

// One thread per term: the table by a binary search over the term ranges.
__global__ void __launch_bounds__(256) cet_terms_kernel(const uint32_t* __restrict__ lists,
                                                        const CetDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ term_desc,
                                                        const uint64_t* __restrict__ term_start, uint32_t nd,
                                                        uint64_t nterms, unsigned long long* __restrict__ terms) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nterms; t += stride) {
    uint32_t lo = 0, hi = nd;  // last range with start <= t
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (term_start[mid] <= t) lo = mid;
      else hi = mid;
    }
    const CetDesc& d = descs[term_desc[lo]];
    const uint64_t e = t - term_start[lo];
    terms[t] = e < d.list_cap ? tuple_term(d.term_qid, lists[d.list_base + e]) : 0ull;
  }
}

// base[d] = chunks of 64 candidates in the tables before d (exclusive scan
// over the tables, one workgroup of 256 threads).
__global__ void __launch_bounds__(256) chunk_base_kernel(const CetDesc* __restrict__ descs, uint32_t ndesc,
                                                         uint32_t* __restrict__ base) {
  __shared__ uint32_t tmp[8];
  uint32_t carry = 0;
  for (uint32_t d0 = 0; d0 < ndesc; d0 += blockDim.x) {
    const uint32_t d = d0 + threadIdx.x;
    const uint32_t ch = d < ndesc ? (descs[d].ncand + WAVE - 1) / WAVE : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan(ch, tmp, &total);
    if (d < ndesc) base[d] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) base[ndesc] = carry;
}

}  // namespace

void launch_cet_count(const DevGraph& g, const uint32_t* bitmaps, const uint32_t* clist, CetDesc* descs,
                      uint32_t ndesc, uint32_t* offsets, int per_cand, uint32_t* chunk_base, uint32_t grid,
                      hipStream_t s, unsigned long long* stats) {
  if (!ndesc) return;
  hipLaunchKernelGGL(chunk_base_kernel, dim3(1), dim3(256), 0, s, descs, ndesc, chunk_base);
  hipLaunchKernelGGL(cet_count_kernel, dim3(grid), dim3(256), 0, s, g.row, g.col, g.ckey, bitmaps, g.W, clist,
                     descs, chunk_base, ndesc, offsets, per_cand, stats);
}

void launch_cet_offsets(CetDesc* descs, uint32_t ndesc, uint32_t* offsets, hipStream_t s) {
  if (!ndesc) return;
  hipLaunchKernelGGL(cet_offsets_kernel, dim3(ndesc), dim3(256), 0, s, descs, offsets);
}

void launch_cet_terms(const uint32_t* lists, const CetDesc* descs, const uint32_t* term_desc,
                      const uint64_t* term_start, uint32_t nterm_desc, uint64_t nterms, unsigned long long* terms,
                      hipStream_t s) {
  if (!nterm_desc || !nterms) return;
  uint64_t grid = (nterms + 255) / 256;
  if (grid > 16384) grid = 16384;
  hipLaunchKernelGGL(cet_terms_kernel, dim3((uint32_t)grid), dim3(256), 0, s, lists, descs, term_desc, term_start,
                     nterm_desc, nterms, terms);
}

void launch_cet_write(const DevGraph& g, const uint32_t* bitmaps, const uint32_t* clist, const CetDesc* descs,
                      uint32_t ndesc, const uint32_t* offsets, uint32_t* lists, uint32_t* chunk_base, uint32_t grid,
                      hipStream_t s, unsigned long long* stats) {
  if (!ndesc) return;
  hipLaunchKernelGGL(chunk_base_kernel, dim3(1), dim3(256), 0, s, descs, ndesc, chunk_base);
  hipLaunchKernelGGL(cet_write_kernel, dim3(grid), dim3(256), 0, s, g.row, g.col, g.ckey, bitmaps, g.W, clist,
                     descs, chunk_base, ndesc, offsets, lists, stats);
}

}  // namespace gpsm
