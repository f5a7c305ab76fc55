// Candidate-vertex filtering on gfx950: the start vertex's label/degree check
// (check_kernel, match/Match.cu:431-461), exploration (explore_kernel,
// :528-659) and neighbourhood refinement (refine_kernel, :761-846), for any
// number of queries per launch.
//
// Layout: candidate sets are bitmaps (1 bit per data vertex, 32-bit words);
// the reference used one byte per (query vertex, data vertex).
// Parallel decomposition: one wave per (op, vertex range).  The ranges are
// fixed per data graph (DevGraph::fchunk): at most 1024 vertices and 4096
// adjacency entries each, so a power-law graph's hubs (which a Chung-Lu or
// R-MAT generator gives the lowest ids) do not pile onto one wave.  The wave
// compacts the range's candidates into LDS (the reference's kernel_collect,
// fused), then walks the concatenated adjacency lists of 64 candidates at a
// time, 64 adjacency entries per lane step (lane t takes entry t; a binary
// search in the 64 degree prefix sums finds its candidate), so hub and leaf
// candidates share lanes evenly instead of one 32-lane warp per candidate.
#include "launch.h"
#include "wave.h"

namespace gpsm {

namespace {

constexpr int MAX_CAND = FCHUNK_VERTS;  // candidates of one vertex range
static_assert(FCHUNK_VERTS <= 32 * (WAVE - 1), "a range's bitmap words fit one wave");

// Start-vertex candidates of every op.  A wave owns 64 consecutive bitmap
// words (2048 vertices); lane l loads the keys of vertices base + l + 64 j,
// j < 32 (each load instruction reads 256 contiguous bytes), keeps them in
// registers, and for each op of its group one ballot per j gives the two
// words 2j and 2j + 1, which lanes 2j and 2j + 1 keep.  The keys are read
// once for all ops, and each op's 64 words leave as one coalesced store.
constexpr uint32_t INIT_OPS = 16;  // ops per wave

__global__ void __launch_bounds__(256) init_kernel(const uint32_t* __restrict__ key, uint32_t n, uint32_t W,
                                                   const InitOp* __restrict__ ops, uint32_t nops,
                                                   uint32_t* __restrict__ bitmaps) {
  const uint32_t lane = lane_id();
  const uint32_t blocks = (W + WAVE - 1) / WAVE, groups = (nops + INIT_OPS - 1) / INIT_OPS;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (wave >= (uint64_t)blocks * groups) return;  // the whole wave
  const uint32_t wb = (uint32_t)(wave % blocks), og = (uint32_t)(wave / blocks);
  const uint32_t v0 = wb * WAVE * 32;
  uint32_t k[32];
#pragma unroll
  for (uint32_t j = 0; j < 32; ++j) {
    const uint32_t v = v0 + j * WAVE + lane;
    k[j] = v < n ? key[v] : 0u;  // key 0 matches no op (dense labels start at 1)
  }
  const uint32_t w = wb * WAVE + lane;
  const uint32_t o_end = min(nops, (og + 1) * INIT_OPS);
  for (uint32_t o = og * INIT_OPS; o < o_end; ++o) {
    const uint32_t lo = uniform(ops[o].lo), hi = uniform(ops[o].hi), slot = uniform(ops[o].slot);
    uint32_t word = 0;
#pragma unroll
    for (uint32_t j = 0; j < 32; ++j) {
      const unsigned long long m = __ballot(k[j] >= lo && k[j] <= hi);
      if (lane == 2 * j) word = (uint32_t)m;
      if (lane == 2 * j + 1) word = (uint32_t)(m >> 32);
    }
    if (w < W) bitmaps[(size_t)slot * W + w] = word;
  }
}

// The (op, vertex range) pairs of a superstep whose range holds at least one
// candidate of the op's C(u), appended to items[] (order arbitrary: a
// superstep's updates are atomic bit operations on distinct slots per
// query, so the order of its items does not change the result).  Once the
// sets are small most ranges are empty, and the superstep then costs one
// coalesced pass over the bitmap words instead of one wave per pair.
// One wave per (op, 256 consecutive ranges); lane l reads words l, l + 64,
// ... of the ranges' span, eight loads in flight, together with the range of
// each word's first vertex (wrange), and a nonzero word marks its range
// (and, when it runs past that range's end, the next ones it has bits in).
// The four waves of a workgroup take consecutive range groups of one op,
// and the list is appended to with one atomic per workgroup: appends to a
// single counter serialise, and one per wave cost as much as the scan.
constexpr uint32_t AR_GROUP = 4 * WAVE;  // vertex ranges per wave
constexpr int AR_UNROLL = 8;             // bitmap words in flight per lane

__global__ void __launch_bounds__(256) active_ranges_kernel(const uint32_t* __restrict__ fchunk,
                                                             const uint32_t* __restrict__ wrange, uint32_t chunks,
                                                             const FilterOp* __restrict__ ops, uint32_t nops,
                                                             const uint32_t* __restrict__ bitmaps, uint32_t W,
                                                             const uint32_t* __restrict__ rflags,
                                                             uint2* __restrict__ items, uint32_t* __restrict__ nitems) {
  __shared__ uint32_t s_lo[4][AR_GROUP + 1];
  __shared__ uint32_t s_act[4][AR_GROUP];
  __shared__ uint32_t s_wtot[4], s_wbase[4];
  const uint32_t lane = lane_id(), wv = threadIdx.x / WAVE;
  const uint32_t groups = (chunks + AR_GROUP - 1) / AR_GROUP;
  const uint32_t wgs_per_op = (groups + 3) / 4;
  const uint32_t o_wg = blockIdx.x / wgs_per_op;  // the workgroup's op
  const uint32_t g = (blockIdx.x % wgs_per_op) * 4 + wv;
  uint32_t* lo_b = s_lo[wv];
  uint32_t* act = s_act[wv];
#pragma unroll
  for (uint32_t j = 0; j < AR_GROUP / WAVE; ++j) act[j * WAVE + lane] = 0;
  const uint32_t o = o_wg;
  uint32_t c0 = 0, nc = 0;
  if (o < nops && g < groups) {  // uniform per wave
    c0 = g * AR_GROUP;
    const uint32_t need = uniform(ops[o].need);
    // an op whose previous refine round removed nothing lists no range
    if (need == ~0u || uniform(rflags[need]) != 0) nc = min(AR_GROUP, chunks - c0);
  }
  if (nc) {
    for (uint32_t j = lane; j <= nc; j += WAVE) lo_b[j] = fchunk[c0 + j];
    wave_sync();
    const uint32_t v_begin = lo_b[0], v_end = lo_b[nc];
    const uint32_t w_begin = v_begin >> 5, w_last = (v_end - 1) >> 5;
    const uint32_t* bm = bitmaps + (size_t)uniform(ops[o].u_slot) * W;
    for (uint32_t w0 = w_begin; w0 <= w_last; w0 += WAVE * AR_UNROLL) {
      uint32_t xs[AR_UNROLL], rs[AR_UNROLL];
#pragma unroll
      for (int u = 0; u < AR_UNROLL; ++u) {
        const uint32_t w = w0 + u * WAVE + lane;
        xs[u] = 0;
        rs[u] = 0;
        if (w <= w_last) {
          xs[u] = bm[w];
          rs[u] = wrange[w];
        }
      }
#pragma unroll
      for (int u = 0; u < AR_UNROLL; ++u) {
        const uint32_t w = w0 + u * WAVE + lane;
        uint32_t x = xs[u];
        if (w == w_begin) x &= ~0u << (v_begin & 31);
        if (w == w_last && (v_end & 31)) x &= (1u << (v_end & 31)) - 1u;
        // range a holds the lowest remaining set bit: the word's first range
        // (c0 or later once the bits below v_begin are masked), then forward
        uint32_t a = rs[u] > c0 ? rs[u] - c0 : 0u;
        while (x) {
          const uint32_t v = w * 32 + (uint32_t)(__ffs(x) - 1);
          while (lo_b[a + 1] <= v) ++a;
          act[a] = 1;
          const uint32_t next = lo_b[a + 1];  // first vertex after range a (> v)
          if (next >= w * 32 + 32) break;
          x &= ~0u << (next - w * 32);
        }
      }
    }
  }
  wave_sync();
  // lane l owns ranges l, l + 64, l + 128, l + 192 of the wave's group
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t j = 0; j < AR_GROUP / WAVE; ++j) mine += act[j * WAVE + lane] != 0;
  const uint32_t incl = wave_incl_scan(mine);
  if (lane == WAVE - 1) s_wtot[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t x = 0; x < blockDim.x / WAVE; ++x) {
      s_wbase[x] = tot;
      tot += s_wtot[x];
    }
    const uint32_t base = tot ? atomicAdd(nitems, tot) : 0u;
    for (uint32_t x = 0; x < blockDim.x / WAVE; ++x) s_wbase[x] += base;
  }
  __syncthreads();
  uint32_t pos = s_wbase[wv] + incl - mine;
#pragma unroll
  for (uint32_t j = 0; j < AR_GROUP / WAVE; ++j)
    if (act[j * WAVE + lane]) items[pos++] = make_uint2(o, c0 + j * WAVE + lane);
}

constexpr int UNROLL = 4;  // adjacency entries in flight per lane

struct WaveLds {
  uint32_t cand[MAX_CAND];
  uint32_t pre[WAVE + 1];
  uint32_t base[WAVE];
  unsigned long long hit[WAVE];
  uint32_t lim[WAVE];
};

__global__ void __launch_bounds__(256) filter_step_kernel(const uint32_t* __restrict__ row,
                                                          const uint32_t* __restrict__ col,
                                                          const uint32_t* __restrict__ ckey, uint32_t W,
                                                          const uint32_t* __restrict__ fchunk,
                                                          const uint2* __restrict__ list,
                                                          const uint32_t* __restrict__ nlist,
                                                          const FilterOp* __restrict__ ops,
                                                          const uint32_t* __restrict__ nbr_slot,
                                                          const uint32_t* __restrict__ nbr_lo,
                                                          const uint32_t* __restrict__ nbr_hi,
                                                          uint32_t* __restrict__ bitmaps,
                                                          uint32_t* __restrict__ rflags,
                                                          unsigned long long* __restrict__ stats) {
  extern __shared__ unsigned char lds_raw[];
  WaveLds& L = reinterpret_cast<WaveLds*>(lds_raw)[threadIdx.x / WAVE];
  const uint32_t lane = lane_id();
  const uint32_t items = uniform(*nlist);
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / WAVE;
  // per-lane tallies for the optional byte accounting, over all items
  unsigned long long t_cand = 0;
  uint32_t t_ent = 0, t_key = 0, t_probe = 0, t_set = 0;

  // the next item's list entry is loaded while the current one runs
  uint2 nxt = wave < items ? list[wave] : make_uint2(0u, 0u);
  for (uint32_t it = wave; it < items; it += nwaves) {
    const uint2 item = nxt;
    if (it + nwaves < items) nxt = list[it + nwaves];
    const uint32_t o = uniform(item.x);
    const uint32_t c = uniform(item.y);
    const uint32_t type = uniform(ops[o].type);
    const uint32_t nv = uniform(ops[o].nv);
    const uint32_t nb = uniform(ops[o].nbr_off);
    uint32_t* bm_u = bitmaps + (size_t)uniform(ops[o].u_slot) * W;
    const unsigned long long full = nv >= 64 ? ~0ULL : ((1ULL << nv) - 1);

    // collect: candidates of vertex range [v_lo, v_hi), ascending; lane l
    // reads bitmap word w_lo + l and keeps the range's bits of it
    const uint32_t v_lo = uniform(fchunk[c]), v_hi = uniform(fchunk[c + 1]);
    const uint32_t w_lo = v_lo >> 5, nw = ((v_hi - 1) >> 5) - w_lo + 1;
    const uint32_t wi = w_lo + lane;
    uint32_t w = 0;
    if (lane < nw) {
      w = bm_u[wi];
      if (lane == 0) w &= ~0u << (v_lo & 31);
      if (lane == nw - 1 && (v_hi & 31)) w &= (1u << (v_hi & 31)) - 1u;
    }
    const uint32_t pc = __popc(w);
    const uint32_t incl = wave_incl_scan(pc);
    const uint32_t C = uniform(__shfl(incl, WAVE - 1, WAVE));
    if (lane == 0) t_cand += C;
    {
      uint32_t pos = incl - pc, x = w;
      while (x) {
        const uint32_t b = __ffs(x) - 1;
        x &= x - 1;
        L.cand[pos++] = wi * 32 + b;
      }
    }
    wave_sync();

    for (uint32_t s0 = 0; s0 < C; s0 += WAVE) {
      // lane = one candidate of the sub-chunk
      const bool mine = s0 + lane < C;
      const uint32_t id = mine ? L.cand[s0 + lane] : 0u;
      const uint32_t off = mine ? row[id] : 0u;
      const uint32_t deg = mine ? row[id + 1] - off : 0u;
      L.hit[lane] = 0;
      uint32_t cur = 0;  // entries of this candidate already examined
      unsigned long long h = 0;
      wave_sync();

      // pass 1: constraints met by each candidate.  Every step re-flattens
      // the remaining adjacency of the candidates not yet satisfying all of
      // them, so a satisfied candidate stops costing lanes.
      for (;;) {
        const uint32_t rem = (mine && h != full) ? deg - cur : 0u;
        const uint32_t rincl = wave_incl_scan(rem);
        const uint32_t T = uniform(__shfl(rincl, WAVE - 1, WAVE));
        if (T == 0) break;
        const uint32_t rpre = rincl - rem;
        L.pre[lane] = rpre;
        L.base[lane] = off + cur;
        wave_sync();
        // refine needs the neighbour ids, explore only their keys
        const uint32_t* __restrict__ src = type == OP_REFINE ? col : ckey;
        uint32_t mvs[UNROLL], cis[UNROLL], es[UNROLL];
        // an entry index past T reads entry T - 1 (in bounds) and is masked
        // out (cis = WAVE): every lane loads, no branch per entry
        uint32_t gi[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) cis[u] = gi[u] = min(u * WAVE + lane, T - 1);
        halving_search<UNROLL>(L.pre, WAVE, cis);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) es[u] = L.base[cis[u]] + (gi[u] - L.pre[cis[u]]);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          mvs[u] = src[es[u]];
          if (u * WAVE + lane >= T) cis[u] = WAVE;
        }
        // constraints each entry meets, the UNROLL entries of a lane side by
        // side: their candidates' known bits read together, then per
        // neighbour v_i one probe (refine) or key comparison (explore) each,
        // the probes' loads in flight together
        unsigned long long kn[UNROLL], bits[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          kn[u] = cis[u] != WAVE ? L.hit[cis[u]] : ~0ULL;
          bits[u] = 0;
        }
        if (type == OP_REFINE) {
          for (uint32_t i = 0; i < nv; ++i) {
            const uint32_t* bm = bitmaps + (size_t)nbr_slot[nb + i] * W;
            uint32_t wv[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) wv[u] = (kn[u] >> i) & 1ULL ? 0u : bm[mvs[u] >> 5];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
              if (!((kn[u] >> i) & 1ULL)) {
                ++t_probe;
                if ((wv[u] >> (mvs[u] & 31)) & 1u) bits[u] |= 1ULL << i;
              }
          }
        } else {
          // explore compares the entry's key (read in place of the column
          // index: no extra bytes) with v_i's label/degree range
          for (uint32_t i = 0; i < nv; ++i) {
            const uint32_t lo = nbr_lo[nb + i], hi = nbr_hi[nb + i];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
              if (mvs[u] >= lo && mvs[u] <= hi) bits[u] |= 1ULL << i;
          }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          if (cis[u] == WAVE) continue;
          ++t_ent;
          if (bits[u]) atomicOr(&L.hit[cis[u]], bits[u]);
        }
        wave_sync();
        // entries of this lane's candidate covered by the UNROLL*64 examined
        const uint32_t span = UNROLL * WAVE;
        cur += rpre >= span ? 0u : min(rem, span - rpre);
        h = L.hit[lane];
      }

      if (type == OP_REFINE) {
        const bool drop = mine && h != full;
        if (drop) clear_bit(bm_u, id);
        if (__ballot(drop) && lane == 0) rflags[uniform(ops[o].mark)] = 1u;
      } else {
        // first neighbour without a match: mu leaves C(u); neighbours before
        // it received all of mu's matching data neighbours (match/Match.cu:547-600)
        const uint32_t ff = h == full ? nv : (uint32_t)(__ffsll((long long)~h) - 1);
        if (mine && ff < nv) clear_bit(bm_u, id);
        const uint32_t lim = mine ? ff : 0u;
        const uint32_t len = lim ? deg : 0u;
        const uint32_t lincl = wave_incl_scan(len);
        const uint32_t T = uniform(__shfl(lincl, WAVE - 1, WAVE));
        L.pre[lane] = lincl - len;
        L.base[lane] = off;
        L.lim[lane] = lim;
        wave_sync();
        // pass 2: every matching data neighbour of a surviving prefix enters C(v_i)
        for (uint32_t f = 0; f < T; f += UNROLL * WAVE) {
          uint32_t mvs[UNROLL], ks[UNROLL], cis[UNROLL], es[UNROLL], gi[UNROLL];
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) cis[u] = gi[u] = min(f + u * WAVE + lane, T - 1);
          halving_search<UNROLL>(L.pre, WAVE, cis);
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) es[u] = L.base[cis[u]] + (gi[u] - L.pre[cis[u]]);
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) {
            mvs[u] = col[es[u]];
            ks[u] = ckey[es[u]];
            if (f + u * WAVE + lane >= T) cis[u] = WAVE;
          }
          // the UNROLL entries of a lane side by side: their prefix limits
          // read together, then per neighbour v_i the bitmap words of the
          // entries that enter C(v_i) loaded together, and a bit set (one
          // atomic) only where it is still clear
          uint32_t lms[UNROLL];
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) {
            lms[u] = cis[u] != WAVE ? L.lim[cis[u]] : 0u;
            if (cis[u] != WAVE) {
              ++t_ent;
              ++t_key;
            }
          }
          for (uint32_t i = 0; i < nv; ++i) {  // lim <= nv
            const uint32_t lo = nbr_lo[nb + i], hi = nbr_hi[nb + i];
            uint32_t* bm = bitmaps + (size_t)nbr_slot[nb + i] * W;
            bool want[UNROLL];
            uint32_t cw[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
              want[u] = i < lms[u] && ks[u] >= lo && ks[u] <= hi;
              cw[u] = want[u] ? __hip_atomic_load(bm + (mvs[u] >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : ~0u;
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
              if (want[u]) {
                ++t_set;
                const uint32_t m = 1u << (mvs[u] & 31);
                if (!(cw[u] & m)) atomicOr(bm + (mvs[u] >> 5), m);
              }
          }
        }
      }
      wave_sync();
    }
  }
  if (stats) {
    // one atomic per counter per workgroup, spread over STAT_COPIES copies
    __shared__ unsigned long long red[4][5];
    const unsigned long long v[5] = {t_cand, wave_sum64(t_ent), wave_sum64(t_key), wave_sum64(t_probe),
                                     wave_sum64(t_set)};
    const uint32_t w = threadIdx.x / WAVE;
    if (lane == 0)
      for (int i = 0; i < 5; ++i) red[w][i] = v[i];
    __syncthreads();
    if (threadIdx.x < 5) {
      unsigned long long tot = 0;
      for (uint32_t x = 0; x < blockDim.x / WAVE; ++x) tot += red[x][threadIdx.x];
      if (tot) atomicAdd(stats + (blockIdx.x % STAT_COPIES) * STAT_STRIDE + threadIdx.x, tot);
    }
  }
}

}  // namespace

void launch_init(const DevGraph& g, const InitOp* ops, uint32_t nops, uint32_t* bitmaps, hipStream_t s) {
  if (!nops) return;
  const uint64_t waves = (uint64_t)((g.W + WAVE - 1) / WAVE) * ((nops + INIT_OPS - 1) / INIT_OPS);
  hipLaunchKernelGGL(init_kernel, dim3((uint32_t)((waves + 3) / 4)), dim3(256), 0, s, g.key, g.n, g.W, ops, nops,
                     bitmaps);
}

void launch_filter_step(const DevGraph& g, const FilterOp* ops, uint32_t nops, const uint32_t* nbr_slot,
                        const uint32_t* nbr_lo, const uint32_t* nbr_hi, uint32_t* bitmaps, uint32_t block,
                        uint32_t max_wg, uint2* items, uint32_t* nitems, uint32_t* rflags,
                        unsigned long long* stats, hipStream_t s) {
  if (!nops) return;
  if (block != 64 && block != 128 && block != 256) block = 256;
  const uint64_t pairs = (uint64_t)nops * g.nfchunk;
  const uint32_t wgs_per_op = ((g.nfchunk + AR_GROUP - 1) / AR_GROUP + 3) / 4;
  // four waves per workgroup: the kernel's LDS and its per-workgroup append assume 256 threads
  hipLaunchKernelGGL(active_ranges_kernel, dim3(nops * wgs_per_op), dim3(256), 0, s, g.fchunk, g.wrange, g.nfchunk,
                     ops, nops, bitmaps, g.W, rflags, items, nitems);
  // the list's length is known on the device only: a grid of at most
  // max_wg workgroups, and at most FILTER_OVERSUB times as many as are
  // resident at once, strides over it (waves past its end leave at once).
  // Workgroups beyond the resident ones start as earlier ones finish, which
  // evens out uneven items (R-MAT hub ranges); with far more, the waves in
  // flight touch more bitmaps at a time.  Measured per batch: 1x resident
  // 10.9 ms (1M Chung-Lu mix) / 24.0 ms (R-MAT 22 shapes), 8x 11.1 / 20.9,
  // 65,536 workgroups 13.3 / 19.5; items taken from a shared counter (16
  // per atomic) 14.1 / 28.4, the atomics serialising across the XCDs.
  const uint32_t wpb = block / WAVE;
  const size_t lds = sizeof(WaveLds) * wpb;
  uint64_t grid = (pairs + wpb - 1) / wpb;
  if (grid > max_wg) grid = max_wg;
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, filter_step_kernel, (int)block, lds) == hipSuccess &&
      ncu > 0 && per_cu > 0)
    grid = std::min<uint64_t>(grid, (uint64_t)ncu * per_cu * FILTER_OVERSUB);
  hipLaunchKernelGGL(filter_step_kernel, dim3((uint32_t)grid), dim3(block), lds, s, g.row, g.col, g.ckey, g.W,
                     g.fchunk, items, nitems, ops, nbr_slot, nbr_lo, nbr_hi, bitmaps, rflags, stats);
}

void filter_chunks(uint32_t n, const uint32_t* row_offset, std::vector<uint32_t>& bounds) {
  bounds.assign(1, 0);
  uint32_t v = 0;
  while (v < n) {
    uint32_t end = v + 1;
    uint64_t entries = row_offset[v + 1] - row_offset[v];
    while (end < n && end - v < FCHUNK_VERTS) {
      const uint64_t d = row_offset[end + 1] - row_offset[end];
      if (entries + d > FCHUNK_ENTRIES) break;
      entries += d;
      ++end;
    }
    bounds.push_back(end);
    v = end;
  }
}

}  // namespace gpsm
