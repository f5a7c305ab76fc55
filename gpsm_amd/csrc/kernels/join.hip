// Join of candidate-edge tables on gfx950 (table_kernel, join_kernel,
// expand_kernel, link_kernel, filter_kernel; match/Match.cu:1141-1569).
//
// The reference materialises every intermediate table breadth-first, finds
// a row's vertex in the candidate list with a warp-linear scan (quadratic in
// rows x candidates), and leaves injectivity to the host.  Here:
//  * a persistent grid pulls tasks (every T-th row of one query's first
//    table) from a device counter, so hundreds of uneven queries share one
//    launch and hub rows are spread over a query's T tasks;
//  * each workgroup walks its slice depth-first through levels k = 1..q-1
//    (k+1 bound vertices) with one bounded column-major buffer per level, so
//    memory is O(cap * q^2) per workgroup whatever the result size;
//  * a row's candidate list is found in O(1) through the rank index of C(x)
//    (set bits before each 256-bit block + popcounts);
//  * a window of 256 rows is expanded flattened: 256 threads take 256
//    consecutive (row, candidate) pairs, so rows of any fan-out share lanes;
//  * join-mode edges become edge-membership checks (binary search in the
//    shorter adjacency) and injectivity is checked in the same pass;
//  * the last level is never written: each embedding adds its tuple hash
//    and 1 to the counters.  The tuple hash is a product of one factor per
//    (query vertex, image) pair, so a row carries the product of its
//    factors and an embedding costs one multiply by its last factor
//    (precomputed per list entry); a wave expanding one long row sums the
//    factors of its valid candidates and multiplies once per row.
#include <algorithm>
#include <type_traits>

#include "launch.h"
#include "wave.h"

namespace gpsm {

namespace {

constexpr uint32_t WROWS = 256;  // rows per window = threads per workgroup
static_assert((WROWS & (WROWS - 1)) == 0, "row indices are masked with WROWS - 1");
// candidates in flight per thread: cyclic plans, tree plans (with the
// waves per SIMD of the two kernels, below)
constexpr int EXP_UNROLL = 4;
constexpr int TREE_UNROLL = 3;
// At the last level a row with at least this many candidates is expanded by
// one wave on its own (no row search per candidate, the row's values
// uniform); shorter rows share lanes through the flattened expansion.
// Measured (join per batch): 96 equals 128 on the 1M Chung-Lu mix and is 1%
// faster on the R-MAT 22 shapes; 192 is 2% slower on the mix.
constexpr uint32_t LONG_ROW = 96;
static_assert((uint64_t)WROWS * LONG_ROW < (1u << 23) && WROWS < 512, "packed scan of short spans and long flags");

__device__ __forceinline__ uint32_t rank_of(const uint32_t* __restrict__ bm, const uint32_t* __restrict__ bp,
                                            uint32_t x) {
  const uint32_t blk = x >> 8;
  uint32_t r = bp[blk];
  const uint32_t w0 = blk * 8, wx = x >> 5;
  for (uint32_t w = w0; w < wx; ++w) r += __popc(bm[w]);
  return r + __popc(bm[wx] & ((1u << (x & 31)) - 1u));
}

// Edge {a, b}: one bit of a hub's adjacency bitmap when either end is a hub,
// else a binary search in the shorter of the two (short) adjacency lists.
template <bool BYTES>
__device__ __forceinline__ bool has_edge(const uint32_t* __restrict__ row, const uint32_t* __restrict__ col,
                                         const uint32_t* __restrict__ hub_of, const uint32_t* __restrict__ hub_bits,
                                         uint32_t W, uint32_t a, uint32_t b, unsigned long long& bytes) {
  const uint32_t ha = hub_of[a], hb = hub_of[b];
  if (BYTES) bytes += 8;
  if (ha != ~0u || hb != ~0u) {
    if (BYTES) bytes += 4;
    return ha != ~0u ? bit_of(hub_bits + (size_t)ha * W, b) != 0 : bit_of(hub_bits + (size_t)hb * W, a) != 0;
  }
  uint32_t lo = row[a], n = row[a + 1] - lo;
  const uint32_t lo2 = row[b], n2 = row[b + 1] - lo2;
  uint32_t key = b;
  if (n2 < n) {
    lo = lo2;
    n = n2;
    key = a;
  }
  if (BYTES) bytes += 16;
  if (n == 0) return false;
  // branch-free search for the last entry <= key: ceil(log2 n) halvings, one
  // comparison at the end (no early exit, so no per-step exec-mask updates)
  if (BYTES) bytes += 4ull * (1 + (n > 1 ? 32 - __clz(n - 1) : 0));  // ceil(log2 n) probes + the last read
  const uint32_t* p = col + lo;
  while (n > 1) {
    const uint32_t half = n >> 1;
    p = p[half] <= key ? p + half : p;
    n -= half;
  }
  return *p == key;
}

// The back-edge tests of a lane's EXP_UNROLL candidates of one row (row i
// of the window, the wave-per-row pass) against the bound vertices of
// positions P.backs[bo..be): those vertices are the same for the whole
// wave, so each one's hub index is read once, and when it is a hub the
// candidates' tests are bits of its bitmap, loaded together (one load per
// candidate and edge).  Otherwise the candidates' own hub indices are read
// together and a pair with no hub falls back to has_edge's binary search.
// A candidate whose test fails skips the later ones.  oks[u] in/out.
template <bool BYTES, int U>
__device__ __forceinline__ void back_edges_of_row(const JoinArgs& a, const JoinPlan& P,
                                                  const uint32_t* __restrict__ val, uint32_t bo, uint32_t be,
                                                  uint32_t i, const uint32_t (&mvs)[U],
                                                  bool (&oks)[U], unsigned long long& bytes) {
  for (uint32_t b = bo; b < be; ++b) {
    const uint32_t bv = uniform(val[P.backs[b] * WROWS + i]);
    const uint32_t hb = uniform(a.hub_of[bv]);
    if (hb != ~0u) {
      const uint32_t* bm = a.hub_bits + (size_t)hb * a.W;
      uint32_t w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = oks[u] ? bm[mvs[u] >> 5] >> (mvs[u] & 31) : 0u;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (BYTES && oks[u]) bytes += 12;  // two hub indices, one bitmap word
        oks[u] = oks[u] && (w[u] & 1u);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (oks[u]) oks[u] = has_edge<BYTES>(a.row, a.col, a.hub_of, a.hub_bits, a.W, mvs[u], bv, bytes);
    }
  }
}

struct JoinShared {
  uint32_t pre[WROWS + 1];
  uint32_t delta[WROWS];  // list position of a row's candidate g is delta[row] + g
  unsigned long long srow[WROWS];  // product of a row's hash factors
  uint32_t tmp[8];
  uint32_t fill[MAXQ], cur_row[MAXQ], cur_cand[MAXQ];
  unsigned long long lrows[MAXQ];
  unsigned long long lwork[MAXQ];
  uint32_t task, k, abort;
  unsigned long long red[4];
  // last level: the window's long rows (row, list start, candidates) and
  // the candidates they hold
  uint32_t lrow[WROWS], lst[WROWS], lub[WROWS];
  uint32_t long_cands;
  // followed by the row window: val[c * WROWS + r], c < qmax
};

// BYTES: account the algorithmic bytes read and written (JoinOut::bytes);
// the kernel without it is the one timed.  CYCLIC: the plans of the launch
// have back edges (join-mode edges); a tree plan has none, and its kernel
// holds no back-edge code.  The two are separate kernels with their own
// register budgets (join_tree_kernel, join_cyclic_kernel below): the
// cyclic one's batched back-edge tests need more registers, and sharing
// them made the tree plans' join 1.2-1.6x slower (spills at seven waves
// per SIMD, or fewer waves).
template <bool BYTES, bool CYCLIC>
__device__ __forceinline__ void join_body(const JoinArgs& a) {
  constexpr int XU = CYCLIC ? EXP_UNROLL : TREE_UNROLL;
  extern __shared__ unsigned char lds_raw[];
  JoinShared& S = *reinterpret_cast<JoinShared*>(lds_raw);
  uint32_t* val = reinterpret_cast<uint32_t*>(lds_raw + ((sizeof(JoinShared) + 15) & ~(size_t)15));
  const uint32_t tid = threadIdx.x;
  uint32_t* lvl_base = a.levels + (uint64_t)blockIdx.x * a.level_stride;

  for (;;) {
    if (tid == 0) S.task = atomicAdd(a.task_counter, 1u);
    __syncthreads();
    const uint32_t t = S.task;
    if (t >= a.ntasks) return;
    const JoinTask task = a.tasks[t];
    const JoinPlan& P = a.plans[task.query];
    const uint32_t q = P.q;
    const CetDesc d0 = a.cets[P.first_cet];
    const uint32_t* off0 = a.offsets + d0.off_base;
    const uint32_t* cl0 = a.clist + d0.clist_off;
    const uint32_t* ls0 = a.lists + d0.list_base;
    unsigned long long my_cnt = 0, my_hash = 0, my_bytes = 0;
    unsigned long long wave_cnt = 0;  // last-level embeddings of this wave (uniform)

    if (tid < MAXQ) {
      S.lrows[tid] = 0;
      S.lwork[tid] = 0;
      S.fill[tid] = 0;
      S.cur_row[tid] = 0;
      S.cur_cand[tid] = 0;
    }
    __syncthreads();
    if (tid == 0) {
      S.fill[1] = (uint32_t)task.count;
      S.lrows[1] = task.count;
      S.lwork[1] = task.count;
      S.k = 1;
      // a query already over its work budget skips its remaining tasks
      S.abort = a.budget ? (uint32_t)__hip_atomic_load(&a.out[task.query].aborted, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT)
                         : 0u;
      if (S.abort) {
        S.fill[1] = S.lwork[1] = 0;
      } else {
        // first-table rows are join work too (the whole join of a 2-vertex query)
        const unsigned long long before = atomicAdd(&a.out[task.query].work, (unsigned long long)task.count);
        if (a.budget && before + task.count > a.budget) {
          a.out[task.query].aborted = 1;
          S.abort = 1;
          S.fill[1] = 0;
        }
      }
    }
    __syncthreads();

    if (q == 2) {
      // the first table is the embedding set
      for (uint64_t i = tid; i < task.count && !S.abort; i += blockDim.x) {
        const uint64_t e = task.first + i * task.stride;
        const uint32_t j = seg_search(off0, d0.ncand + 1, (uint32_t)e);
        const uint32_t mu = cl0[j], mv = ls0[e];
        my_hash += tuple_term(P.order[0], mu) * tuple_term(P.order[1], mv);
        ++my_cnt;
        if (a.emit) {
          const unsigned long long pos = atomicAdd(a.emit_ctr, 1ULL);
          if (pos < a.emit_cap) {
            a.emit[pos * 2 + P.order[0]] = mu;
            a.emit[pos * 2 + P.order[1]] = mv;
          }
        }
      }
    } else {
      for (;;) {
        if (S.abort) break;  // uniform: set by thread 0 before the last barrier
        const uint32_t k = S.k;
        if (S.cur_row[k] >= S.fill[k]) {
          __syncthreads();
          if (k == 1) break;
          if (tid == 0) S.k = k - 1;
          __syncthreads();
          continue;
        }
        const uint32_t r0 = S.cur_row[k];
        const uint32_t nw = min(WROWS, S.fill[k] - r0);
        const uint32_t nk = k + 1;              // level being produced
        const bool final = nk == q - 1;
        const uint32_t p = P.par[nk];
        const CetDesc dn = a.cets[P.cet[nk]];
        const uint32_t* bm_x = a.bitmaps + (size_t)dn.x_slot * a.W;
        const uint32_t* bp_x = a.bprefix + (size_t)dn.x_slot * a.NB;
        const uint32_t* offn = a.offsets + dn.off_base;
        const uint32_t bo = P.back_off[nk], be = CYCLIC ? P.back_off[nk + 1] : bo;  // a tree plan has none
        const uint64_t inj = P.inj[nk];
        const uint32_t wn = P.order[nk];
        const uint32_t cap = a.cap;
        // level j (j >= 2, j+1 columns) starts at cap * (j(j+1)/2 - 3)
        const uint32_t* Lk = k >= 2 ? lvl_base + (uint64_t)cap * ((k * (k + 1)) / 2 - 3) : nullptr;
        uint32_t* Ln = final ? nullptr : lvl_base + (uint64_t)cap * ((nk * (nk + 1)) / 2 - 3);

        // load the window and each row's candidate range
        uint32_t ub = 0, st = 0;
        if (tid < nw) {
          const uint32_t r = r0 + tid;
          if (k == 1) {
            const uint64_t e = task.first + (uint64_t)r * task.stride;
            const uint32_t j = seg_search(off0, d0.ncand + 1, (uint32_t)e);
            val[0 * WROWS + tid] = cl0[j];
            val[1 * WROWS + tid] = ls0[e];
          } else {
            for (uint32_t c = 0; c <= k; ++c) val[c * WROWS + tid] = Lk[(uint64_t)c * cap + r];
          }
          unsigned long long s = 1;
          for (uint32_t c = 0; c <= k; ++c) s *= tuple_term(P.order[c], val[c * WROWS + tid]);
          S.srow[tid] = s;
          // row values, rank index (prefix + words) and the offset pair
          if (BYTES) my_bytes += (k == 1 ? 8 : 4 * (k + 1)) + 4 + 4 * (((val[p * WROWS + tid] >> 5) & 7) + 1) + 8;
          const uint32_t j = rank_of(bm_x, bp_x, val[p * WROWS + tid]);
          st = offn[j];
          ub = offn[j + 1] - st;
          if (tid == 0) {
            ub -= S.cur_cand[k];
            st += S.cur_cand[k];
          }
        }
        uint32_t T, nlong = 0;
        if (!final) {
          const uint32_t ex = block_excl_scan(ub, S.tmp, &T);
          S.pre[tid] = tid < nw ? ex : T;
          S.delta[tid] = st - ex;  // wraps; delta + g is in range for g in the row's span
          if (tid == 0) S.fill[nk] = 0;
        } else {
          // last level: long rows are listed for the wave-per-row pass, the
          // others get spans of the flattened pass (T = their candidates);
          // one scan of (short candidates << 9 | long flag)
          const bool lng = tid < nw && ub >= LONG_ROW;
          uint32_t tot;
          const uint32_t ex2 = block_excl_scan(((lng ? 0u : ub) << 9) | (lng ? 1u : 0u), S.tmp, &tot);
          T = tot >> 9;
          nlong = tot & 511u;
          S.pre[tid] = tid < nw ? ex2 >> 9 : T;
          S.delta[tid] = st - (ex2 >> 9);
          if (lng) {
            S.lrow[ex2 & 511u] = tid;
            S.lst[ex2 & 511u] = st;
            S.lub[ex2 & 511u] = ub;
          }
          if (tid == 0) S.long_cands = 0;
        }
        __syncthreads();

        // expand: XU candidates per thread per step, list loads in
        // flight together; rows of the next level are appended per wave (one
        // LDS atomic per wave; row order inside a level does not matter).
        // The last level and the others are separate instantiations.
        const uint32_t* lst = a.lists + dn.list_base;
        const uint32_t lane = lane_id();
        const unsigned long long below = lane ? (~0ULL >> (WAVE - lane)) : 0ULL;
        auto expand = [&](auto fin) -> uint32_t {
          constexpr bool FINAL = decltype(fin)::value;
          const unsigned long long* trm_n = FINAL ? a.terms + dn.term_base : nullptr;
          uint32_t f = 0;
          for (; f < T; f += WROWS * XU) {
            if constexpr (!FINAL) {
              // every thread reads fill before any wave appends in this step
              const uint32_t fl = S.fill[nk];
              __syncthreads();
              if (fl + WROWS * XU > cap) break;  // uniform
            }
            uint32_t mvs[XU], is[XU];
            unsigned long long trm[XU];
            // wave w takes the step's candidates g0 .. g0 + 255 (lane l: g0 +
            // 64u + l), so one pair of searches per wave brackets the rows they
            // fall in and each candidate searches only inside the bracket
            // (T > 0 here; indices past T are clamped to T - 1 throughout, so
            // the bracket holds the row of every clamped index)
            const uint32_t g0 = f + (tid / WAVE) * (WAVE * XU);
            uint32_t rlo, rn;
            {
              const uint32_t glast = min(g0 + WAVE * XU - 1, T - 1);
              uint32_t r = lane == 0 ? min(g0, T - 1) : glast;
              halving_search<1>(S.pre, uniform(nw), &r);
              rlo = uniform(__shfl(r, 0, WAVE));
              rn = uniform(__shfl(r, 1, WAVE)) - rlo + 1;
            }
            // each candidate's row inside the bracket, the XU
            // searches of a lane stepping together
            // a candidate index past T searches for T - 1 (its loads then
            // re-read the last candidate's entry, in bounds); the candidate
            // itself is masked out (is = WROWS)
            uint32_t rb[XU];
#pragma unroll
            for (int u = 0; u < XU; ++u) rb[u] = min(g0 + u * WAVE + lane, T - 1);
            halving_search<XU>(S.pre + rlo, rn, rb);
            // LDS reads of the XU candidates are issued together
            // (row offsets, then the bound vertices the checks compare
            // with), so a lane waits once per kind rather than once per
            // candidate
            uint32_t es[XU];
#pragma unroll
            for (int u = 0; u < XU; ++u) {
              const uint32_t g = g0 + u * WAVE + lane;
              const uint32_t i = rlo + rb[u];
              is[u] = g < T ? i : WROWS;
              es[u] = S.delta[i] + min(g, T - 1);
              trm[u] = 0;
            }
#pragma unroll
            for (int u = 0; u < XU; ++u) {
              mvs[u] = lst[es[u]];
              // last level: the entry's hash term, computed once per list entry
              if constexpr (FINAL) trm[u] = trm_n[es[u]];
            }
            bool oks[XU];
#pragma unroll
            for (int u = 0; u < XU; ++u) oks[u] = is[u] != WROWS;
            // injectivity: only the positions bound to a vertex of the same
            // label that no edge test already separates
            for (uint64_t m = inj; m; m &= m - 1) {
              const uint32_t c = (uint32_t)(__ffsll((long long)m) - 1);
#pragma unroll
              for (int u = 0; u < XU; ++u) oks[u] &= val[c * WROWS + (is[u] & (WROWS - 1))] != mvs[u];
            }
            if constexpr (!FINAL) {
              // join-mode edges, then one append to the next level for the
              // lane's XU candidates together: one LDS atomic per
              // wave and step instead of one per candidate
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                if (oks[u]) {
                  if (BYTES) my_bytes += 4;  // the list entry
                  for (uint32_t b = bo; b < be && oks[u]; ++b)
                    oks[u] = has_edge<BYTES>(a.row, a.col, a.hub_of, a.hub_bits, a.W, mvs[u],
                                             val[P.backs[b] * WROWS + is[u]], my_bytes);
                }
              }
              unsigned long long bal[XU];
              uint32_t nbal = 0;
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                bal[u] = __ballot(oks[u]);
                nbal += (uint32_t)__popcll(bal[u]);
              }
              uint32_t wbase = 0;
              if (lane == 0 && nbal) wbase = atomicAdd(&S.fill[nk], nbal);
              wbase = __shfl(wbase, 0, WAVE);
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                if (oks[u]) {
                  const uint64_t r = wbase + __popcll(bal[u] & below);
                  for (uint32_t c = 0; c <= k; ++c) Ln[(uint64_t)c * cap + r] = val[c * WROWS + is[u]];
                  Ln[(uint64_t)nk * cap + r] = mvs[u];
                  if (BYTES) my_bytes += 4 * (nk + 1);
                }
                wbase += (uint32_t)__popcll(bal[u]);
              }
            }
            if constexpr (FINAL) {
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                const uint32_t i = is[u], mv = mvs[u];
                bool ok = oks[u];
                if (ok) {
                  if (BYTES) my_bytes += 4;  // the list entry (its hash term is a derived copy of it)
                  for (uint32_t b = bo; b < be && ok; ++b)
                    ok = has_edge<BYTES>(a.row, a.col, a.hub_of, a.hub_bits, a.W, mv, val[P.backs[b] * WROWS + i],
                                         my_bytes);
                }
                wave_cnt += __popcll(__ballot(ok));
                if (ok) {
                  my_hash += S.srow[i] * trm[u];
                  if (a.emit) {
                    const unsigned long long pos = atomicAdd(a.emit_ctr, 1ULL);
                    if (pos < a.emit_cap) {
                      uint32_t* t = a.emit + pos * q;
                      for (uint32_t c = 0; c <= k; ++c) t[P.order[c]] = val[c * WROWS + i];
                      t[wn] = mv;
                    }
                  }
                }
              }
            }
            if constexpr (!FINAL) __syncthreads();
          }
          return f;
        };
        if (final && nlong) {
          // wave-per-row pass over the window's long rows: lanes take
          // consecutive list entries of one row, whose values are uniform
          const unsigned long long* trm_n = a.terms + dn.term_base;
          for (uint32_t j = tid / WAVE; j < nlong; j += WROWS / WAVE) {
            const uint32_t i = uniform(S.lrow[j]);
            const uint32_t st_r = uniform(S.lst[j]), n_r = uniform(S.lub[j]);
            const unsigned long long srow = S.srow[i];
            unsigned long long rsum = 0;  // this lane's factors of the row's valid candidates
            if (lane == 0) atomicAdd(&S.long_cands, n_r);
            for (uint32_t g0 = 0; g0 < n_r; g0 += WAVE * XU) {
              uint32_t mvs[XU];
              unsigned long long trm[XU];
              bool oks[XU];
              // every lane loads (a lane past the row's end re-reads its
              // last entry, n_r >= LONG_ROW > 0): no branch per candidate
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                const uint32_t g = g0 + u * WAVE + lane;
                oks[u] = g < n_r;
                const uint32_t e = st_r + min(g, n_r - 1);
                mvs[u] = lst[e];
                // without back edges the hash factors load beside the entries
                trm[u] = 0;
                if (!CYCLIC || bo == be) trm[u] = trm_n[e];
              }
              for (uint64_t m = inj; m; m &= m - 1) {
                const uint32_t vc = val[(uint32_t)(__ffsll((long long)m) - 1) * WROWS + i];
#pragma unroll
                for (int u = 0; u < XU; ++u) oks[u] &= mvs[u] != vc;
              }
#pragma unroll
              for (int u = 0; u < XU; ++u)
                if (BYTES && oks[u]) my_bytes += 4;  // the list entry
              if (CYCLIC && bo != be) {
                // with back edges: the tests first, then the factors of the
                // candidates that passed them
                back_edges_of_row<BYTES, XU>(a, P, val, bo, be, i, mvs, oks, my_bytes);
#pragma unroll
                for (int u = 0; u < XU; ++u)
                  if (oks[u]) trm[u] = trm_n[st_r + g0 + u * WAVE + lane];
              }
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                const bool ok = oks[u];
                wave_cnt += __popcll(__ballot(ok));
                rsum += ok ? trm[u] : 0ull;
                if (ok) {
                  if (a.emit) {
                    const unsigned long long pos = atomicAdd(a.emit_ctr, 1ULL);
                    if (pos < a.emit_cap) {
                      uint32_t* t = a.emit + pos * q;
                      for (uint32_t c = 0; c <= k; ++c) t[P.order[c]] = val[c * WROWS + i];
                      t[wn] = mvs[u];
                    }
                  }
                }
              }
            }
            my_hash += srow * rsum;
          }
        }
        const uint32_t f = final ? expand(std::true_type{}) : expand(std::false_type{});
        if (final) __syncthreads();  // long_cands complete
        if (!final && tid == 0) S.lrows[nk] += S.fill[nk];
        // advance level k's cursor past what was expanded
        if (tid == 0) {
          // join work: candidates examined (a window cut short by a full
          // buffer is reloaded later from the cursor), and the budget on it
          const unsigned long long done = (unsigned long long)min(f, T) + (final ? S.long_cands : 0u);
          const unsigned long long before = atomicAdd(&a.out[task.query].work, done);
          S.lwork[nk] += done;
          if (a.budget && before + done > a.budget) {
            a.out[task.query].aborted = 1;
            S.abort = 1;
          }
          if (f >= T) {
            S.cur_row[k] = r0 + nw;
            S.cur_cand[k] = 0;
          } else {
            const uint32_t i = seg_search(S.pre, nw, f);
            const uint32_t extra = (i == 0) ? S.cur_cand[k] : 0u;
            S.cur_row[k] = r0 + i;
            S.cur_cand[k] = extra + (f - S.pre[i]);
          }
          if (!final && S.fill[nk] > 0) {
            S.k = nk;
            S.cur_row[nk] = 0;
            S.cur_cand[nk] = 0;
          }
        }
        __syncthreads();
      }
    }

    // flush this task's counters
    my_cnt = wave_sum64(my_cnt) + wave_cnt;
    my_hash = wave_sum64(my_hash);
    __syncthreads();
    if (lane_id() == 0) S.red[tid / WAVE] = my_cnt;
    __syncthreads();
    if (tid == 0) {
      unsigned long long c = 0;
      for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) c += S.red[w];
      S.lrows[q - 1] = (q == 2) ? c : S.lrows[q - 1] + c;
      if (c) atomicAdd(&a.out[task.query].count, c);
    }
    __syncthreads();
    if (lane_id() == 0) S.red[tid / WAVE] = my_hash;
    __syncthreads();
    if (tid == 0) {
      unsigned long long h = 0;
      for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) h += S.red[w];
      atomicAdd(&a.out[task.query].hash, h);
    }
    __syncthreads();
    if constexpr (BYTES) {
      my_bytes = wave_sum64(my_bytes);
      if (lane_id() == 0) S.red[tid / WAVE] = my_bytes;
      __syncthreads();
      if (tid == 0) {
        unsigned long long bsum = 0;
        for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) bsum += S.red[w];
        atomicAdd(&a.out[task.query].bytes, bsum);
      }
    }
    if (tid >= 1 && tid < q && S.lrows[tid]) atomicAdd(&a.out[task.query].level_rows[tid], S.lrows[tid]);
    if (tid >= 1 && tid < q && S.lwork[tid]) atomicAdd(&a.out[task.query].level_work[tid], S.lwork[tid]);
    __syncthreads();
  }
}

// Eight waves per SIMD for tree plans, three candidates per thread: 64
// VGPRs with 4 spilled.  On the 1M Chung-Lu mix (join per batch, learned
// order) 438 ms, against 445 ms with four candidates at eight waves (7
// spilled), 457 ms with four at seven waves (no spill), 452 ms with two at
// eight and 483-571 ms with six or eight candidates at five to seven waves:
// the waves in flight hide the list loads better than more loads per wave.
// Five waves for cyclic plans, whose batched back-edge tests spill heavily
// beyond (R-MAT 22 cycles and cliques 2.05 s per batch at five, 2.07 s at
// six, 2.33 s at eight).
template <bool BYTES>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) join_tree_kernel(JoinArgs a) {
  join_body<BYTES, false>(a);
}

template <bool BYTES>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) join_cyclic_kernel(JoinArgs a) {
  join_body<BYTES, true>(a);
}

template <bool BYTES>
const void* join_fn(bool cyclic) {
  return cyclic ? reinterpret_cast<const void*>(join_cyclic_kernel<BYTES>)
                : reinterpret_cast<const void*>(join_tree_kernel<BYTES>);
}

}  // namespace

uint64_t join_level_words(uint32_t qmax, uint32_t cap) {
  // levels k = 2 .. qmax-2 hold k+1 columns; level k starts at cap*(k(k+1)/2 - 3)
  if (qmax < 4) return 0;
  const uint64_t k = qmax - 1;  // one past the last materialised level
  return (uint64_t)cap * (k * (k + 1) / 2 - 3);
}

namespace {
size_t join_lds_bytes(uint32_t qmax) {
  return ((sizeof(JoinShared) + 15) & ~(size_t)15) + sizeof(uint32_t) * WROWS * qmax;
}
}  // namespace

uint32_t join_resident_per_cu(uint32_t qmax, bool cyclic) {
  const size_t lds = join_lds_bytes(std::max<uint32_t>(qmax, 2));
  const void* fn = join_fn<false>(cyclic);
  if (lds > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)WROWS, lds) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  return (uint32_t)per_cu;
}

void launch_join(const JoinArgs& a, uint32_t nwg, bool cyclic, hipStream_t s) {
  const size_t lds = join_lds_bytes(a.qmax);
  const void* fn = a.count_bytes ? join_fn<true>(cyclic) : join_fn<false>(cyclic);
  if (lds > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (a.count_bytes) {
    if (cyclic) hipLaunchKernelGGL(join_cyclic_kernel<true>, dim3(nwg), dim3(WROWS), lds, s, a);
    else hipLaunchKernelGGL(join_tree_kernel<true>, dim3(nwg), dim3(WROWS), lds, s, a);
  } else {
    if (cyclic) hipLaunchKernelGGL(join_cyclic_kernel<false>, dim3(nwg), dim3(WROWS), lds, s, a);
    else hipLaunchKernelGGL(join_tree_kernel<false>, dim3(nwg), dim3(WROWS), lds, s, a);
  }
}

}  // namespace gpsm
