// Host launchers of the device kernels (one translation unit per phase).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "gpsm_dev.h"

namespace gpsm {

struct DevGraph {
  uint32_t n = 0;
  uint64_t adj = 0;
  uint32_t W = 0;   // bitmap words per query vertex
  uint32_t NB = 0;  // rank-index entries per query vertex
  uint32_t* row = nullptr;
  uint32_t* col = nullptr;
  uint32_t* key = nullptr;
  // ckey[e] = key[col[e]]: the filter's label/degree tests stream the keys
  // of a row's neighbours beside the row instead of gathering them
  uint32_t* ckey = nullptr;
  // filter work items: vertex ranges [fchunk[i], fchunk[i+1]) of at most
  // FCHUNK_VERTS vertices and FCHUNK_ENTRIES adjacency entries (a vertex
  // with more entries is a range of its own), so that no wave of the filter
  // walks far more adjacency than the others (power-law hubs)
  uint32_t* fchunk = nullptr;
  uint32_t nfchunk = 0;
  // wrange[w]: the vertex range holding vertex 32w (one per bitmap word)
  uint32_t* wrange = nullptr;
  // hubs (degree above hub_degree): hub_of[v] = index of v's adjacency
  // bitmap (W words each, in hub_bits) or ~0u, so an edge test touching a
  // hub is one bit; other edge tests search the shorter adjacency list
  uint32_t* hub_of = nullptr;
  uint32_t* hub_bits = nullptr;
  uint32_t nhub = 0;
  uint32_t hub_degree = 0;
};

constexpr uint32_t FCHUNK_VERTS = 1024;
constexpr uint32_t FCHUNK_ENTRIES = 4096;

// Boundaries of the filter's work items over a CSR (host side).
void filter_chunks(uint32_t n, const uint32_t* row_offset, std::vector<uint32_t>& bounds);

// build.hip: the undirected CSR of an edge list (host arrays src/dst of m
// edges, either orientation; duplicates and self-loops dropped), built on
// the device.  *row (n+1) and *col (*adj) are new device allocations.
// GPSM_EINVAL if an endpoint is >= n.
int build_csr_from_edges(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, hipStream_t s,
                         uint32_t** row, uint32_t** col, uint64_t* adj);

// Sets the adjacency bitmaps of the hubs: hub_bits (nhub * W words) must be
// zero; hubs[h] is the vertex of bitmap h; block_base[h] = 64-entry blocks
// of the adjacency of hubs before h (nhub + 1 entries, last = nblocks).
void launch_neighbour_keys(const DevGraph& g, hipStream_t s);
void launch_hub_bits(const DevGraph& g, const uint32_t* hubs, const uint32_t* block_base, uint32_t nblocks,
                     hipStream_t s);

// filter.hip ------------------------------------------------------------------
// Start-vertex candidates: C(slot) = {v : lo <= key[v] <= hi} (check_kernel).
void launch_init(const DevGraph& g, const InitOp* ops, uint32_t nops, uint32_t* bitmaps, hipStream_t s);
// One superstep: every op (one per query) explores or refines its C(u).
// stats (optional, STAT_COPIES x STAT_STRIDE u64, zeroed by the caller):
// copy c, entries 0..4 = candidates, adjacency entries read, key gathers,
// bitmap probes, bits set, added by the workgroups with blockIdx % copies
// == c (the sum over copies is the launch's total).
// The superstep visits only the (op, vertex range) pairs whose range holds a
// candidate of C(u): a first kernel lists them in items (nops x g.nfchunk
// capacity) and counts them in *nitems (zeroed by the caller), the second
// processes the list with at most max_wg workgroups, and at most
// FILTER_OVERSUB times as many as are resident.  rflags: the refine
// rounds' change flags (FilterOp::mark / need), zeroed by the caller once
// per batch; an op whose need flag is clear lists no range.
constexpr uint32_t STAT_COPIES = 32, STAT_STRIDE = 8;
constexpr uint32_t FILTER_OVERSUB = 8;
void launch_filter_step(const DevGraph& g, const FilterOp* ops, uint32_t nops, const uint32_t* nbr_slot,
                        const uint32_t* nbr_lo, const uint32_t* nbr_hi, uint32_t* bitmaps, uint32_t block,
                        uint32_t max_wg, uint2* items, uint32_t* nitems, uint32_t* rflags,
                        unsigned long long* stats, hipStream_t s);

// scan.hip ---------------------------------------------------------------------
// Bitmap rows are compacted in segments of SEG_WORDS words, one workgroup
// per (row, segment).
constexpr uint32_t SEG_WORDS = 2048;
inline uint32_t segs_of(uint32_t W) { return (W + SEG_WORDS - 1) / SEG_WORDS; }
// segc[r * segs_of(W) + i] = set bits of bitmap row r before segment i
// (exclusive prefix of the segments' popcounts); counts[r] = popcount of row r.
void launch_row_popcount(const DevGraph& g, const uint32_t* bitmaps, uint32_t nrows, uint32_t* segc,
                         uint32_t* counts, hipStream_t s);
// Compaction of bitmap rows to ascending id lists (row r at clist + coff[r])
// plus the rank index, from the segment counts of launch_row_popcount.
void launch_compact_rows(const DevGraph& g, const uint32_t* bitmaps, uint32_t nrows, const uint32_t* segc,
                         const uint64_t* coff, uint32_t* clist, uint32_t* bprefix, hipStream_t s);
// Exclusive scan of n u32 values; out may alias in.  total (device, u64) optional.
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, unsigned long long* total,
                           uint32_t* scratch, uint64_t scratch_words, hipStream_t s);
uint64_t scan_scratch_words(uint64_t n);
// Ascending ids of set bits of one bitmap; count written to *count (device
// u64).  scratch: compact_scratch_words(nbits) words, 8-byte aligned.
uint64_t compact_scratch_words(uint64_t nbits);
void launch_compact_bitmap(const uint32_t* bits, uint64_t nbits, uint32_t* ids, unsigned long long* count,
                           uint32_t* scratch, hipStream_t s);

// text.hip ---------------------------------------------------------------------
// Embedding lines "(j, k) ... \n" of IO::output.  R = text_rows_per_wg(q)
// rows per workgroup; wg_len / wg_off: ceil(rows / R) words.
constexpr uint32_t TEXT_STAGE_BYTES = 61440;
uint32_t text_fixed_length(uint32_t q);
uint32_t text_rows_per_wg(uint32_t q);
void launch_text_len(const uint32_t* tuples, uint64_t rows, uint32_t q, const uint8_t* keep, uint32_t* wg_len,
                     hipStream_t s);
void launch_text_write(const uint32_t* tuples, uint64_t rows, uint32_t q, const uint8_t* keep,
                       const uint32_t* wg_off, char* out, hipStream_t s);

// cet.hip ----------------------------------------------------------------------
// Per CetDesc: |CET| into desc.total (atomic); per-candidate counts into
// offsets (when per_cand != 0) at desc.off_base.
// chunk_base: device scratch of ndesc+1 words; grid: workgroups (persistent).
// stats (optional, STAT_COPIES x STAT_STRIDE u64, zeroed by the caller):
// entries 0..3 = candidates, adjacency entries examined, bitmap probes, list
// entries written (cet_write only), summed over the copies.
void launch_cet_count(const DevGraph& g, const uint32_t* bitmaps, const uint32_t* clist, CetDesc* descs,
                      uint32_t ndesc, uint32_t* offsets, int per_cand, uint32_t* chunk_base, uint32_t grid,
                      hipStream_t s, unsigned long long* stats = nullptr);
// Segmented exclusive scan of the per-candidate counts of every desc.
void launch_cet_offsets(CetDesc* descs, uint32_t ndesc, uint32_t* offsets, hipStream_t s);
// Hash terms of the last join levels' lists: for table term_desc[i] (terms
// [term_start[i], term_start[i+1]) of terms), tuple_term(term_qid, list[e]).
void launch_cet_terms(const uint32_t* lists, const CetDesc* descs, const uint32_t* term_desc,
                      const uint64_t* term_start, uint32_t nterm_desc, uint64_t nterms, unsigned long long* terms,
                      hipStream_t s);
// Writes the lists.
void launch_cet_write(const DevGraph& g, const uint32_t* bitmaps, const uint32_t* clist, const CetDesc* descs,
                      uint32_t ndesc, const uint32_t* offsets, uint32_t* lists, uint32_t* chunk_base, uint32_t grid,
                      hipStream_t s, unsigned long long* stats = nullptr);

// join.hip ---------------------------------------------------------------------
struct JoinArgs {
  const uint32_t* row;
  const uint32_t* col;
  const uint32_t* hub_of;    // DevGraph::hub_of / hub_bits
  const uint32_t* hub_bits;
  const uint32_t* bitmaps;
  const uint32_t* bprefix;
  const uint32_t* clist;
  const uint32_t* offsets;
  const uint32_t* lists;
  const unsigned long long* terms;  // hash terms of the last level's lists (CetDesc::term_base)
  const CetDesc* cets;
  const JoinPlan* plans;
  const JoinTask* tasks;
  uint32_t ntasks;
  uint32_t W, NB;
  uint32_t cap;       // rows per level buffer
  uint32_t qmax;      // widest query of the batch
  uint32_t* levels;   // per workgroup: (qmax-2) buffers of cap rows, column-major
  uint64_t level_stride;  // words per workgroup
  unsigned int* task_counter;
  JoinOut* out;
  // optional tuple output (query-id order), emit_cap tuples of q words
  uint32_t* emit;
  unsigned long long* emit_ctr;
  uint64_t emit_cap;
  // 0 or the candidate rows a query may examine before it is abandoned
  unsigned long long budget;
  // nonzero: account algorithmic bytes in JoinOut::bytes (a slower variant)
  uint32_t count_bytes;
};
// One persistent join launch over tasks [0, ntasks) of a.tasks; cyclic: the
// plans of those tasks have back edges (join_cyclic_kernel), else none
// (join_tree_kernel, which holds no back-edge tests).
void launch_join(const JoinArgs& a, uint32_t nwg, bool cyclic, hipStream_t s);
uint64_t join_level_words(uint32_t qmax, uint32_t cap);
// Join workgroups (of the timed variant of the tree or cyclic kernel) that
// are resident on one CU at once for a batch whose widest query has qmax
// vertices (LDS and registers); a persistent grid gains nothing from more.
uint32_t join_resident_per_cu(uint32_t qmax, bool cyclic);

}  // namespace gpsm
