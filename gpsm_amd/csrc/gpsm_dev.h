// Device-side data layout shared by the HIP kernels and their host launchers.
//
// Data graph (resident in HBM for the life of a gpsm_graph):
//   row[n+1], col[A]   undirected CSR, rows ascending
//   key[n]             (dense label << KEY_DEG_BITS) | min(degree, KEY_DEG_MASK):
//                      one 4-byte gather answers "label == L and degree >= d"
//                      for d <= KEY_DEG_MASK as lo <= key <= hi.
// Per batch of queries (one arena, bump-allocated):
//   bitmaps[slot][W]   candidate set C(u) of every query vertex ("slot" =
//                      batch-wide vertex index), W = ceil(n/32) words
//   bprefix[slot][NB]  set bits before each 256-bit block (rank index)
//   clist[coff[slot]]  ascending ids of C(u)
//   per directed query edge x->w used by a join: off[|C(x)|+1] and
//                      list[|CET|], the candidate-edge table
#pragma once
#include <cstdint>

namespace gpsm {

constexpr uint32_t KEY_DEG_BITS = 20;
constexpr uint32_t KEY_DEG_MASK = (1u << KEY_DEG_BITS) - 1;
constexpr uint32_t MAX_LABELS = (1u << (32 - KEY_DEG_BITS)) - 1;
constexpr uint32_t MAXQ = 64;
constexpr uint32_t BLOCK_BITS = 256;  // rank-index block

inline uint32_t words_of(uint32_t n) { return (n + 31) / 32; }
inline uint32_t nblocks_of(uint32_t n) { return (words_of(n) + 7) / 8 + 1; }

enum FilterOpType : uint32_t { OP_EXPLORE = 1, OP_REFINE = 2 };

// One filter operation of one query in a superstep.
struct FilterOp {
  uint32_t type;
  uint32_t u_slot;   // C(u) iterated
  uint32_t nv;       // neighbour entries
  uint32_t nbr_off;  // into nbr_slot / nbr_lo / nbr_hi
  // Refine rounds (match/Match.cu:863-887) after one that removed nothing
  // would remove nothing either: the state is then a fixed point of the
  // round.  mark: flag word set when this op removes a candidate (~0u: none);
  // need: flag word of the query's previous round, which must be set for
  // this op to run (~0u: always runs).
  uint32_t mark;
  uint32_t need;
};

// Initial label/degree check of a query's start vertex.
struct InitOp {
  uint32_t slot;
  uint32_t lo, hi;  // key range
};

// One directed candidate-edge table x -> w.
struct CetDesc {
  uint32_t x_slot, w_slot;
  uint32_t ncand;      // |C(x)|
  uint32_t list_cap;   // words allocated for the list (offsets are clamped to it)
  uint64_t clist_off;  // C(x) ids in clist
  uint64_t off_base;   // offsets [ncand+1] (u32) in the offsets arena
  uint64_t list_base;  // lists in the list arena
  uint64_t total;      // |CET|
  // the last join level's table also gets the hash term of every list entry
  // (tuple_term(term_qid, list[e]) at term_base + e); ~0 for other tables
  uint64_t term_base;
  uint32_t term_qid;
  // key range of w (as the filter used it): an entry whose neighbour key is
  // outside it is not in C(w), so its bitmap probe is skipped; w_hi == 0
  // means no range is known (every entry is probed)
  uint32_t w_lo, w_hi;
  uint32_t pad;
};

// Join plan of one query: vertex order and, per level, how to extend.
struct JoinPlan {
  uint32_t q;
  uint32_t vbase;             // slot of query vertex 0
  uint32_t first_cet;         // CetDesc index of the first table (pi0 -> pi1)
  uint32_t nback_total;
  // injectivity checks of level k: positions c < k whose query vertex has
  // the label of position k's, other than the parent and the back
  // positions.  A candidate of another label cannot equal the vertex bound
  // there, and one adjacent to the vertex at c cannot equal it (no self-loops).
  uint64_t inj[MAXQ];
  uint32_t order[MAXQ];       // position -> query id
  uint32_t par[MAXQ];         // parent position of level k (k >= 2)
  uint32_t cet[MAXQ];         // CetDesc index used to extend level k (k >= 2)
  uint32_t back_off[MAXQ + 1];  // back positions of level k in backs[]
  uint8_t backs[MAXQ * MAXQ / 2];  // back positions (excluding the parent)
};

// A slice of the first table of one query: rows [lo, hi).
// Rows first, first + stride, ..., (count of them) of one query's first
// table.  A query's tasks interleave its rows, so the rows of hub
// candidates (the table's first rows on a power-law graph) are spread over
// all of its tasks instead of making the first task the longest.
struct JoinTask {
  uint32_t query;
  uint32_t stride;
  uint64_t first, count;
};

// Per-query accumulators written by the join.
struct JoinOut {
  unsigned long long count;
  unsigned long long hash;
  unsigned long long level_rows[MAXQ];
  unsigned long long bytes;  // algorithmic bytes the join read and wrote for this query
  unsigned long long work;   // candidate rows examined (checked against the work budget)
  unsigned long long aborted;
  unsigned long long level_work[MAXQ];  // [1]: first-table rows; [k]: candidates examined for level k
};

}  // namespace gpsm
