/*
 * oracle.h — CPU oracle of the GpSM filter-and-join path.  TEST
 * INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU-baseline leg of
 * bench.py load it as the checker; the product library never does.
 *
 *   O2  host restatement of Match::match (bookug/GpSM match/Match.cu), step by
 *       step: spanning-tree filter order, check/explore, 4+1 refine rounds,
 *       candidate-edge tables, greedy join.  Candidate sets are byte arrays as
 *       in the reference (match/Match.cu:1764).
 *   O1  independent backtracking enumerator (label + edge + injectivity only),
 *       sharing no code with O2.
 *   scan / compact / MLP references for the device kernels.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   O1, O2 final sets  closed forms (K_k in K_n, C_k in C_n, stars, paths,
 *                      bipartite odd cycles), brute force over all injective
 *                      maps on graphs <= 8 vertices, O1 == O2 on random inputs.
 *   O2 filter stages   soundness (every embedding image stays a candidate),
 *                      refine monotonicity, hand-worked fixtures in
 *                      tests/golden/, CET invariants.  The reference GPU code
 *                      cannot run here, so the exact candidate bytes after each
 *                      kernel are pinned only by these checks and the
 *                      hand-worked fixture ("parity partially pinned").
 *   scan / compact     numpy cumsum / nonzero.
 *   MLP                finite differences in FP64 and torch autograd.
 */
#ifndef GPSM_ORACLE_H
#define GPSM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXQ 64
#define ORC_MAXE 2048

typedef struct orc_graph {
  uint32_t n;
  const uint32_t* row_offset; /* [n+1] */
  const uint32_t* col;        /* sorted rows, no duplicates, no self-loops */
  const uint32_t* labels;     /* [n] */
} orc_graph;

typedef struct orc_query {
  uint32_t n, m;
  const uint32_t* labels;
  const uint32_t* src; /* file order */
  const uint32_t* dst;
} orc_query;

/* O2 status */
#define ORC_OK 0
#define ORC_EMPTY_INIT 1    /* a candidate set empty in InitializeCandidateVertices */
#define ORC_EMPTY_REFINE 2  /* a candidate set empty in RefineCandidateVertices */
#define ORC_EMPTY_JOIN 3    /* a join step left no rows */
#define ORC_ERR (-1)

typedef struct orc_o2_out {
  int32_t status;
  uint64_t count, hash;
  uint32_t filter_order[ORC_MAXQ];
  uint32_t n_edges;
  uint32_t edge_from[ORC_MAXE], edge_to[ORC_MAXE], cet_size[ORC_MAXE];
  uint32_t cand[ORC_MAXQ];        /* |C(u)| after refine */
  uint32_t order[ORC_MAXQ];       /* join position -> query id */
  uint64_t level_rows[ORC_MAXQ];  /* rows with k+1 columns after all steps at that width */
  uint32_t n_steps;               /* edge steps (first table = step 0) */
  uint32_t step_edge[ORC_MAXE];
  uint8_t step_join[ORC_MAXE];    /* 1 = join mode (both ends bound), 0 = expand/table */
  uint64_t step_rows[ORC_MAXE];   /* R after each edge step, as the reference's tables */
  uint32_t* tuples;               /* caller buffer [tuple_cap * q] or NULL */
  uint64_t tuple_cap, n_tuples;   /* tuples written sorted; n_tuples may exceed cap (then truncated) */
  uint64_t join_work;             /* first-table rows + list entries visited by expand steps */
  uint64_t level_work[ORC_MAXQ];  /* [1] first-table rows, [k] entries visited to reach width k+1 */
} orc_o2_out;

/* mode 0: expand rows are kept injective (the build's semantics);
 * mode 1: rows are homomorphisms as in the reference, injectivity applied at
 *         the end as IO::verify does (io/IO.cpp:146-160).
 * order: NULL = the reference heuristic (with the expand-mode fix), else an
 *        explicit vertex order. */
int orc_o2_match(const orc_graph* g, const orc_query* q, int mode, const uint32_t* order,
                 orc_o2_out* out);

/* O2 (mode 0, heuristic order) over nq queries, OpenMP across queries
 * (threads <= 0: all cores): the all-cores CPU baseline.  counts/hashes
 * [nq]; returns 0 or the first query's error. */
int orc_o2_match_many(const orc_graph* g, const orc_query* qs, uint32_t nq, int threads, uint64_t* counts,
                      uint64_t* hashes);

/* Candidate byte sets after InitializeCandidateVertices (init, may be NULL)
 * and after RefineCandidateVertices (refined), each [q->n * g->n].
 * Returns the O2 status of the filter phase. */
int orc_o2_filter(const orc_graph* g, const orc_query* q, uint8_t* init, uint8_t* refined,
                  uint32_t* filter_order);

/* Forward candidate-edge tables in the reference layout
 * [C(u) ids | offsets (|C|+1) | lists], back to back; table_off[e] start of
 * table e.  Returns total words needed (written only if <= cap). */
int64_t orc_o2_cet(const orc_graph* g, const orc_query* q, uint32_t* buf, uint64_t cap,
                   uint64_t* table_off, uint32_t* edge_from, uint32_t* edge_to, uint32_t* cand_num,
                   uint32_t* n_edges);

/* O1: all injective label- and edge-preserving maps.  threads <= 0: all cores.
 * tuples (optional, cap rows) receive the sorted embedding list. */
int orc_o1_enumerate(const orc_graph* g, const orc_query* q, int threads, uint64_t* count,
                     uint64_t* hash, uint32_t* tuples, uint64_t cap, uint64_t* n_tuples);

/* Tuple hash: prod_j (fmix64((j<<32|t[j]) ^ 0x9E3779B97F4A7C15) | 1) mod 2^64;
 * the set hash is the sum of tuple hashes mod 2^64. */
uint64_t orc_tuple_hash(const uint32_t* t, uint32_t q);

/* Exclusive scan (thrust::exclusive_scan semantics); returns the total. */
uint64_t orc_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n);
/* Ascending ids of set bits of a bitmap (32-bit little-endian words). */
uint64_t orc_compact_bitmap(const uint32_t* bits, uint64_t nbits, uint32_t* ids);

/* MLP 32->128->128->1 (ReLU).  params in gpsm_mlp_get_params order.
 * bf16 != 0 rounds weights, inputs and hidden activations to BF16 (round to
 * nearest even) exactly where the device kernels do; accumulation in FP64
 * here, FP32 on the device. */
int orc_mlp_forward(const float* params, const float* x, uint32_t B, int bf16, float* y);
int orc_mlp_grad(const float* params, const float* x, const float* t, uint32_t B, int bf16,
                 double* loss, double* grad);

#ifdef __cplusplus
}
#endif
#endif
