/*
 * gpsm.h — C ABI of the MI355X (gfx950) GpSM filter-and-join subgraph matcher.
 *
 * Semantics: undirected, vertex-labelled, non-induced, injective matching of a
 * small query graph (at most GPSM_MAX_QV vertices, connected, at least one edge)
 * against a data graph that stays resident in HBM.  Vertex ids start at 0,
 * labels at 1.  The result of a query is its embedding set: every injective map
 * f from query vertices to data vertices that preserves labels and maps every
 * query edge to a data edge.  It is reported as an exact 64-bit count plus an
 * order-independent 64-bit set hash (gpsm_tuple_hash summed mod 2^64), and,
 * on request, as the sorted list of int32 tuples in query-vertex-id order.
 *
 * Conventions: every entry point returns GPSM_OK (0) or a negative GPSM_E*
 * code and never exits the process.  Input arrays are caller-owned and only
 * read during the call.  Result buffers (gpsm_result.tuples) are library-owned
 * and released by gpsm_result_free.  One gpsm_ctx per host thread; a context
 * owns one HIP stream and all device scratch it uses.
 *
 * Each declaration names the reference interface it replaces
 * (bookug/GpSM, paths relative to the reference root).
 */
#ifndef GPSM_H
#define GPSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPSM_MAX_QV 64 /* maximum query vertices handled by the device join */

/* error codes */
#define GPSM_OK 0
#define GPSM_EINVAL (-1)      /* malformed argument or graph */
#define GPSM_ENOMEM (-2)      /* device or host allocation failed */
#define GPSM_EDEVICE (-3)     /* HIP runtime error */
#define GPSM_ETOOBIG (-4)     /* query larger than GPSM_MAX_QV */
#define GPSM_EOVERFLOW (-5)   /* tuple materialisation exceeded tuple_cap */
#define GPSM_ENODEV (-6)      /* no HIP device */
#define GPSM_EDISCONN (-7)    /* query not connected or without edges */
#define GPSM_EFORMAT (-8)     /* text reader: malformed or unterminated input */
#define GPSM_EBUDGET (-9)     /* the join exceeded gpsm_opts.work_budget (result incomplete) */
#define GPSM_ESCRATCH (-10)   /* input larger than the fixed scratch of an allocation-free entry point */

typedef struct gpsm_ctx gpsm_ctx;
typedef struct gpsm_graph gpsm_graph;

/* A query graph as the reference's text format lists it: vertex labels and the
 * "e src dst label" lines in file order (io/IO.cpp:46-100).  Edge direction
 * and edge labels are ignored (undirected semantics). */
typedef struct gpsm_query {
  uint32_t n;              /* vertices, ids 0..n-1 */
  uint32_t m;              /* edges in file order */
  const uint32_t* labels;  /* [n], >= 1 */
  const uint32_t* src;     /* [m] */
  const uint32_t* dst;     /* [m] */
} gpsm_query;

/* join-order selection */
#define GPSM_ORDER_HEURISTIC 0 /* the reference's greedy min-|CET| order (match/Match.cu:1599-1661) */
#define GPSM_ORDER_LEARNED 1   /* candidate orders ranked by MLP-predicted join cost (queries with little
                                  predicted join work keep the heuristic order); needs gpsm_set_planner */
#define GPSM_ORDER_EXPLICIT 2  /* caller-given vertex order in gpsm_opts.explicit_order */

typedef struct gpsm_opts {
  int32_t order_mode;             /* GPSM_ORDER_* */
  const uint32_t* explicit_order; /* EXPLICIT: query ids, every prefix of length >= 2 connected;
                                     gpsm_match reads [n], gpsm_match_batch reads query b's order
                                     at explicit_order + b*GPSM_MAX_QV */
  uint32_t want_tuples;           /* gpsm_match only: materialise the sorted tuple list */
  uint64_t tuple_cap;             /* maximum tuples to materialise (GPSM_EOVERFLOW beyond) */
  uint32_t use_graph;             /* gpsm_match_batch: capture the phases in hipGraphs */
  uint32_t block;                 /* 0 = autotuned; else workgroup size of the filter/collect kernels */
  uint32_t join_block;            /* 0 = autotuned; else workgroup size of the join kernel */
  uint32_t join_cap;              /* 0 = autotuned; else rows per DFS level buffer of the join */
  uint64_t work_budget;           /* 0 = none; else a query whose join examines more candidate
                                     rows than this is abandoned with status GPSM_EBUDGET */
} gpsm_opts;

typedef struct gpsm_result {
  int32_t status;                   /* GPSM_OK or error for this query */
  uint32_t q;                       /* query vertices */
  uint64_t count;                   /* embeddings */
  uint64_t hash;                    /* set hash */
  uint32_t order[GPSM_MAX_QV];      /* join position -> query vertex id */
  uint64_t level_rows[GPSM_MAX_QV]; /* rows with k+1 bound vertices after step k (k>=1) */
  uint32_t cand[GPSM_MAX_QV];       /* |C(u)| after refinement, by query vertex id */
  uint32_t* tuples;                 /* library-owned, n_tuples x q, sorted */
  uint64_t n_tuples;
  double device_ms;                 /* device time of this call (batch: whole batch) */
  uint64_t join_work;               /* first-table rows + candidates the join examined at every
                                       later level (the work opts.work_budget bounds) */
  uint64_t level_work[GPSM_MAX_QV]; /* [1]: first-table rows; [k >= 2]: candidates examined
                                       to produce the rows of step k (sum = join_work) */
} gpsm_result;

/* --- lifecycle ---------------------------------------------------------- */

/* Replaces Match::initGPU (match/Match.cu:15-70): selects the device, creates
 * the context's stream and warms the runtime.  Error code instead of exit. */
int gpsm_init(int device, gpsm_ctx** out);
void gpsm_shutdown(gpsm_ctx* ctx);
const char* gpsm_strerror(int code);
/* Name of the HIP device the context runs on (empty when none). */
const char* gpsm_device_name(const gpsm_ctx* ctx);

/* Replaces Match::copyGraphToGPU (match/Match.cu:128-143), which re-uploaded
 * the CSR on every query: the graph is uploaded once and stays resident.
 * Input is the undirected CSR of Graph::transformToCSR (graph/Graph.cpp:30-104):
 * row_offset[n+1], col[row_offset[n]] with each row sorted ascending, no
 * self-loops and no duplicates, labels[n] >= 1. */
int gpsm_graph_upload(gpsm_ctx* ctx, uint32_t n, const uint32_t* row_offset,
                      const uint32_t* col, const uint32_t* labels, gpsm_graph** out);
/* The same resident graph built on the device from an edge list of m edges
 * (src[i], dst[i]) in either orientation, duplicates and self-loops dropped
 * (no counterpart in the reference, which builds the CSR on the host,
 * graph/Graph.cpp:30-104; its note at graph/Graph.cpp:26-28 asks for a GPU
 * build): one radix sort of 64-bit (row, column) keys and one unique pass.
 * GPSM_EINVAL if an endpoint is >= n or 2m >= 2^32. */
int gpsm_graph_upload_edges(gpsm_ctx* ctx, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                            const uint32_t* labels, gpsm_graph** out);
/* Vertices and adjacency entries (2 x undirected edges) of a resident graph. */
int gpsm_graph_info(const gpsm_graph* g, uint32_t* n, uint64_t* adj);
/* Copies a resident graph's CSR to host buffers row_offset[n+1], col[adj]
 * (e.g. to write a binary cache of a graph built on the device). */
int gpsm_graph_csr(const gpsm_graph* g, uint32_t* row_offset, uint32_t* col);
int gpsm_graph_free(gpsm_graph* g);
/* Bytes of device memory the resident graph occupies. */
uint64_t gpsm_graph_device_bytes(const gpsm_graph* g);
/* Hub adjacency bitmaps for graphs uploaded after this call (no counterpart
 * in the reference, whose join scans candidate lists linearly): every vertex
 * of degree > D gets an n-bit adjacency bitmap in device memory, so a
 * join-mode edge test touching it is one bit instead of a binary search.
 * D starts at min_degree (default 64) and doubles until the bitmaps fit
 * max_bytes (default: 48 GiB or a quarter of the device's memory, whichever
 * is less; 0 = no hubs).  Results do not depend on it. */
int gpsm_set_hub_policy(gpsm_ctx* ctx, uint32_t min_degree, uint64_t max_bytes);
/* The hubs of an uploaded graph: their number and the degree D above which
 * a vertex is one (0 when there are none). */
int gpsm_graph_hubs(const gpsm_graph* g, uint32_t* nhub, uint32_t* hub_degree);

/* --- embedding text ------------------------------------------------------ */

/* Replaces the line loop of IO::output (io/IO.cpp:207-213) with device
 * formatting: for every row i of tuples (n_rows x q, column j = query vertex
 * j) with keep == NULL or keep[i] != 0, the line "(0, t0) (1, t1) ... \n"
 * ("(%u, %u) " per column, then a newline).  Rows are formatted in chunks
 * (bounded device and host staging).  gpsm_format_tuples copies the text to
 * out[0..cap) and returns GPSM_EOVERFLOW when it is longer (out holds the
 * first cap bytes); *bytes (may be NULL) receives the full length. */
int gpsm_format_tuples(gpsm_ctx* ctx, const uint32_t* tuples, uint64_t n_rows, uint32_t q, const uint8_t* keep,
                       char* out, uint64_t cap, uint64_t* bytes);
/* The same text written to a file descriptor chunk by chunk (GPSM_EINVAL
 * when a write fails). */
int gpsm_write_tuples(gpsm_ctx* ctx, int fd, const uint32_t* tuples, uint64_t n_rows, uint32_t q,
                      const uint8_t* keep, uint64_t* bytes);

/* --- directed, edge-labelled final check -------------------------------- */

/* Attaches the directed, edge-labelled out-adjacency of the data graph to a
 * resident graph (Graph::vertices[v].out, graph/Graph.h:91-158, as read by
 * io/IO.cpp:77-86): out_row[n+1], out_col/out_label[out_row[n]], each row
 * sorted by (target, label); duplicate arcs and self-loops allowed.  Replaces
 * an earlier attachment.  GPSM_EINVAL on unsorted rows or targets >= n. */
int gpsm_graph_set_directed(gpsm_graph* g, const uint32_t* out_row, const uint32_t* out_col,
                            const uint32_t* out_label);
/* Replaces IO::verify (io/IO.cpp:146-180) with Graph::isEdgeContained
 * (graph/Graph.cpp:107-118) on the device: tuple i (tuples[i*q .. i*q+q),
 * column j = query vertex j) is kept iff its q vertices are distinct and for
 * every query arc e the data graph has the arc tuple[src[e]] -> tuple[dst[e]]
 * with label elabel[e].  keep[n_tuples] receives 1/0, *kept (may be NULL)
 * the number kept.  Needs gpsm_graph_set_directed first (GPSM_EINVAL). */
int gpsm_verify_directed(gpsm_ctx* ctx, const gpsm_graph* g, uint32_t q, uint32_t m, const uint32_t* src,
                         const uint32_t* dst, const uint32_t* elabel, const uint32_t* tuples, uint64_t n_tuples,
                         uint8_t* keep, uint64_t* kept);

/* --- matching ----------------------------------------------------------- */

/* Replaces Match::match (match/Match.cu:1710-1966) together with the
 * injectivity half of IO::verify (io/IO.cpp:146-160): one query, launched
 * step by step on the context's stream. */
int gpsm_match(gpsm_ctx* ctx, const gpsm_graph* g, const gpsm_query* q,
               const gpsm_opts* opts, gpsm_result* out);

/* Batched form of gpsm_match with no reference counterpart (the reference
 * loops over queries in main/run.cpp:74-93): nq queries of uneven size are
 * filtered in lock-step supersteps and joined by one persistent launch fed by
 * a device work queue.  out[nq]. */
int gpsm_match_batch(gpsm_ctx* ctx, const gpsm_graph* g, uint32_t nq,
                     const gpsm_query* qs, const gpsm_opts* opts, gpsm_result* out);

void gpsm_result_free(gpsm_result* r);

/* Autotuner (no reference counterpart; the reference hard-codes block sizes
 * 1024/256/32, match/Match.cu:455-1618): times gpsm_match_batch on the given
 * sample for every candidate filter workgroup size (64/128/256), join level
 * buffer (2048/4096/8192 rows), persistent join workgroups per CU (2/4/6/8,
 * never more than are resident) and join tasks per workgroup (8/16/32/64),
 * one parameter at a time (best of three runs each; the filter workgroup
 * size timed by the batch's wall time, the join settings by the join
 * kernel's HIP-event time), and keeps
 * the fastest in the context.  The caller's profiling statistics are left
 * as they were.  Results of every trial are checked to agree.  *best_ms
 * receives the join kernel's time (ms) with the chosen setting.  Later
 * calls with opts->block / join_cap == 0 use it. */
int gpsm_autotune(gpsm_ctx* ctx, const gpsm_graph* g, uint32_t nq, const gpsm_query* qs,
                  const gpsm_opts* opts, double* best_ms);
/* The tuned values: filter workgroup size, join buffer rows, join workgroups
 * per CU, join tasks per workgroup (any pointer may be NULL). */
int gpsm_tuning(const gpsm_ctx* ctx, uint32_t* filter_block, uint32_t* join_cap, uint32_t* join_wg_per_cu,
                uint32_t* tasks_per_wg);
/* Sets the same values, e.g. a tuning saved from an earlier gpsm_autotune
 * (filter_block 64/128/256, join_cap 2048..2^20, join_wg_per_cu 1..16,
 * tasks_per_wg 1..256; GPSM_EINVAL otherwise).  A batch launches at most as
 * many join workgroups per CU as are resident at once (about 7 for queries
 * of up to 16 vertices), whatever join_wg_per_cu asks.  No counterpart in
 * the reference, whose launch shapes are constants. */
int gpsm_set_tuning(gpsm_ctx* ctx, uint32_t filter_block, uint32_t join_cap, uint32_t join_wg_per_cu,
                    uint32_t tasks_per_wg);

/* --- pipeline stages (parity entry points) ------------------------------ */

/* InitializeCandidateVertices + RefineCandidateVertices
 * (match/Match.cu:661-719, 848-890): candidate bitmaps after the filter.
 * bitmaps: [q->n][words] with words = (n+31)/32, bit v of row u set iff v in C(u);
 * counts[q->n] = |C(u)|. */
int gpsm_filter(gpsm_ctx* ctx, const gpsm_graph* g, const gpsm_query* q,
                uint32_t* bitmaps, uint32_t* counts);

/* FindCandidateEdges (match/Match.cu:1082-1139) for the query edge list in
 * the reference's order (match/Match.cu:1735-1754, from < to).
 * edge_from/edge_to[e]: the edges; cet_size[e] = |CET(from,to)|.
 * If lists != NULL (capacity list_cap entries) the forward tables are written
 * back to back in the reference layout [C(u) ids | offsets | neighbour lists]
 * (match/Match.cu:1119-1121), table e starting at lists[table_off[e]]. */
int gpsm_candidate_edges(gpsm_ctx* ctx, const gpsm_graph* g, const gpsm_query* q,
                         uint32_t* edge_from, uint32_t* edge_to, uint32_t* cet_size,
                         uint32_t* lists, uint64_t list_cap, uint64_t* table_off);

/* thrust::exclusive_scan as used by kernel_collect / kernel_join
 * (match/Match.cu:504-508) and gpsm::scan::prefixSum (util/prefix_scan.cu:410-427):
 * out[i] = sum(in[0..i)), returns the total in *total. Host buffers. */
int gpsm_scan_u32(gpsm_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* total);

/* The scan of gpsm_scan_u32 on device buffers without allocating anything:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  d_in and d_out are device pointers of n u32 values
 * (d_out may alias d_in), written by work the caller has completed; *total
 * (host) receives the u64 sum.  The block sums live in a fixed device array
 * of GPSM_SCAN_NOALLOC_SCRATCH_BYTES owned by the library (one per device,
 * held until the process ends), which bounds n at GPSM_SCAN_NOALLOC_MAX
 * (GPSM_ESCRATCH beyond it).  Runs on the context's stream and waits for it.
 * Calls that overlap on one device share that array and race. */
#define GPSM_SCAN_NOALLOC_SCRATCH_BYTES (256u * 1024u)
/* 2 words hold the total, 2 words per block of 4096 values hold its sum:
 * (65536 - 2) / 2 = 32767 blocks */
#define GPSM_SCAN_NOALLOC_MAX (32767ull * 4096ull)
int gpsm_scan_u32_noalloc(gpsm_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n, uint64_t* total);

/* kernel_collect (match/Match.cu:488-526) / gpsm::scan::compact
 * (util/compact_scan.cu:52-70) on a bitmap: ids of set bits, ascending. */
int gpsm_compact_bitmap(gpsm_ctx* ctx, const uint32_t* bits, uint64_t nbits,
                        uint32_t* ids, uint64_t* count);

/* --- learned planner (no reference counterpart) -------------------------- */

#define GPSM_FEAT 32   /* features per candidate join step */
#define GPSM_HIDDEN 128

/* Per-step features for planning: for the state "vertices in prefix placed",
 * for each candidate next edge/vertex.  Computed on the host from the filter
 * statistics the device returns (|C(u)|, |CET(e)|, label frequencies, degrees). */
typedef struct gpsm_mlp gpsm_mlp;

/* MLP 32 -> 128 -> 128 -> 1 with ReLU, BF16 weights, FP32 accumulation,
 * trained with Adam on the device. */
int gpsm_mlp_create(gpsm_ctx* ctx, uint64_t seed, gpsm_mlp** out);
void gpsm_mlp_free(gpsm_mlp* m);
/* Parameter count and (de)serialisation of the FP32 master weights in the
 * order W1[128x32], b1[128], W2[128x128], b2[128], W3[128], b3[1]. */
uint64_t gpsm_mlp_num_params(void);
int gpsm_mlp_get_params(gpsm_mlp* m, float* out);
int gpsm_mlp_set_params(gpsm_mlp* m, const float* in);
/* Fixed input standardisation applied in FP32 before the BF16 rounding of
 * the inputs: x' = (x - mu) * inv_sd (NULL restores the identity).  Not
 * trained; part of the planner's saved state. */
int gpsm_mlp_set_input_norm(gpsm_mlp* m, const float* mu, const float* inv_sd);
int gpsm_mlp_get_input_norm(gpsm_mlp* m, float* mu, float* inv_sd);
/* Forward: x[B x 32] (host, FP32, standardised, rounded to BF16 on the device) -> y[B]. */
int gpsm_mlp_forward(gpsm_mlp* m, const float* x, uint32_t B, float* y);
/* Loss 0.5*mean((y-t)^2) and its gradient w.r.t. every parameter (same
 * order as gpsm_mlp_get_params) for one mini-batch. */
int gpsm_mlp_grad(gpsm_mlp* m, const float* x, const float* t, uint32_t B, float* loss, float* grad);
/* One Adam step on a mini-batch; *loss receives the pre-step loss. */
int gpsm_mlp_train_step(gpsm_mlp* m, const float* x, const float* t, uint32_t B, float lr, float* loss);
/* Installs (or, with NULL, removes) the network used by GPSM_ORDER_LEARNED. */
int gpsm_set_planner(gpsm_ctx* ctx, gpsm_mlp* m);

/* Training rows for the planner (host only, no device work).  For a vertex
 * order of q (every prefix connected) and the rows measured at each level
 * (level_rows[k] = rows with k+1 bound vertices, as gpsm_result reports),
 * writes for every level k = 2..q->n-1 the GPSM_FEAT features of extending
 * the prefix order[0..k) by order[k] (log1p of the measured rows of level k-1
 * among them) into feat[(k-2)*GPSM_FEAT ..] and log1p(level_rows[k]) into
 * target[k-2].  cand[u] = |C(u)|, cet_size[e] = |CET| of the query edges in
 * the order gpsm_candidate_edges reports them.  Returns the number of rows. */
int gpsm_plan_features(const gpsm_graph* g, const gpsm_query* q, const uint32_t* cand,
                       const uint64_t* cet_size, const uint32_t* order, const uint64_t* level_rows,
                       float* feat, float* target);

/* --- profiling (no reference counterpart; the reference times whole
 * queries with a millisecond wall clock, util/Util.cpp:21-27) ------------- */

typedef struct gpsm_kstat {
  char name[32];      /* kernel family */
  uint64_t launches;
  double total_ms;    /* sum of HIP-event durations around each launch */
  double bytes;       /* algorithmic bytes of all launches (0 if not accounted) */
} gpsm_kstat;

/* Per-launch statistics on the context's stream: mode 0 off, 1 HIP-event
 * timing of every launch, 2 timing plus device counters of the work the
 * filter and join kernels did (candidates, adjacency entries, key reads,
 * probes, list entries, searches; a few atomics per workgroup), from which
 * their algorithmic bytes are computed (the join runs a variant that counts
 * them), 3 HIP-event timing of the join launches only (no events between
 * the other launches).  Setting a mode clears the statistics.  Launches captured in a
 * hipGraph are not timed.  Entries named "wall:<phase>" are host wall-clock
 * times of a batch's phases, synchronisations included. */
int gpsm_profile_enable(gpsm_ctx* ctx, int mode);
/* Copies up to max kernel-family statistics; *n receives the number. */
int gpsm_profile_read(gpsm_ctx* ctx, gpsm_kstat* out, uint32_t max, uint32_t* n);
/* With opts.use_graph, a batch's two captured phases (the filter; the
 * candidate-edge tables) run through executable hipGraphs that the context
 * keeps; the join is launched directly.  A later batch whose launches have the
 * same arguments (buffer addresses, sizes, counts: e.g. the same queries
 * again, or queries of the same shapes) relaunches them without capturing;
 * otherwise its capture updates them in place when its topology is the same
 * and instantiates them again when it is not.  These counters (since
 * gpsm_init, per phase launch) say which happened; any pointer may be NULL.
 * No counterpart in the reference (no graphs). */
int gpsm_graph_cache_stats(const gpsm_ctx* ctx, uint64_t* instantiations, uint64_t* updates, uint64_t* replays);

/* --- set hash ------------------------------------------------------------- */

/* Hash of one tuple t[0..q) in query-id order: the product mod 2^64 over j of
 * ( fmix64(((uint64)j << 32 | t[j]) ^ 0x9E3779B97F4A7C15) | 1 ),
 * fmix64 = MurmurHash3's 64-bit finaliser (every factor is odd, so the
 * product is an odd 64-bit value).  The set hash is the sum of tuple hashes
 * mod 2^64, so it is independent of enumeration order; the product is
 * independent of the order in which a join binds the vertices. */
uint64_t gpsm_tuple_hash(const uint32_t* t, uint32_t q);

#ifdef __cplusplus
}
#endif
#endif /* GPSM_H */
