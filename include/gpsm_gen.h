/*
 * gpsm_gen.h — seeded synthetic inputs and text I/O for the matcher.
 *
 * The reference reads its graphs from gSpan-like text files
 * (io/IO.cpp:46-100) and ships no data; these generators produce the
 * workloads the benchmarks are quoted on, bit-identically on every x86-64
 * host (SplitMix64 streams, no std::*_distribution).  All functions are host
 * code; buffers they return are released with gpsm_gen_free.
 */
#ifndef GPSM_GEN_H
#define GPSM_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* An undirected CSR data graph: rows sorted ascending, no self-loops, no
 * duplicate edges (the form Graph::transformToCSR builds, graph/Graph.cpp:30-104,
 * plus deduplication). */
typedef struct gpsm_csr {
  uint32_t n;
  uint64_t adj;          /* row_offset[n] = 2 * undirected edges */
  uint32_t* row_offset;  /* [n+1] */
  uint32_t* col;         /* [adj] */
  uint32_t* labels;      /* [n], >= 1 */
} gpsm_csr;

/* A list of queries stored back to back. */
typedef struct gpsm_qset {
  uint32_t nq;
  uint32_t* v_off;   /* [nq+1] offsets into labels */
  uint32_t* labels;  /* [v_off[nq]] */
  uint32_t* e_off;   /* [nq+1] offsets into src/dst */
  uint32_t* src;     /* [e_off[nq]] */
  uint32_t* dst;     /* [e_off[nq]] */
} gpsm_qset;

/* Builds a CSR from an undirected edge list (either orientation, duplicates
 * and self-loops dropped).  labels may be NULL (all 1). */
int gpsm_gen_csr_from_edges(uint32_t n, uint64_t m, const uint32_t* a, const uint32_t* b,
                            const uint32_t* labels, gpsm_csr* out);

/* Erdos-Renyi G(n, m): m distinct undirected edges chosen uniformly. */
int gpsm_gen_er(uint32_t n, uint64_t m, uint64_t seed, gpsm_csr* out);
/* Chung-Lu with power-law weights w_i ~ (i + i0)^(-1/(gamma-1)) capped at
 * sqrt(sum w); m distinct edges, endpoints drawn independently by weight. */
int gpsm_gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint64_t seed, gpsm_csr* out);
/* R-MAT on 2^scale vertices with quadrant probabilities a, b, c (d = 1-a-b-c);
 * m distinct undirected edges after dropping self-loops and duplicates. */
int gpsm_gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed,
                  gpsm_csr* out);
/* Labels 1..L: zipf_s == 0 draws uniformly, otherwise P(k) ~ k^-zipf_s. */
int gpsm_gen_labels(gpsm_csr* g, uint32_t L, double zipf_s, uint64_t seed);

/* Query kinds */
#define GPSM_Q_WALK_TREE 0    /* random walk, edges = first-visit edges (a spanning tree) */
#define GPSM_Q_WALK_INDUCED 1 /* random walk, edges = all data edges among the walk vertices */
#define GPSM_Q_WALK_HALF 2    /* even slots tree, odd slots induced */
#define GPSM_Q_CYCLE 3        /* a k-cycle found in the data graph (labels copied) */
#define GPSM_Q_CLIQUE 4       /* a k-clique found in the data graph (labels copied) */

/* nq queries with k drawn uniformly from [kmin, kmax] per slot.  Every query
 * is an induced-or-tree pattern of an actual data subgraph, so it has at least
 * one embedding.  Feasibility rule: a plain label+edge backtracking search
 * (vertices in walk order) must finish within node_budget examined adjacency
 * entries; Knuth's estimator (64 seeded probes) of that cost screens draws
 * first; the decision is an exact upper bound on the embedding count (the
 * label- and degree-preserving homomorphisms of the pattern's BFS spanning
 * tree) against node_budget.  A rejected query makes the slot draw again,
 * with a fresh size uniform in [kmin, kmax] (sizes are never shrunk, so a
 * slot's query follows the size distribution conditioned on the rule);
 * after 3200 rejected draws of one slot the call fails (GPSM_EINVAL).
 * node_budget 0 disables the rule.  Slots are independent streams (seed,
 * slot), so the set is the same for any thread count. */
int gpsm_gen_queries(const gpsm_csr* g, uint32_t nq, uint32_t kmin, uint32_t kmax, int kind,
                     uint64_t seed, uint64_t node_budget, gpsm_qset* out);

void gpsm_gen_free_csr(gpsm_csr* g);
void gpsm_gen_free_qset(gpsm_qset* s);

/* Text format of io/IO.cpp:46-100: "t # id", "n m Lv Le", "v id label",
 * "e from to label", terminated by "t # -1".  write: one graph per call
 * (edges written once, from < to); the terminator is appended by
 * gpsm_io_write_end.  read: next graph of the stream; returns 1 on a graph,
 * 0 at the terminator, GPSM_EFORMAT (-8) on malformed or unterminated input
 * (the reference loops forever on a missing terminator). */
int gpsm_io_write_graph(const char* path, int append, int gid, uint32_t n, const uint32_t* labels,
                        uint64_t m, const uint32_t* src, const uint32_t* dst);
int gpsm_io_write_end(const char* path);
/* Reads a whole graph stream with the driver's reader (IO::input,
 * io/IO.cpp:46-100): *ngraphs receives the number of graphs and *nverts /
 * *nedges (optional) the totals.  Returns GPSM_EFORMAT on a malformed line or
 * a missing "t # -1" terminator, GPSM_EINVAL if the file cannot be opened. */
int gpsm_io_count_graphs(const char* path, uint32_t* ngraphs, uint64_t* nverts, uint64_t* nedges);

/* Binary CSR cache of one data graph (no counterpart in the reference, whose
 * text reader io/IO.cpp:46-100 parses token by token): a 64-byte header
 * ("GPSMCSR" + version 1, n, adj) followed by row_offset[n+1], col[adj] and
 * labels[n] as little-endian u32.  map: read-only mmap of such a file; out's
 * arrays point into the mapping until gpsm_io_unmap_csr.  The header, the
 * file length and the offsets (monotone, ending at adj) are checked
 * (GPSM_EFORMAT otherwise); rows are validated by gpsm_graph_upload. */
int gpsm_io_write_csr(const char* path, uint32_t n, const uint32_t* row_offset, const uint32_t* col,
                      const uint32_t* labels);
int gpsm_io_map_csr(const char* path, gpsm_csr* out);
void gpsm_io_unmap_csr(gpsm_csr* g);

#ifdef __cplusplus
}
#endif
#endif /* GPSM_GEN_H */
