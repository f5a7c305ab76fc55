// Internal state behind the opaque C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "gpsm.h"
#include "kernels/launch.h"
#include "plan.h"

namespace gpsm {

// A growable device buffer, reused across calls of one context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t nbytes);
  void release();
};

// Pinned host staging (hipHostMalloc): copies from and to it are queued on
// the stream like kernels, where a pageable copy waits on the host.
struct HostBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t nbytes);
  void release();
};

// Pinned staging for a batch's uploads: take() hands out fresh regions (in
// blocks of at least 1 MB, kept for later batches) until reset(), which the
// batch calls at its start once the stream is idle, so no queued copy still
// reads a region that is handed out again.
struct PinnedArena {
  std::vector<HostBuf> blocks;
  size_t cur = 0, used = 0;
  void* take(size_t nbytes);
  void reset() { cur = used = 0; }
  void release() {
    for (auto& b : blocks) b.release();
    blocks.clear();
    reset();
  }
};

enum BufId {
  B_SCRATCH,
  B_BITMAPS,
  B_INIT,
  B_FOPS,
  B_NBR_SLOT,
  B_NBR_LO,
  B_NBR_HI,
  B_COUNTS,
  B_COFF,
  B_CLIST,
  B_BPREFIX,
  B_FDESC,
  B_JDESC,
  B_CHUNKS,
  B_OFFSETS,
  B_LISTS,
  B_PLANS,
  B_TASKS,
  B_JOUT,
  B_CTR,
  B_LEVELS,
  B_EMIT,
  B_MLP,
  B_FSTATS,
  B_TERMS,       // hash terms of the last join level's lists (u64)
  B_TERM_DESC,   // tables that have them
  B_TERM_START,  // their ranges
  B_SEGC,        // popcounts of the candidate bitmaps' segments
  B_TEXT_IN,     // embedding writer: tuples of a chunk
  B_TEXT_KEEP,   // ... their keep flags
  B_TEXT_WG,     // ... scan scratch, per-workgroup lengths and offsets
  B_TEXT_OUT,    // ... the chunk's text
  B_FITEMS,      // filter superstep: (op, vertex range) pairs with candidates
  B_FNITEMS,     // ... their number, one counter per superstep
  B_CDESC,       // candidate-edge tables of reversed plan edges (counted again)
  B_RFLAGS,      // refine rounds that removed a candidate, 4 per query
  B_COUNT_
};

// Launch parameters chosen by the autotuner (defaults until it runs).
// The defaults are the values the autotuner settles on for the 100 walk
// queries on the 1M-vertex Chung-Lu graph on an MI355X: 8 join workgroups
// per CU (out of 4/8/12/16) and 32 tasks per workgroup (out of 8-64) in
// every run; with 4 and 8 the join of that batch took 15.6 ms instead of 10.
struct Tuning {
  uint32_t filter_block = 128;   // workgroup size of the filter supersteps
  uint32_t join_cap = 4096;      // rows per join level buffer
  uint32_t join_wg_per_cu = 8;   // persistent join workgroups per CU
  uint32_t task_min = 256;       // minimum first-table rows per join task
  uint32_t tasks_per_wg = 32;    // join tasks per persistent workgroup over a batch
};

struct BatchHooks {
  void* user = nullptr;
  // after phase A; bitmaps are device pointers, counts host.  Nonzero stops.
  int (*after_filter)(void* user, gpsm_ctx* c, const uint32_t* bitmaps, const uint32_t* counts, uint32_t nslots,
                      uint32_t W) = nullptr;
  // after the forward |CET| counts (host descs, totals filled).  Nonzero stops.
  int (*after_cet_count)(void* user, gpsm_ctx* c, const CetDesc* fdesc, uint32_t nfdesc) = nullptr;
  void (*after_join)(void* user, gpsm_ctx* c, const uint32_t* unused, uint32_t nq) = nullptr;
};

int run_batch(gpsm_ctx* c, const gpsm_graph* g, uint32_t nq, const gpsm_query* qs, const gpsm_opts* opts,
              gpsm_result* out, const BatchHooks* hooks);

// scan.hip
void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);

}  // namespace gpsm

namespace gpsm {

// Per-launch timing on the context stream while profiling is on.
struct Prof {
  int mode = 0;  // 0 off, 1 HIP-event timing, 2 timing + the filter's byte counters, 3 the join only
  struct Pending {
    int kind;
    hipEvent_t a, b;
    double bytes;
    int stats_slot;  // >= 0: bytes come from the device counters of this slot
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  struct Acc {
    uint64_t launches = 0;
    double ms = 0, bytes = 0;
  };
  Acc acc[12];
  hipEvent_t get();
};

// K_WALL_*: host wall-clock time of a batch's phases, synchronisations and
// host work included (filter; compaction + forward |CET|; join orders;
// candidate-edge tables + join + results)
enum ProfKind {
  K_FILTER = 0,
  K_JOIN,
  K_CET_COUNT,
  K_CET_WRITE,
  K_COMPACT,
  K_INIT,
  K_OTHER,
  K_WALL_FILTER,
  K_WALL_CET,
  K_WALL_PLAN,
  K_WALL_JOIN,
  K_KINDS_
};
extern const char* const kProfNames[K_KINDS_];

}  // namespace gpsm

struct gpsm_ctx {
  gpsm::Prof prof;
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // the cyclic plans' join runs on stream2 beside the tree plans' (forked
  // from and joined back into stream with ev_fork / ev_join), so that its
  // workgroups fill the CUs the tree launch's last tasks leave idle; both
  // streams at the default priority
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  char name[256] = {0};
  gpsm::DevBuf bufs[gpsm::B_COUNT_];
  gpsm::Tuning tune;
  gpsm_mlp* planner = nullptr;
  std::vector<uint32_t> h_counts;
  gpsm::HostBuf h_stage[3];  // pinned readbacks of a batch: counts, |CET| descriptors, join results
  gpsm::PinnedArena arena;   // pinned sources of a batch's uploads
  // Instantiated hipGraphs of a batch's two captured phases (filter;
  // candidate-edge tables), kept across batches: a batch whose launches
  // have the argument list a graph was captured with (gsig) relaunches it
  // without a capture (graph_replays); otherwise a new capture updates it in
  // place when its topology is the same (hipGraphExecUpdate, graph_updates)
  // and is instantiated afresh when it is not (graph_instantiations).
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<uint64_t> gsig[2];
  // this device's copy of the allocation-free scan's static scratch array
  // (hipGetSymbolAddress at gpsm_init)
  uint32_t* noalloc_scratch = nullptr;
  uint64_t graph_instantiations = 0, graph_updates = 0, graph_replays = 0;
  // hub adjacency bitmaps of later uploads: vertices of degree > the hub
  // degree, starting at hub_min_degree and doubled until they fit
  // hub_max_bytes (gpsm_init: 48 GiB or a quarter of the device's memory,
  // whichever is less)
  uint32_t hub_min_degree = 64;
  uint64_t hub_max_bytes = 8ull << 30;
};

namespace gpsm {

// Directed, edge-labelled out-adjacency of a resident graph, rows sorted by
// (target, label): only the final-set check reads it.
struct DirectedGraph {
  uint32_t* row = nullptr;    // [n+1]
  uint32_t* col = nullptr;    // [arcs]
  uint32_t* label = nullptr;  // [arcs]
  uint64_t arcs = 0;
  void release() {
    hipFree(row);
    hipFree(col);
    hipFree(label);
    *this = DirectedGraph();
  }
};

}  // namespace gpsm

struct gpsm_graph {
  gpsm_ctx* ctx = nullptr;
  gpsm::DevGraph dg;
  gpsm::GraphStats stats;
  uint64_t bytes = 0;
  gpsm::DirectedGraph directed;
};
