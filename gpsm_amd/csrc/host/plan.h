// Host-side planning of one query: the reference's decisions that run on the
// CPU in Match::match (query edge list, spanning-tree filter order, refine
// schedule, join order), restated for the batched device pipeline.
#pragma once
#include <cstdint>
#include <vector>

#include "gpsm.h"

namespace gpsm {

// Data-graph statistics planning needs: the inverse label list of
// Graph::transformToCSR (graph/Graph.cpp:67-101) and the sizes.
struct GraphStats {
  uint32_t n = 0;
  uint64_t adj = 0;
  std::vector<uint32_t> label_value;  // distinct labels, ascending
  std::vector<uint32_t> label_count;  // vertices per label
  // dense label id (1-based index into label_value) or 0 when absent
  uint32_t dense(uint32_t label) const;
  uint32_t count(uint32_t label) const;
};

// A query in the reference's representation (Graph of io/IO.cpp:46-100 plus
// Graph::transformToCSR): in/out lists in file order, sorted undirected CSR.
struct HostQuery {
  uint32_t n = 0;
  std::vector<uint32_t> lab;
  std::vector<std::vector<uint32_t>> in, out;
  std::vector<uint32_t> ro, ci;
  std::vector<uint32_t> ef, et;  // edge list of match/Match.cu:1735-1754 (from < to)
  uint32_t deg(uint32_t u) const { return ro[u + 1] - ro[u]; }
  // GPSM_OK, GPSM_EINVAL, GPSM_ETOOBIG or GPSM_EDISCONN
  int build(const gpsm_query* q);
};

struct FilterStep {
  uint32_t type;  // OP_EXPLORE / OP_REFINE
  uint32_t u;
  uint32_t round = 0;            // refine: round 0..3
  std::vector<uint32_t> v;       // neighbour query ids, in the reference's order
  std::vector<uint32_t> lo, hi;  // explore: key range of each neighbour
};

struct FilterSchedule {
  std::vector<uint32_t> filter_order;
  uint32_t start = 0, start_lo = 0, start_hi = 0;
  std::vector<FilterStep> steps;
};

// Key range [lo, hi] of "label of u and degree >= deg(u)".
void key_range(const HostQuery& q, const GraphStats& s, uint32_t u, uint32_t& lo, uint32_t& hi);

// generateSpanningTree (match/Match.cu:338-429), InitializeCandidateVertices
// (:661-719) and RefineCandidateVertices (:848-890) as a list of device steps.
void make_filter_schedule(const HostQuery& q, const GraphStats& s, FilterSchedule& out);

// A join order: positions -> query ids, and for every level k >= 2 the
// parent position (the expand edge) and the other earlier neighbours (the
// join-mode edges, checked as edge membership).
struct OrderPlan {
  std::vector<uint32_t> order;
  uint32_t first_edge = 0;
  std::vector<uint32_t> par_pos, par_edge;       // [q], valid for k >= 2
  std::vector<std::vector<uint32_t>> backs;      // [q] earlier positions, parent excluded
};

// JoinCandidateEdges (match/Match.cu:1599-1700): first edge = smallest
// forward |CET| (strict <, lowest index); afterwards every edge with both
// ends bound is joined before the next expand, which takes the smallest
// |CET| among edges with exactly one bound end (the reference also admits
// edges with none, a defect fixed here).
void heuristic_order(const HostQuery& q, const std::vector<uint64_t>& cet_size, uint64_t big, OrderPlan& out);

// Completes a given vertex order (every prefix of length >= 2 connected):
// the expand edge of a vertex is its smallest-|CET| edge to an earlier one.
// Returns false if the order is invalid.
bool order_from_vertices(const HostQuery& q, const std::vector<uint64_t>& cet_size, const uint32_t* order,
                         OrderPlan& out);

// Fixed-width planner features of extending a prefix by query edge e
// (parent bound, child unbound).  Defined in plan.cpp next to the order
// search that uses them.
struct PlanContext {
  const HostQuery* q;
  const GraphStats* s;
  const std::vector<uint32_t>* cand;       // |C(u)| per query vertex
  const std::vector<uint64_t>* cet_size;   // |CET| per query edge
};
void step_features(const PlanContext& c, const std::vector<uint8_t>& placed, uint32_t nplaced, uint32_t edge,
                   uint32_t child, double log_rows, float* f);
void first_features(const PlanContext& c, uint32_t edge, uint32_t from, float* f);

// predict(user, x, B, y) evaluates the network on B feature rows (0 = ok).
typedef int (*PredictFn)(void* user, const float* x, uint32_t B, float* y);

// Greedy search over predicted log cardinalities (first edge and every next
// vertex minimise the predicted rows of the next level), one decision at a
// time, so that a batch of queries shares one network evaluation per round.
struct LearnedState {
  const PlanContext* c = nullptr;
  bool done = false;
  int round = 0;
  std::vector<uint32_t> order;
  std::vector<uint8_t> placed;
  double lr = 0;                                          // log1p rows of the current level
  std::vector<std::pair<uint32_t, uint32_t>> first;       // round 0: (edge, first vertex)
  std::vector<std::pair<size_t, size_t>> span;            // round 0: feature rows per choice
  std::vector<uint32_t> child;                            // later rounds: candidate vertices
};
void learned_start(const PlanContext& c, LearnedState& s);
size_t learned_emit(LearnedState& s, std::vector<float>& feats);  // appends rows, returns their number
void learned_consume(LearnedState& s, const float* y);
bool learned_finish(const LearnedState& s, OrderPlan& out);

// Candidate vertex orders for the learned cost model to choose among:
// the heuristic order first, then, for a query with at most kAllOrders
// connected vertex orders (every prefix connected), all of them in
// lexicographic order; for a larger one, from each of the kFirst
// smallest-|CET| first edges in both orientations, a min-|CET| continuation
// and a most-constrained continuation (most bound neighbours, then min
// |CET|).  Duplicates are dropped.
constexpr size_t kAllOrders = 64;
void candidate_orders(const HostQuery& q, const std::vector<uint64_t>& cet_size, uint64_t big,
                      std::vector<std::vector<uint32_t>>& out);

// Predicted join cost of one vertex order, in units of one examined
// candidate: the rows of the first table plus the candidates every later
// level examines, each weighted by its edge-membership checks,
//   cost = R_1 + sum_{k=2}^{n-1} R_{k-1} * |CET(e_k)| / |C(parent_k)| * (1 + kBackWeight * b_k),
// with e_k the expand edge of position k, b_k its join-mode (back) edges
// and R_j the rows of level j:
// R_1 = |CET| of the first edge exactly, R_2 .. R_{n-2} chained through the
// network (each level's features take the previous level's predicted log
// rows).  The final level's examined candidates are counted; its rows, the
// embedding count, are the same for every order.  Evaluated one level per
// round so that all candidates of a batch share one network call per round.
// Join time of one back-edge check (a hub-bitmap bit or a binary search in
// the shorter of two adjacency lists) in units of one examined candidate,
// fitted on an MI355X to the join kernel's time over batches of the 100 walk
// queries on the 1M-vertex Chung-Lu graph, each query taking one of its
// candidate orders at random (gpsm_amd/calibrate_planner.py: 24 batches,
// 23.0-27.8 ms, largest error 3.1 ms; first-table rows and materialised rows
// got weight 0).
constexpr double kBackWeight = 3.9;

struct PlanCost {
  const PlanContext* c = nullptr;
  OrderPlan plan;
  std::vector<uint8_t> placed;
  uint32_t k = 0;
  double lr = 0;    // log1p R_{k-1}
  double rows = 0;  // R_{k-1}
  double cost = 0;
  bool done = false;
};
bool cost_start(const PlanContext& c, const uint32_t* order, PlanCost& s);  // false: invalid order
size_t cost_emit(PlanCost& s, std::vector<float>& feats);                    // 0 or 1 rows
void cost_consume(PlanCost& s, const float* y);

// A candidate other than the heuristic order is taken only when its
// predicted cost is below this fraction of the heuristic order's.
constexpr double kLearnedMargin = 0.9;

// Queries whose heuristic plan has a predicted cost (analytic_cost) below
// this many examined candidates keep the heuristic order without ranking
// (learned_plans' rank flags).
// On an MI355X the join examines about 3e8 candidates per millisecond
// (3.56e-9 ms each, fitted by gpsm_amd/calibrate_planner.py on the 100 walk
// queries of the 1M-vertex Chung-Lu graph), so such a query's whole join
// takes a few microseconds, less than ranking its candidate orders on the
// host.
constexpr double kLearnedMinWork = 1e6;

// The learned planner for a batch: greedy learned orders, then the cheapest
// of each query's candidates (candidate_orders plus the greedy order) by
// chained predicted cost.  out[b] is always a valid plan; chosen[b] (if
// given) is the index of the candidate taken, 0 = the heuristic order.
// Returns 0 or the first non-zero code of predict.
// rank (optional, one flag per query): queries with a zero flag skip the
// greedy search and the candidate orders and keep the heuristic order, whose
// chained cost is still evaluated (so that every query's cost comes from the
// same network).
int learned_plans(const std::vector<const PlanContext*>& qs, PredictFn predict, void* user, uint64_t big,
                  std::vector<OrderPlan>& out, std::vector<uint32_t>* chosen, std::vector<double>* cost = nullptr,
                  const std::vector<uint8_t>* rank = nullptr);

// The same cost of a complete plan with the analytic chain estimate in place
// of the network: each level's log rows = previous log rows + log(|CET| of
// the expand edge / |C(parent)|) + the log densities of its back edges
// (feature 27 of step_features).  Used to size the join's tasks.
double analytic_cost(const PlanContext& c, const OrderPlan& plan);

}  // namespace gpsm
