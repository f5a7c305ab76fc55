// Test driver for the learned planner's host logic (csrc/host/plan.cpp):
// candidate orders, the batched greedy search and the choice by chained
// predicted cost, run with synthetic predictors instead of the network so
// that a test can recompute every cost independently.  Built with g++ by
// tests/test_planner_host.py; no device is involved.
//
// stdin:  n m / n labels / m edges (a b) / n |C(u)| / m |CET| (per input edge)
//         / predictor ("zero", "cet" or "chain")
// stdout: "edges", "heur", "cand" (one line each), "greedy", "chosen", "plan"
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "plan.h"

using namespace gpsm;

static std::string mode;

// zero:  every level predicted empty (log1p rows 0)
// cet:   log rows = log1p |CET| of the expand edge
// chain: log rows = previous log rows + log1p |CET| - log1p |C(parent)|
static int predict(void*, const float* x, uint32_t B, float* y) {
  for (uint32_t i = 0; i < B; ++i) {
    const float* f = x + (size_t)i * GPSM_FEAT;
    if (mode == "zero") y[i] = 0.0f;
    else if (mode == "cet") y[i] = f[2];
    else y[i] = f[0] + f[2] - f[3];
  }
  return 0;
}

static void print(const char* tag, const std::vector<uint32_t>& v) {
  std::printf("%s", tag);
  for (uint32_t x : v) std::printf(" %u", x);
  std::printf("\n");
}

int main() {
  uint32_t n, m;
  if (std::scanf("%u %u", &n, &m) != 2) return 2;
  std::vector<uint32_t> lab(n), src(m), dst(m), cand(n);
  std::vector<unsigned long> cet_in(m);
  bool ok = true;
  for (auto& x : lab) ok &= std::scanf("%u", &x) == 1;
  for (uint32_t e = 0; e < m; ++e) ok &= std::scanf("%u %u", &src[e], &dst[e]) == 2;
  for (auto& x : cand) ok &= std::scanf("%u", &x) == 1;
  for (auto& x : cet_in) ok &= std::scanf("%lu", &x) == 1;
  if (!ok) return 2;
  char buf[32];
  if (std::scanf("%31s", buf) != 1) return 2;
  mode = buf;

  gpsm_query gq{n, m, lab.data(), src.data(), dst.data()};
  HostQuery q;
  if (q.build(&gq) != GPSM_OK) return 3;
  std::vector<uint64_t> cs(q.ef.size());
  for (size_t i = 0; i < q.ef.size(); ++i)
    for (uint32_t e = 0; e < m; ++e)
      if (std::min(src[e], dst[e]) == q.ef[i] && std::max(src[e], dst[e]) == q.et[i]) cs[i] = cet_in[e];
  std::printf("edges");
  for (size_t i = 0; i < q.ef.size(); ++i) std::printf(" %u:%u:%lu", q.ef[i], q.et[i], (unsigned long)cs[i]);
  std::printf("\n");

  GraphStats st;
  st.n = 100000;
  st.adj = 1600000;
  for (uint32_t l = 0; l < 8; ++l) {
    st.label_value.push_back(l);
    st.label_count.push_back(12500);
  }
  const uint64_t big = st.adj + 1;
  OrderPlan h;
  heuristic_order(q, cs, big, h);
  print("heur", h.order);
  std::vector<std::vector<uint32_t>> cands;
  candidate_orders(q, cs, big, cands);
  for (const auto& od : cands) print("cand", od);

  PlanContext pc{&q, &st, &cand, &cs};
  LearnedState ls;
  learned_start(pc, ls);
  std::vector<float> feats, y;
  while (!ls.done) {
    feats.clear();
    const size_t R = learned_emit(ls, feats);
    y.assign(R + 1, 0.0f);
    predict(nullptr, feats.data(), (uint32_t)R, y.data());
    learned_consume(ls, y.data());
  }
  print("greedy", ls.order);

  std::vector<const PlanContext*> qs{&pc};
  std::vector<OrderPlan> out;
  std::vector<uint32_t> chosen;
  if (learned_plans(qs, predict, nullptr, big, out, &chosen)) return 4;
  std::printf("chosen %u\n", chosen[0]);
  print("plan", out[0].order);
  return 0;
}
