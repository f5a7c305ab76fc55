// GpSM command-line driver: GpSM <data> <query> [output=ans.txt] [dev=0]
// (the reference's main/run.cpp:20-118).  Matches every query against every
// data graph of the data file and writes the embeddings; "match used" covers
// matching and output of all queries of a data graph, excluding reading and
// CSR construction, as in the reference.  With --count, only embedding
// counts and set hashes are printed (no tuples are materialised).
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Graph.h"
#include "IO.h"
#include "Match.h"

static long now_ms() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec * 1000 + tv.tv_usec / 1000;
}

int main(int argc, const char* argv[]) {
  bool count_only = false;
  std::vector<std::string> args;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--count") == 0) count_only = true;
    else args.push_back(argv[i]);
  }
  if (args.size() < 2 || args.size() > 4) {
    std::fprintf(stderr, "usage: GpSM [--count] <data> <query> [output=ans.txt] [dev=0]\n");
    return 2;
  }
  const std::string output = args.size() >= 3 ? args[2] : "ans.txt";
  const int dev = args.size() == 4 ? std::atoi(args[3].c_str()) : 0;
  if (!Match::initGPU(dev)) return 1;
  IO io(args[1], args[0], output);
  if (!io.ok()) return 1;
  std::vector<Graph*> queries;
  if (!io.input(queries)) {
    std::fprintf(stderr, "query file: malformed or missing terminator\n");
    return 1;
  }
  std::fprintf(stderr, "input ok!\n");
  Graph* data = nullptr;
  int rc = 0;
  while (io.input(data)) {
    long start = now_ms();
    for (size_t i = 0; i < queries.size(); ++i) {
      Match m(queries[i], data);
      io.output((int)i);
      if (count_only) {
        uint64_t cnt, hash;
        if (m.count(cnt, hash) != GPSM_OK) rc = 1;
        std::printf("query %zu: %llu embeddings, hash %016llx\n", i, (unsigned long long)cnt,
                    (unsigned long long)hash);
        continue;
      }
      unsigned* result = nullptr;
      unsigned rows = 0, cols = 0;
      int* id_map = nullptr;
      m.match(io, result, rows, cols, id_map);
      if (m.last_status() != GPSM_OK) rc = 1;
      io.output(result, rows, cols, id_map, queries[i], data);
      io.flush();
      delete[] result;
    }
    std::fprintf(stderr, "match used: %ld ms\n", now_ms() - start);
    delete data;
  }
  if (io.error()) {
    std::fprintf(stderr, "data file: malformed or missing terminator\n");
    rc = 1;
  }
  for (Graph* q : queries) delete q;
  io.flush();
  Match::releaseGPU();
  return rc;
}
