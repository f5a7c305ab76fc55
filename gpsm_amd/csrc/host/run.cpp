// GpSM command-line driver: GpSM <data> <query> [output=ans.txt] [dev=0]
// (the reference's main/run.cpp:20-118).  Matches every query against every
// data graph of the data file and writes the embeddings; "match used" covers
// matching and output of all queries of a data graph, excluding reading and
// CSR construction, as in the reference.  With --count, only embedding
// counts and set hashes are printed (no tuples are materialised).  With
// --directed, embeddings are written only when every query arc maps to a
// data arc with the same edge label (the reference's IO::verify,
// io/IO.cpp:146-180, done on the device); without it the semantics are
// undirected and edge labels are ignored.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <string>
#include <vector>

#include "Graph.h"
#include "IO.h"
#include "Match.h"
#include "gpsm_gen.h"

static long now_ms() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec * 1000 + tv.tv_usec / 1000;
}

// The first 8 bytes of a binary CSR cache (gpsm_io_write_csr).
static bool is_cache(const std::string& path) {
  char m[8] = {0};
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  const bool ok = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, "GPSMCSR\x01", 8) == 0;
  std::fclose(f);
  return ok;
}

// GpSM --write-cache <data.g> <out.csr>: the first data graph of a text
// file as a binary CSR cache (no device needed).
static int write_cache(const std::string& data, const std::string& out) {
  IO io(data, data, "/dev/null");
  if (!io.ok()) return 1;
  Graph* g = nullptr;
  if (!io.input(g)) {
    std::fprintf(stderr, "data file: malformed or missing terminator\n");
    return 1;
  }
  const int rc = gpsm_io_write_csr(out.c_str(), g->vertex_num, g->undirected_row_offset,
                                   g->undirected_column_index, g->vertex_value);
  std::fprintf(stderr, "%s: %u vertices, %u adjacency entries -> %s (%s)\n", data.c_str(), g->vertex_num,
               g->eSize(), out.c_str(), gpsm_strerror(rc));
  delete g;
  return rc == GPSM_OK ? 0 : 1;
}

int main(int argc, const char* argv[]) {
  bool count_only = false, directed = false;
  std::vector<std::string> args;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--count") == 0) count_only = true;
    else if (std::strcmp(argv[i], "--directed") == 0) directed = true;
    else args.push_back(argv[i]);
  }
  if (args.size() == 3 && args[0] == "--write-cache") return write_cache(args[1], args[2]);
  if (args.size() < 2 || args.size() > 4) {
    std::fprintf(stderr,
                 "usage: GpSM [--count | --directed] <data> <query> [output=ans.txt] [dev=0]\n"
                 "       GpSM --write-cache <data.g> <out.csr>\n"
                 "<data> is a text graph stream or a binary CSR cache (one data graph).\n");
    return 2;
  }
  const std::string output = args.size() >= 3 ? args[2] : "ans.txt";
  const int dev = args.size() == 4 ? std::atoi(args[3].c_str()) : 0;
  const bool cache = is_cache(args[0]);
  if (directed && (count_only || cache)) {
    std::fprintf(stderr, "--directed needs the text data format and tuples (not --count)\n");
    return 2;
  }
  Match::initGPU(dev);
  if (!Match::gpuReady()) return 1;
  // a cache is read below; the text reader then only serves the queries
  IO io(args[1], cache ? args[1] : args[0], output);
  if (!io.ok()) return 1;
  io.set_row_writer(&Match::write_rows);  // embedding lines formatted on the device
  std::vector<Graph*> queries;
  if (!io.input(queries)) {
    std::fprintf(stderr, "query file: malformed or missing terminator\n");
    return 1;
  }
  std::fprintf(stderr, "input ok!\n");
  int rc = 0;
  auto run_data = [&](Graph* data) {
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
      else if (directed && rows && m.verify_directed(result, rows) != GPSM_OK) {
        std::fprintf(stderr, "verify: %s\n", gpsm_strerror(m.last_status()));
        rc = 1;
        rows = 0;
      }
      if (!io.output(result, rows, cols, id_map, queries[i], data)) rc = 1;
      io.flush();
      delete[] result;
    }
    std::fprintf(stderr, "match used: %ld ms\n", now_ms() - start);
    Match::forget(data);  // its device copy is keyed by the address about to be freed
    delete data;
  };
  if (cache) {
    gpsm_csr c;
    const int mrc = gpsm_io_map_csr(args[0].c_str(), &c);
    if (mrc != GPSM_OK) {
      std::fprintf(stderr, "data cache: %s\n", gpsm_strerror(mrc));
      return 1;
    }
    Graph* data = new Graph;
    data->setCSR(c.n, c.row_offset, c.col, c.labels);
    gpsm_io_unmap_csr(&c);
    io.next_data();
    run_data(data);
  } else {
    // the next data graph is read and its CSR built on another thread while
    // the device matches the current one (the reference reads, builds and
    // matches them one after another, main/run.cpp:60-98)
    auto read = [&io]() { return io.read_data(); };
    std::future<Graph*> next = std::async(std::launch::async, read);
    for (;;) {
      Graph* data = next.get();
      if (!data) break;
      next = std::async(std::launch::async, read);
      io.next_data();
      run_data(data);
    }
    if (io.error()) {
      std::fprintf(stderr, "data file: malformed or missing terminator\n");
      rc = 1;
    }
  }
  for (Graph* q : queries) delete q;
  io.flush();
  Match::releaseGPU();
  return rc;
}
