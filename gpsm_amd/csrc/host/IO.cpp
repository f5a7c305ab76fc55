#include "IO.h"

#include <cstring>

IO::IO(const std::string& query, const std::string& data, const std::string& file) {
  qfp = std::fopen(query.c_str(), "r");
  dfp = std::fopen(data.c_str(), "r");
  ofp = std::fopen(file.c_str(), "w+");
  if (!qfp || !dfp || !ofp) std::fprintf(stderr, "IO: cannot open %s\n", !qfp ? query.c_str() : !dfp ? data.c_str() : file.c_str());
}

IO::~IO() {
  if (qfp) std::fclose(qfp);
  if (dfp) std::fclose(dfp);
  if (ofp) std::fclose(ofp);
}

Graph* IO::input(FILE* fp) {
  err = false;
  if (!fp) {
    err = true;
    return nullptr;
  }
  Graph* g = nullptr;
  char buf[512];
  long pos = std::ftell(fp);
  while (std::fgets(buf, sizeof buf, fp)) {
    char* s = buf;
    while (*s == ' ' || *s == '\t') ++s;
    if (*s == '\n' || *s == '\r' || *s == 0) {
      pos = std::ftell(fp);
      continue;
    }
    if (*s == 't') {
      if (g) {  // next graph starts: rewind to this line
        std::fseek(fp, pos, SEEK_SET);
        return g;
      }
      int id = 0;
      if (std::sscanf(s, "t # %d", &id) != 1) break;
      if (id == -1) return nullptr;
      g = new Graph;
      // "n m Lv Le" header line
      if (!std::fgets(buf, sizeof buf, fp)) break;
    } else if (*s == 'v') {
      int id, lb;
      if (!g || std::sscanf(s, "v %d %d", &id, &lb) != 2 || id != (int)g->vertices.size()) break;
      g->addVertex(lb);
    } else if (*s == 'e') {
      int a, b, lb = 0;
      if (!g || std::sscanf(s, "e %d %d %d", &a, &b, &lb) < 2) break;
      if (a < 0 || b < 0 || a >= (int)g->vertices.size() || b >= (int)g->vertices.size()) break;
      g->addEdge(a, b, lb);
    } else {
      break;
    }
    pos = std::ftell(fp);
  }
  // malformed line, or end of file before "t # -1"
  delete g;
  err = true;
  return nullptr;
}

bool IO::input(std::vector<Graph*>& query_list) {
  for (;;) {
    Graph* g = input(qfp);
    if (!g) return !err;
    g->transformToCSR();
    query_list.push_back(g);
  }
}

Graph* IO::read_data() {
  Graph* g = input(dfp);
  if (g) g->transformToCSR();
  return g;
}

bool IO::input(Graph*& data_graph) {
  data_graph = input(dfp);
  if (!data_graph) return false;
  ++data_id;
  data_graph->transformToCSR();
  return true;
}

bool IO::output(int qid) {
  std::fprintf(ofp, "query graph:%d    data graph:%d\n%s\n", qid, data_id, line.c_str());
  return true;
}

bool IO::output(const unsigned* final_result, unsigned rows, unsigned cols, const int* id_map, Graph*, Graph*) {
  bool identity = true;
  for (unsigned j = 0; j < cols; ++j) identity = identity && id_map[j] == (int)j;
  if (writer && identity && rows) {
    std::fflush(ofp);
    if (writer(fileno(ofp), final_result, rows, cols) != 0) return false;
    std::fprintf(ofp, "\n\n\n");
    return true;
  }
  for (unsigned i = 0; i < rows; ++i) {
    const unsigned* ans = final_result + (size_t)i * cols;
    for (unsigned j = 0; j < cols; ++j) std::fprintf(ofp, "(%u, %u) ", j, ans[id_map[j]]);
    std::fprintf(ofp, "\n");
  }
  std::fprintf(ofp, "\n\n\n");
  return true;
}

bool IO::output() {
  if (!ofp) return false;
  std::fprintf(ofp, "\n\n\n");
  return true;
}

bool IO::output(int* m, int size) {
  if (!ofp || (size > 0 && !m)) return false;
  for (int i = 0; i < size; ++i) std::fprintf(ofp, "(%d, %d) ", i, m[i]);
  std::fprintf(ofp, "\n");
  return true;
}

void IO::flush() {
  if (ofp) std::fflush(ofp);
}

#include "gpsm.h"
#include "gpsm_gen.h"

extern "C" int gpsm_io_count_graphs(const char* path, uint32_t* ngraphs, uint64_t* nverts, uint64_t* nedges) {
  if (!path || !ngraphs) return GPSM_EINVAL;
  FILE* fp = std::fopen(path, "r");
  if (!fp) return GPSM_EINVAL;
  IO io;
  uint32_t n = 0;
  uint64_t nv = 0, ne = 0;
  for (;;) {
    Graph* g = io.input(fp);
    if (!g) break;
    ++n;
    nv += g->vertices.size();
    for (const Vertex& v : g->vertices) ne += v.out.size();
    delete g;
  }
  std::fclose(fp);
  if (io.error()) return GPSM_EFORMAT;
  *ngraphs = n;
  if (nverts) *nverts = nv;
  if (nedges) *nedges = ne;
  return GPSM_OK;
}
