#include "IO.h"

#include <sys/stat.h>

#include <algorithm>
#include <cstring>

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// A number as scanf's %d reads it: leading white space, an optional sign, at
// least one digit; s advances past it.  (One sscanf per line took most of
// the 4 s that a 135 MB graph file took to read.)
bool read_int(const char*& s, int& out) {
  while (is_space(*s)) ++s;
  bool neg = false;
  if (*s == '-' || *s == '+') neg = *s++ == '-';
  if (*s < '0' || *s > '9') return false;
  long long v = 0;
  for (; *s >= '0' && *s <= '9'; ++s)
    if (v < (1LL << 40)) v = v * 10 + (*s - '0');
  out = (int)(neg ? -v : v);
  return true;
}

// v in decimal at p (as "%u" prints it); returns the end.
char* put_u32(char* p, unsigned v) {
  char t[10];
  int k = 0;
  do {
    t[k++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (k) *p++ = t[--k];
  return p;
}

// A graph's edges in file order, added to its vertices' in/out lists when
// the graph ends, each list reserved at its final size first: one
// allocation per list instead of growth by doubling (about 2.3 s of 2.7 s
// for 8M edges).  The lists hold what Graph::addEdge in file order gives.
struct EdgeBuf {
  std::vector<int> a, b, lb;
};

void add_edges(Graph* g, const EdgeBuf& e) {
  const size_t n = g->vertices.size(), m = e.a.size();
  std::vector<unsigned> nout(n, 0), nin(n, 0);
  for (size_t i = 0; i < m; ++i) {
    ++nout[e.a[i]];
    ++nin[e.b[i]];
  }
  for (size_t v = 0; v < n; ++v) {
    g->vertices[v].out.reserve(nout[v]);
    g->vertices[v].in.reserve(nin[v]);
  }
  for (size_t i = 0; i < m; ++i) g->addEdge(e.a[i], e.b[i], e.lb[i]);
}

}  // namespace

IO::IO(const std::string& query, const std::string& data, const std::string& file) {
  qfp = std::fopen(query.c_str(), "r");
  dfp = std::fopen(data.c_str(), "r");
  ofp = std::fopen(file.c_str(), "w+");
  // 1 MiB stdio buffers for the graph streams (fewer read calls)
  if (qfp) std::setvbuf(qfp, nullptr, _IOFBF, 1 << 20);
  if (dfp) std::setvbuf(dfp, nullptr, _IOFBF, 1 << 20);
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
  EdgeBuf edges;
  char buf[512];
  while (std::fgets(buf, sizeof buf, fp)) {
    const char* s = buf;
    while (*s == ' ' || *s == '\t') ++s;
    if (*s == '\n' || *s == '\r' || *s == 0) continue;
    if (*s == 't') {
      if (g) {  // next graph starts: back to the start of this line
        std::fseek(fp, -(long)std::strlen(buf), SEEK_CUR);
        add_edges(g, edges);
        return g;
      }
      // "t # id" (scanf's "t # %d": white space optional around '#')
      int id = 0;
      ++s;
      while (is_space(*s)) ++s;
      if (*s++ != '#' || !read_int(s, id)) break;
      if (id == -1) return nullptr;
      g = new Graph;
      // "n m Lv Le" header line: its counts only size the arrays
      if (!std::fgets(buf, sizeof buf, fp)) break;
      // (at most what the file's size can hold: a "v" or "e" line takes at
      // least 6 bytes, so a damaged header cannot ask for more; 2^20 on a
      // stream of unknown size, which then grows as lines arrive)
      const char* h = buf;
      int hn, hm;
      if (read_int(h, hn) && read_int(h, hm) && hn > 0 && hm > 0) {
        long long cap = 1LL << 20;
        struct stat st;
        if (fstat(fileno(fp), &st) == 0 && S_ISREG(st.st_mode))
          cap = std::min<long long>(1LL << 26, st.st_size / 6 + 1);
        g->vertices.reserve((size_t)std::min<long long>(hn, cap));
        const size_t em = (size_t)std::min<long long>(hm, cap);
        edges.a.reserve(em);
        edges.b.reserve(em);
        edges.lb.reserve(em);
      }
    } else if (*s == 'v') {
      int id, lb;
      ++s;
      if (!g || !read_int(s, id) || !read_int(s, lb) || id != (int)g->vertices.size()) break;
      g->addVertex(lb);
    } else if (*s == 'e') {
      // "e from to [label]": the label is optional (0 when absent)
      int a, b, lb = 0;
      ++s;
      if (!g || !read_int(s, a) || !read_int(s, b)) break;
      if (!read_int(s, lb)) lb = 0;
      if (a < 0 || b < 0 || a >= (int)g->vertices.size() || b >= (int)g->vertices.size()) break;
      edges.a.push_back(a);
      edges.b.push_back(b);
      edges.lb.push_back(lb);
    } else {
      break;
    }
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
  // the same text as fprintf("(%u, %u) ") per pair, formatted by hand into
  // a 1 MiB buffer (a row takes at most 25 bytes per column plus a newline)
  std::vector<char> buf(1 << 20);
  const size_t row_max = (size_t)cols * 25 + 1;
  if (row_max > buf.size()) buf.resize(row_max);
  size_t n = 0;
  for (unsigned i = 0; i < rows; ++i) {
    if (n + row_max > buf.size()) {
      if (std::fwrite(buf.data(), 1, n, ofp) != n) return false;
      n = 0;
    }
    const unsigned* ans = final_result + (size_t)i * cols;
    char* p = buf.data() + n;
    for (unsigned j = 0; j < cols; ++j) {
      *p++ = '(';
      p = put_u32(p, j);
      *p++ = ',';
      *p++ = ' ';
      p = put_u32(p, ans[id_map[j]]);
      *p++ = ')';
      *p++ = ' ';
    }
    *p++ = '\n';
    n = (size_t)(p - buf.data());
  }
  if (n && std::fwrite(buf.data(), 1, n, ofp) != n) return false;
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
  std::setvbuf(fp, nullptr, _IOFBF, 1 << 20);
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
