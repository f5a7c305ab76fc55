// Host-only check of the C++ boundary classes (Graph, IO) written the way the
// reference's callers use them: the CSR fields of Graph are plain unsigned
// arrays (graph/Graph.h:99-119), and IO has the reference's four output
// overloads (io/IO.h:24-28).  Prints what it read and writes the output file.
#include <cstdio>

#include "Graph.h"
#include "IO.h"

// Takes the arrays as pointers, as gpsm_graph_upload and the reference's
// Match::copyGraphToGPU (match/Match.cu:128-143) do.
static unsigned sum_csr(unsigned n, const unsigned* row, const unsigned* col, const unsigned* lab) {
  unsigned s = 0;
  for (unsigned v = 0; v < n; ++v) {
    s += lab[v];
    for (unsigned e = row[v]; e < row[v + 1]; ++e) s += col[e] * (v + 1);
  }
  return s;
}

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  IO io(argv[1], argv[2], argv[3]);
  if (!io.ok()) return 3;
  Graph* data = nullptr;
  if (!io.input(data)) return 4;
  unsigned* labels = data->vertex_value;
  unsigned* row = data->undirected_row_offset;
  unsigned* col = data->undirected_column_index;
  std::printf("n %u adj %u labels %u checksum %u\n", data->vSize(), data->eSize(), data->label_num,
              sum_csr(data->vertex_num, row, col, labels));
  for (unsigned i = 0; i < data->label_num; ++i) std::printf("label %u x%u\n", data->inverse_label[i], data->inverse_num[i]);
  unsigned num = 0, off = 0;
  data->getNeighbor(0, num, off);
  std::printf("deg0 %u off0 %u deg_last %d\n", num, off, data->getVertexDegree(data->vertex_num - 1));
  // a copy of the CSR through setCSR (the binary-cache path)
  Graph copy;
  copy.setCSR(data->vertex_num, row, col, labels);
  std::printf("copy checksum %u\n", sum_csr(copy.vertex_num, copy.undirected_row_offset,
                                             copy.undirected_column_index, copy.vertex_value));
  io.output(0);
  int m[3] = {4, 0, 2};
  io.output(m, 3);
  unsigned rows[6] = {1, 2, 3, 3, 2, 1};
  int id_map[3] = {2, 0, 1};
  io.output(rows, 2, 3, id_map, nullptr, data);
  io.output();
  io.flush();
  delete data;
  return 0;
}
