#include "Graph.h"

#include <algorithm>
#include <cstring>

Graph::~Graph() { free_csr(); }

void Graph::free_csr() {
  delete[] vertex_value;
  delete[] undirected_row_offset;
  delete[] undirected_column_index;
  delete[] inverse_label;
  delete[] inverse_num;
  vertex_value = undirected_row_offset = undirected_column_index = inverse_label = inverse_num = nullptr;
  label_num = 0;
}

void Graph::addVertex(int label) { vertices.emplace_back(label); }

void Graph::addEdge(int from, int to, int elb) {
  vertices[from].out.emplace_back(to, elb);
  vertices[to].in.emplace_back(from, elb);
}

void Graph::transformToCSR() {
  free_csr();
  vertex_num = (unsigned)vertices.size();
  vertex_value = new unsigned[vertex_num ? vertex_num : 1];
  undirected_row_offset = new unsigned[vertex_num + 1];
  undirected_row_offset[0] = 0;
  std::vector<unsigned> cols;
  std::vector<unsigned> rows;
  for (unsigned v = 0; v < vertex_num; ++v) {
    vertex_value[v] = (unsigned)vertices[v].label;
    rows.clear();
    for (const Neighbor& nb : vertices[v].in)
      if ((unsigned)nb.vid != v) rows.push_back((unsigned)nb.vid);
    for (const Neighbor& nb : vertices[v].out)
      if ((unsigned)nb.vid != v) rows.push_back((unsigned)nb.vid);
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    cols.insert(cols.end(), rows.begin(), rows.end());
    undirected_row_offset[v + 1] = (unsigned)cols.size();
  }
  undirected_column_index = new unsigned[cols.empty() ? 1 : cols.size()];
  std::copy(cols.begin(), cols.end(), undirected_column_index);
  build_inverse_labels();
}

void Graph::setCSR(unsigned n, const unsigned* row_offset, const unsigned* col, const unsigned* labels) {
  free_csr();
  vertex_num = n;
  vertex_value = new unsigned[n ? n : 1];
  undirected_row_offset = new unsigned[n + 1];
  undirected_column_index = new unsigned[row_offset[n] ? row_offset[n] : 1];
  std::memcpy(vertex_value, labels, sizeof(unsigned) * n);
  std::memcpy(undirected_row_offset, row_offset, sizeof(unsigned) * (n + 1));
  std::memcpy(undirected_column_index, col, sizeof(unsigned) * row_offset[n]);
  build_inverse_labels();
}

void Graph::build_inverse_labels() {
  // inverse label list: labels ascending with their vertex counts
  std::vector<unsigned> lv(vertex_value, vertex_value + vertex_num);
  std::sort(lv.begin(), lv.end());
  std::vector<unsigned> lab, num;
  for (size_t i = 0; i < lv.size(); ++i) {
    if (i == 0 || lv[i] != lv[i - 1]) {
      lab.push_back(lv[i]);
      num.push_back(0);
    }
    num.back()++;
  }
  label_num = (unsigned)lab.size();
  delete[] inverse_label;
  delete[] inverse_num;
  inverse_label = new unsigned[label_num ? label_num : 1];
  inverse_num = new unsigned[label_num ? label_num : 1];
  std::copy(lab.begin(), lab.end(), inverse_label);
  std::copy(num.begin(), num.end(), inverse_num);
}

bool Graph::isEdgeContained(int from, int to, int label) const {
  for (const Neighbor& nb : vertices[from].out)
    if (nb.vid == to && nb.elb == label) return true;
  return false;
}
