#include "Graph.h"

#include <algorithm>

void Graph::addVertex(int label) { vertices.emplace_back(label); }

void Graph::addEdge(int from, int to, int elb) {
  vertices[from].out.emplace_back(to, elb);
  vertices[to].in.emplace_back(from, elb);
}

void Graph::transformToCSR() {
  vertex_num = (unsigned)vertices.size();
  vertex_value.resize(vertex_num);
  undirected_row_offset.assign(vertex_num + 1, 0);
  std::vector<unsigned> tmp;
  std::vector<unsigned> rows;
  undirected_column_index.clear();
  for (unsigned v = 0; v < vertex_num; ++v) {
    vertex_value[v] = (unsigned)vertices[v].label;
    rows.clear();
    for (const Neighbor& nb : vertices[v].in)
      if ((unsigned)nb.vid != v) rows.push_back((unsigned)nb.vid);
    for (const Neighbor& nb : vertices[v].out)
      if ((unsigned)nb.vid != v) rows.push_back((unsigned)nb.vid);
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    undirected_column_index.insert(undirected_column_index.end(), rows.begin(), rows.end());
    undirected_row_offset[v + 1] = (unsigned)undirected_column_index.size();
  }
  build_inverse_labels();
}

void Graph::setCSR(unsigned n, const unsigned* row_offset, const unsigned* col, const unsigned* labels) {
  vertex_num = n;
  vertex_value.assign(labels, labels + n);
  undirected_row_offset.assign(row_offset, row_offset + n + 1);
  undirected_column_index.assign(col, col + row_offset[n]);
  build_inverse_labels();
}

void Graph::build_inverse_labels() {
  // inverse label list: labels ascending with their vertex counts
  std::vector<unsigned> lv(vertex_value);
  std::sort(lv.begin(), lv.end());
  inverse_label.clear();
  inverse_num.clear();
  for (size_t i = 0; i < lv.size(); ++i) {
    if (i == 0 || lv[i] != lv[i - 1]) {
      inverse_label.push_back(lv[i]);
      inverse_num.push_back(0);
    }
    inverse_num.back()++;
  }
  label_num = (unsigned)inverse_label.size();
}

bool Graph::isEdgeContained(int from, int to, int label) const {
  for (const Neighbor& nb : vertices[from].out)
    if (nb.vid == to && nb.elb == label) return true;
  return false;
}
