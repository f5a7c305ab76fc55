// Host graph of the matcher's C++ driver.  Same entry points as the
// reference's Graph (graph/Graph.h:91-158): vertices with in/out neighbour
// lists as read from the text format, and the undirected CSR plus inverse
// label list built by transformToCSR.  As in the reference, the CSR and the
// inverse label list are plain arrays (new[]) owned by the Graph and freed by
// its destructor, so code written against the reference's fields (e.g.
// passing data->vertex_value as an unsigned*) compiles unchanged.  Difference:
// rows are deduplicated (A->B plus B->A is one undirected edge).
#pragma once
#include <cstdint>
#include <vector>

struct Neighbor {
  int vid;
  int elb;
  Neighbor(int v = -1, int e = -1) : vid(v), elb(e) {}
};

struct Vertex {
  int label = -1;
  std::vector<Neighbor> in, out;
  explicit Vertex(int lb = -1) : label(lb) {}
  unsigned eSize() const { return (unsigned)(in.size() + out.size()); }
};

class Graph {
 public:
  Graph() = default;
  ~Graph();
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  std::vector<Vertex> vertices;
  void addVertex(int label);
  void addEdge(int from, int to, int elb);

  // CSR (valid after transformToCSR)
  unsigned vertex_num = 0;
  unsigned* vertex_value = nullptr;             // [n] labels
  unsigned* undirected_row_offset = nullptr;    // [n+1]
  unsigned* undirected_column_index = nullptr;  // [row_offset[n]] sorted, deduplicated rows
  unsigned label_num = 0;
  unsigned* inverse_label = nullptr;            // [label_num] distinct labels, ascending
  unsigned* inverse_num = nullptr;              // [label_num] vertices per label

  void transformToCSR();
  // The CSR fields (and the inverse label list) from an existing undirected
  // CSR, e.g. a mapped binary cache; the in/out lists stay empty.
  void setCSR(unsigned n, const unsigned* row_offset, const unsigned* col, const unsigned* labels);

 private:
  void build_inverse_labels();
  void free_csr();

 public:
  // directed, edge-labelled membership (graph/Graph.cpp:107-118)
  bool isEdgeContained(int from, int to, int label) const;
  unsigned eSize() const { return undirected_row_offset ? undirected_row_offset[vertex_num] : 0; }
  unsigned vSize() const { return vertex_num; }
  void getNeighbor(unsigned id, unsigned& num, unsigned& offset) const {
    offset = undirected_row_offset[id];
    num = undirected_row_offset[id + 1] - offset;
  }
  int getVertexDegree(unsigned id) const { return (int)(undirected_row_offset[id + 1] - undirected_row_offset[id]); }
};
