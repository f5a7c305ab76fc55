// Text input/output of the C++ driver, same entry points as the reference's
// IO (io/IO.h:15-39).  Format (io/IO.cpp:46-100): "t # id" opens a graph,
// then "n m Lv Le", "v id label" lines and "e from to label" lines; "t # -1"
// ends the stream.  Unlike the reference, a missing terminator or a
// malformed line is an error (the reference loops forever on the former).
#pragma once
#include <cstdio>
#include <string>
#include <vector>

#include "Graph.h"

class IO {
 public:
  IO() = default;
  IO(const std::string& query, const std::string& data, const std::string& file);
  ~IO();
  bool ok() const { return qfp && dfp && ofp; }
  // all query graphs (CSR built)
  bool input(std::vector<Graph*>& query_list);
  // next data graph (CSR built); false at the terminator or on error
  bool input(Graph*& data_graph);
  // the same without counting it (data ids are assigned with next_data):
  // safe on another thread while this one writes output
  Graph* read_data();
  // next graph of a stream; nullptr at the terminator, sets error() on failure
  Graph* input(FILE* fp);
  bool error() const { return err; }
  // a data graph read by other means (a binary CSR cache) counts as the next
  void next_data() { ++data_id; }
  bool output(int qid);
  // separator after a data graph's results: three newlines (io/IO.cpp:138-142)
  bool output();
  // one embedding given as m[i] = image of query vertex i: "(i, m[i]) " pairs
  // and a newline (io/IO.cpp:223-232)
  bool output(int* m, int size);
  // one line per embedding: "(query id, data id)" pairs in query-id order
  bool output(const unsigned* final_result, unsigned rows, unsigned cols, const int* id_map, Graph* query,
              Graph* data);
  void flush();
  // Formats rows (query-id order) into a file descriptor, e.g. on the device
  // (gpsm_write_tuples); output() uses it for identity id maps.  Returns 0 on
  // success.
  using RowWriter = int (*)(int fd, const unsigned* rows, unsigned n, unsigned cols);
  void set_row_writer(RowWriter w) { writer = w; }

 private:
  int data_id = -1;
  FILE* qfp = nullptr;
  FILE* dfp = nullptr;
  FILE* ofp = nullptr;
  bool err = false;
  RowWriter writer = nullptr;
  std::string line = "============================================================";
};
