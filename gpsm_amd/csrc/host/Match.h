// Matcher front end of the C++ driver, same entry points as the reference's
// Match (match/Match.h:80-125): initGPU(dev), Match(query, data) and
// match(io, final_result, rows, cols, id_map).  It calls the device through
// the C ABI (gpsm.h) and keeps each data graph resident across queries.
//
// Result convention: final_result (new[]-allocated, caller delete[]s) holds
// the embedding set, injective rows only, sorted, one column per query vertex
// in query-id order, so id_map is the identity (owned by the Match object,
// valid while it lives).  The reference returned join-order columns with
// non-injective rows for IO::verify to drop.
#pragma once
#include <cstdint>

#include "Graph.h"
#include "IO.h"
#include "gpsm.h"

class Match {
 public:
  Match(Graph* query, Graph* data);
  ~Match();
  // count-only variant: embedding count and set hash, nothing materialised
  int count(uint64_t& embeddings, uint64_t& hash);
  void match(IO& io, unsigned*& final_result, unsigned& result_row_num, unsigned& result_col_num, int*& id_map);
  // The reference's directed, edge-labelled final check (io/IO.cpp:146-180)
  // on the device: keeps the rows of final_result (as match returns them)
  // whose query arcs are data arcs with the same label, in place.  Needs the
  // data graph's in/out lists (a text input, not a binary cache).
  int verify_directed(unsigned* final_result, unsigned& result_row_num);
  // IO::output's embedding lines formatted on the device into fd
  // (gpsm_write_tuples); an IO::RowWriter.
  static int write_rows(int fd, const unsigned* rows, unsigned n, unsigned cols);
  // Selects the device (match/Match.h:87).  The reference exit()s when none is
  // usable; here the failure is printed and gpuReady() returns false.
  static void initGPU(int dev);
  static bool gpuReady();
  static void releaseGPU();
  // Frees the device copy of a data graph (call before deleting the Graph:
  // resident copies are keyed by its address).
  static void forget(const Graph* data);
  int last_status() const { return status; }

 private:
  Graph* query;
  Graph* data;
  int* id2pos = nullptr;
  int status = GPSM_OK;
  gpsm_graph* resident();
};
