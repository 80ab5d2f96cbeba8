// Edge-list loader (reference semantics of src/graph_io/graph_io.cpp:12-106).
#pragma once

#include <string>
#include <vector>

namespace beam_host {

struct Graph {
  int n = 0;
  long long arcs = 0;             // 2 x edges (CSRGraph::m)
  long long input_edges_raw = 0;  // edge lines read
  std::vector<long long> row_ptr;
  std::vector<int> col;
};

Graph LoadEdgeList(const std::string& path);

}  // namespace beam_host
