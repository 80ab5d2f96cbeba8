// Golden-vector harness for the exact merge gain (test infrastructure only).
//
// Linked against the reference's own QuotientGraph / CostOracle / graph loader
// by oracle/ref.mk; it never ships.  It replays a script of operations on a
// graph and prints the reference's answers, which tests/golden/ commits as
// fixtures:
//
//   M keep remove   QuotientGraph::Merge(keep, remove)      (quotient_graph.cpp:628-742)
//   G a b           ExactMergeGainPersistent(a, b)         (quotient_graph.cpp:340-470)
//                   -> "G a b gain before"
//   F a b           CostOracle::ExactMergeGain on labels   (cost_oracle.cpp:126-168)
//                   -> "F a b gain"
//   C               QuotientGraph::ExactCost and CostOracle::ExactPartitionCost
//                   -> "C quotient_cost oracle_cost"
//
// Usage: gain_harness <edge-list> <script>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cost_oracle.hpp"
#include "graph_io/graph_io.hpp"
#include "quotient_graph.hpp"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: gain_harness <edge-list> <script>\n";
    return 2;
  }
  try {
    const CSRGraph graph = LoadUndirectedEdgeList(argv[1]);
    QuotientGraph quotient(graph, false, false);
    CostOracle oracle;
    std::vector<int> labels(static_cast<size_t>(graph.n));
    for (int v = 0; v < graph.n; ++v) labels[static_cast<size_t>(v)] = v;

    std::ifstream script(argv[2]);
    std::string line;
    while (std::getline(script, line)) {
      std::istringstream in(line);
      std::string op;
      if (!(in >> op)) continue;
      if (op == "M") {
        int keep = 0, remove = 0;
        in >> keep >> remove;
        quotient.Merge(keep, remove);
        for (int& label : labels) {
          if (label == remove) label = keep;
        }
      } else if (op == "G") {
        int a = 0, b = 0;
        in >> a >> b;
        const QuotientMergeGainResult r = quotient.ExactMergeGainPersistent(a, b);
        std::cout << "G " << a << ' ' << b << ' ' << r.gain << ' ' << r.before_cost
                  << '\n';
      } else if (op == "F") {
        int a = 0, b = 0;
        in >> a >> b;
        std::cout << "F " << a << ' ' << b << ' '
                  << oracle.ExactMergeGain(graph, labels, a, b) << '\n';
      } else if (op == "C") {
        std::cout << "C " << quotient.ExactCost() << ' '
                  << oracle.ExactPartitionCost(graph, labels) << '\n';
      } else {
        std::cerr << "unknown op: " << op << '\n';
        return 2;
      }
    }
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << '\n';
    return 1;
  }
  return 0;
}
