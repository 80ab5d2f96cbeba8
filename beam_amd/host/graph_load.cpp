// Edge-list loader with the reference's semantics (src/graph_io/graph_io.cpp:12-106):
// '#' comment and blank lines are skipped, each other line needs at least two
// integers "u v" (further tokens are ignored), negative ids, ids above INT_MAX
// and self-loops are rejected with the reference's messages, both arcs of
// every edge are stored, n = max id + 1, each row is sorted and duplicates are
// kept.  The file is read in one block and parsed without iostreams.
#include "graph_load.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace beam_host {

namespace {

bool IsSpace(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Parses one signed integer the way `istream >> int64_t` does after skipping
// whitespace; returns false when no integer starts at *p.
bool ParseInt(const char*& p, const char* end, long long* out) {
  while (p < end && IsSpace(*p)) ++p;
  if (p >= end) return false;
  bool neg = false;
  if (*p == '+' || *p == '-') {
    neg = *p == '-';
    ++p;
  }
  if (p >= end || *p < '0' || *p > '9') return false;
  unsigned long long v = 0;
  bool overflow = false;
  while (p < end && *p >= '0' && *p <= '9') {
    const unsigned d = static_cast<unsigned>(*p - '0');
    if (v > (static_cast<unsigned long long>(std::numeric_limits<long long>::max()) - d) / 10) overflow = true;
    v = v * 10 + d;
    ++p;
  }
  if (overflow) return false;
  *out = neg ? -static_cast<long long>(v) : static_cast<long long>(v);
  return true;
}

}  // namespace

Graph LoadEdgeList(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("Failed to open file: " + path);
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<char> buf(static_cast<size_t>(std::max<long>(size, 0)));
  if (size > 0 && std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
    std::fclose(f);
    throw std::runtime_error("Failed to open file: " + path);
  }
  std::fclose(f);

  std::vector<int> us, vs;
  long long max_node = -1, line_no = 0;
  const char* p = buf.data();
  const char* end = p + buf.size();
  if (buf.size() >= 16 && std::memcmp(buf.data(), "BEAMEDG1", 8) == 0) {
    // Binary edge list (not a reference format): "BEAMEDG1", uint64 edge
    // count, uint32 (u, v) pairs; edge k is reported as line k + 1.
    uint64_t m = 0;
    std::memcpy(&m, buf.data() + 8, sizeof(m));
    if (m > (buf.size() - 16) / 8 || buf.size() != 16 + 8 * m) {
      throw std::runtime_error("Malformed binary edge list: size does not match the edge count.");
    }
    us.resize(m);
    vs.resize(m);
    for (uint64_t k = 0; k < m; ++k) {
      uint32_t e[2];
      std::memcpy(e, buf.data() + 16 + 8 * k, sizeof(e));
      const long long u = e[0], v = e[1];
      if (u > std::numeric_limits<int>::max() || v > std::numeric_limits<int>::max()) {
        throw std::runtime_error("Node id exceeds int range at line " + std::to_string(k + 1) + ".");
      }
      if (u == v) throw std::runtime_error("Self-loop is not supported at line " + std::to_string(k + 1) + ".");
      us[k] = static_cast<int>(u);
      vs[k] = static_cast<int>(v);
      max_node = std::max({max_node, u, v});
    }
    std::vector<char>().swap(buf);
    p = end = nullptr;
  }
  while (p < end) {
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    if (!eol) eol = end;
    ++line_no;
    const char* q = p;
    while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q < eol && *q != '#') {
      long long u = 0, v = 0;
      const char* r = q;
      if (!ParseInt(r, eol, &u) || !ParseInt(r, eol, &v)) {
        throw std::runtime_error("Malformed edge list line " + std::to_string(line_no) +
                                 ": expected at least two integers 'u v'.");
      }
      if (u < 0 || v < 0) throw std::runtime_error("Negative node id at line " + std::to_string(line_no) + ".");
      if (u > std::numeric_limits<int>::max() || v > std::numeric_limits<int>::max()) {
        throw std::runtime_error("Node id exceeds int range at line " + std::to_string(line_no) + ".");
      }
      if (u == v) throw std::runtime_error("Self-loop is not supported at line " + std::to_string(line_no) + ".");
      us.push_back(static_cast<int>(u));
      vs.push_back(static_cast<int>(v));
      max_node = std::max({max_node, u, v});
    }
    p = eol + 1;
  }
  if (max_node + 1 <= 0) throw std::runtime_error("Invalid node count.");
  Graph g;
  g.n = static_cast<int>(max_node + 1);
  g.input_edges_raw = static_cast<long long>(us.size());
  g.arcs = 2 * g.input_edges_raw;
  g.row_ptr.assign(static_cast<size_t>(g.n) + 1, 0);
  for (size_t i = 0; i < us.size(); ++i) {
    g.row_ptr[static_cast<size_t>(us[i]) + 1]++;
    g.row_ptr[static_cast<size_t>(vs[i]) + 1]++;
  }
  for (int i = 0; i < g.n; ++i) g.row_ptr[static_cast<size_t>(i) + 1] += g.row_ptr[static_cast<size_t>(i)];
  g.col.assign(static_cast<size_t>(g.arcs), 0);
  std::vector<long long> cur(g.row_ptr.begin(), g.row_ptr.end() - 1);
  for (size_t i = 0; i < us.size(); ++i) {
    g.col[static_cast<size_t>(cur[static_cast<size_t>(us[i])]++)] = vs[i];
    g.col[static_cast<size_t>(cur[static_cast<size_t>(vs[i])]++)] = us[i];
  }
  for (int v = 0; v < g.n; ++v) {
    std::sort(g.col.begin() + g.row_ptr[static_cast<size_t>(v)], g.col.begin() + g.row_ptr[static_cast<size_t>(v) + 1]);
  }
  return g;
}

}  // namespace beam_host

// C entry points for callers that do not link C++ (the Python wrapper).
namespace {
thread_local std::string g_load_error;
}

extern "C" {

void* beam_graph_load(const char* path) {
  try {
    return new beam_host::Graph(beam_host::LoadEdgeList(path));
  } catch (const std::exception& e) {
    g_load_error = e.what();
    return nullptr;
  }
}

const char* beam_graph_last_error(void) { return g_load_error.c_str(); }

void beam_graph_info(void* g, int* n, long long* arcs, long long* input_edges_raw) {
  const auto* gr = static_cast<beam_host::Graph*>(g);
  *n = gr->n;
  *arcs = gr->arcs;
  *input_edges_raw = gr->input_edges_raw;
}

void beam_graph_csr(void* g, long long* row_ptr, int* col) {
  const auto* gr = static_cast<beam_host::Graph*>(g);
  std::memcpy(row_ptr, gr->row_ptr.data(), gr->row_ptr.size() * sizeof(long long));
  std::memcpy(col, gr->col.data(), gr->col.size() * sizeof(int));
}

void beam_graph_free(void* g) { delete static_cast<beam_host::Graph*>(g); }

}  // extern "C"
