// CPU restatement of the reference's merge-scoring path -- TEST INFRASTRUCTURE.
//
// This oracle is the checker for the HIP path.  Only tests/, the smoke check and
// the benchmark's CPU-baseline leg may load it; the product (beam_amd/) never
// links, loads or calls it.  Each function cites the reference file:line it
// restates (paths relative to the reference repository root).
//
// Parity pins (see tests/test_oracle.py): golden end-to-end summaries and
// per-iteration counters written by the compiled reference (tests/golden/),
// per-pair gain vectors written by the reference's own QuotientGraph through
// oracle/ref_gain_harness.cpp, the full-recompute CostOracle identities, and
// brute force on tiny graphs.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace oracle {

struct Graph {  // graph.hpp:8-29
  int n = 0;
  int64_t m = 0;  // arcs (2 x edges)
  int64_t input_edges_raw = 0;
  std::vector<int64_t> row_ptr;
  std::vector<int> col_idx;
};

Graph LoadEdgeList(const std::string& path);  // graph_io.cpp:12-106
Graph FromEdges(int n, const std::vector<std::pair<int, int>>& edges);

int64_t BlockCost(int64_t sa, int64_t sb, int64_t e, bool internal);  // cost_oracle.cpp:93-110
int64_t ExactPartitionCost(const Graph& g, const std::vector<int>& labels);  // :112-124
int64_t FullMergeGain(const Graph& g, const std::vector<int>& labels, int a, int b);  // :126-168

struct GainResult {
  int64_t gain = 0;
  int64_t before = 0;
};

// Persistent supernode-to-supernode state (quotient_graph.hpp:143-207).
class Quotient {
 public:
  using Row = std::vector<std::pair<int, int64_t>>;
  void Rebuild(const Graph& g, const std::vector<int>& labels);  // quotient_graph.cpp:80-145
  GainResult Gain(int a, int b, uint64_t* raw_entries = nullptr) const;  // :340-470
  void Merge(int keep, int remove);  // :628-742
  void MergeBulk(const std::vector<std::pair<int, int>>& pairs);  // :759-856
  int64_t ExactCost() const;  // :240-257
  uint64_t Nnz() const;  // :259-270
  bool Active(int r) const { return active_[static_cast<size_t>(r)] != 0; }
  int64_t Size(int r) const { return sizes_[static_cast<size_t>(r)]; }
  int64_t Internal(int r) const { return internal_[static_cast<size_t>(r)]; }
  const Row& RowOf(int r) const { return rows_[static_cast<size_t>(r)]; }
  int N() const { return static_cast<int>(active_.size()); }
  bool SameAs(const Quotient& other) const;

 private:
  std::vector<int64_t> sizes_, internal_;
  std::vector<Row> rows_;
  std::vector<uint8_t> active_;
};

struct Options {  // main.cpp:14-56 for the flags this path reads
  int top_k = 16;
  int group_batch_size = 64;
  int divide_hash_dims = 16;
  int divide_max_group = 512;
  bool ea_use_threshold = true;
  uint64_t seed = 1;
};

struct Stats {  // subset of RuntimeStats (sweg.hpp:221-386) used by the checks
  int64_t divide_max_group_size = 0;
  int64_t divide_fallback_splits = 0;
  int64_t merge_candidate_pairs = 0;
  int64_t merge_selected_pairs = 0;
  int64_t merge_total_local_gain = 0;
  uint64_t merge_group_count = 0;
  uint64_t merge_group_max_size = 0;
  uint64_t merge_raw_pair_count = 0;
  uint64_t merge_candidate_pairs_after_prune = 0;
  uint64_t merge_exact_gain_calls = 0;
  uint64_t merge_positive_gain_pairs = 0;
  uint64_t merge_rejected_by_overlap = 0;
  uint64_t merge_rejected_by_threshold = 0;
  uint64_t selected_merges_for_validation = 0;
  uint64_t accepted_merges_after_validation = 0;
  uint64_t rejected_nonpositive = 0;
  uint64_t validation_exact_calls = 0;
  uint64_t validation_exact_row_entry_work = 0;
  int64_t isolated_gain_sum = 0;
};

struct IterationStats {  // IterationProfile (sweg.hpp:83-178) columns used by the checks
  int iteration = 0;
  int64_t active_supernodes = 0;
  int64_t group_count = 0;
  uint64_t candidate_proxy_pairs_examined = 0;
  uint64_t candidate_pairs_after_topk = 0;
  uint64_t exact_gain_calls = 0;
  uint64_t exact_gain_positive_count = 0;
  uint64_t above_threshold_count = 0;
  uint64_t matching_selected_count = 0;
  uint64_t actual_merge_count = 0;
};

struct CandidatePair {  // EaCandidatePair (sweg.hpp:477-482)
  int a_idx = -1;
  int b_idx = -1;
  int64_t gain = 0;
  int64_t before = 0;
};

struct Encoding {  // EncodingResult (sweg.hpp:212-219)
  std::vector<std::vector<int>> supernodes;
  std::vector<int> supernode_sizes;
  std::vector<int> node_to_supernode;
  std::vector<std::pair<int, int>> P, Cp, Cm;
};

class Sweg {
 public:
  Sweg(const Graph& g, const Options& opt);  // sweg.cpp:102-194
  void Run(int iterations);  // sweg.cpp:651-709
  void Divide(int iter);  // sweg.cpp:853-890
  void Merge(int iter);  // sweg.cpp:953-1145 (blocked branch)

  // Legacy EA-proxy candidate generation for one group (sweg.cpp:1598-1999).
  std::vector<CandidatePair> PrepareGroup(const std::vector<int>& q, uint64_t* proxy_examined) const;

  uint64_t PartitionHash() const;  // sweg.cpp:2920-2928
  std::vector<int> Labels() const;  // sweg.cpp:2902-2918
  Encoding Encode() const;  // sweg.cpp:3017-3163 (mags-compatible branch)
  // G.txt / P.txt / Cp.txt / Cm.txt of an encoding (sweg.cpp:3255-3287).
  static void WriteOutput(const std::string& dir, const Encoding& e);
  int64_t ExactCurrentPartitionCost() const;

  const std::vector<std::pair<int, int>>& Groups() const { return groups_; }  // (start, length)
  const std::vector<int>& GroupOrder() const { return active_reps_; }
  const Quotient& Q() const { return q_; }
  Quotient& MutableQ() { return q_; }
  const Stats& stats() const { return stats_; }
  const std::vector<IterationStats>& rounds() const { return rounds_; }
  // Commit log: per block of every iteration, the accepted (keep, remove) pairs.
  const std::vector<std::vector<std::pair<int, int>>>& commit_log() const { return commit_log_; }
  void set_record_commits(bool on) { record_commits_ = on; }
  // Applies Update_S for an externally chosen merge (sweg.cpp:2977-3015).
  void UpdateS(int keep, int remove);

 private:
  uint64_t HashRank64(int node, int dim, int iter) const;  // sweg.cpp:731-737
  uint64_t DimRank(int dim, int iter) const;  // :739-744
  uint64_t FallbackRank(int rep, int iter) const;  // :746-752
  uint64_t SupernodeMinHash(int rep, int dim, int iter);  // :754-784
  void FallbackSplit(int l, int r, int iter);  // :786-810
  void Split(int l, int r, int depth, int iter);  // :812-851
  void ScoreFilterSelect(const std::vector<int>& q, std::vector<CandidatePair>* pairs, double theta,
                         std::vector<std::pair<int, int>>* selected);
  void Commit(const std::vector<std::pair<int, int>>& pairs);  // :2665-2769

  const Graph* g_;
  Options opt_;
  int n_;
  std::vector<int> I_, J_, S_, sizes_by_rep_;
  std::vector<int> active_reps_;
  std::vector<std::pair<int, int>> groups_;
  std::vector<int> dim_order_;
  std::unordered_map<uint64_t, uint64_t> minhash_cache_;  // (rep, dim) memo, per Divide
  Quotient q_;
  Stats stats_;
  std::vector<IterationStats> rounds_;
  IterationStats cur_;
  bool record_commits_ = false;
  std::vector<std::vector<std::pair<int, int>>> commit_log_;
};

uint64_t SplitMix64(uint64_t x);  // sweg.cpp:2290-2295

}  // namespace oracle
