// C entry points over the CPU oracle -- TEST INFRASTRUCTURE (loaded by tests/
// through ctypes and by the smoke check as the checker; never by the product).
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "oracle.hpp"

namespace {

thread_local std::string g_error;

struct Handle {
  oracle::Graph graph;
  std::unique_ptr<oracle::Sweg> sweg;
};

template <typename F>
int Guard(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_error.c_str(); }

void* orc_create(const char* path, int top_k, int group_batch, int hash_dims, int max_group,
                 unsigned long long seed) {
  try {
    auto h = std::make_unique<Handle>();
    h->graph = oracle::LoadEdgeList(path);
    oracle::Options opt;
    opt.top_k = top_k;
    opt.group_batch_size = group_batch;
    opt.divide_hash_dims = hash_dims;
    opt.divide_max_group = max_group;
    opt.seed = seed;
    h->sweg = std::make_unique<oracle::Sweg>(h->graph, opt);
    return h.release();
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void orc_free(void* h) { delete static_cast<Handle*>(h); }

int orc_n(void* h) { return static_cast<Handle*>(h)->graph.n; }

long long orc_arcs(void* h) { return static_cast<Handle*>(h)->graph.m; }

// CSR of the loaded graph (row_ptr: n+1 int64, col_idx: arcs int32).
int orc_graph_csr(void* h, long long* row_ptr, int* col_idx) {
  auto* s = static_cast<Handle*>(h);
  std::memcpy(row_ptr, s->graph.row_ptr.data(), s->graph.row_ptr.size() * sizeof(long long));
  std::memcpy(col_idx, s->graph.col_idx.data(), s->graph.col_idx.size() * sizeof(int));
  return 0;
}

void orc_record_commits(void* h, int on) { static_cast<Handle*>(h)->sweg->set_record_commits(on != 0); }

int orc_run(void* h, int iterations) {
  return Guard([&] {
    static_cast<Handle*>(h)->sweg->Run(iterations);
    return 0;
  });
}

unsigned long long orc_partition_hash(void* h) { return static_cast<Handle*>(h)->sweg->PartitionHash(); }

int orc_labels(void* h, int* out) {
  return Guard([&] {
    const auto labels = static_cast<Handle*>(h)->sweg->Labels();
    std::memcpy(out, labels.data(), labels.size() * sizeof(int));
    return 0;
  });
}

long long orc_exact_partition_cost(void* h) { return static_cast<Handle*>(h)->sweg->ExactCurrentPartitionCost(); }

long long orc_quotient_cost(void* h) { return static_cast<Handle*>(h)->sweg->Q().ExactCost(); }

// Summary counters in a fixed order (documented in tests/oracle_lib.py).
int orc_summary(void* h, long long* out) {
  const oracle::Stats& s = static_cast<Handle*>(h)->sweg->stats();
  const long long v[] = {
      s.divide_max_group_size, s.divide_fallback_splits, s.merge_candidate_pairs,
      s.merge_selected_pairs, s.merge_total_local_gain, static_cast<long long>(s.merge_group_count),
      static_cast<long long>(s.merge_group_max_size), static_cast<long long>(s.merge_raw_pair_count),
      static_cast<long long>(s.merge_candidate_pairs_after_prune),
      static_cast<long long>(s.merge_exact_gain_calls), static_cast<long long>(s.merge_positive_gain_pairs),
      static_cast<long long>(s.merge_rejected_by_overlap), static_cast<long long>(s.merge_rejected_by_threshold),
      static_cast<long long>(s.selected_merges_for_validation),
      static_cast<long long>(s.accepted_merges_after_validation), static_cast<long long>(s.rejected_nonpositive),
      static_cast<long long>(s.validation_exact_calls), static_cast<long long>(s.validation_exact_row_entry_work),
      s.isolated_gain_sum};
  std::memcpy(out, v, sizeof(v));
  return static_cast<int>(sizeof(v) / sizeof(v[0]));
}

// Per-iteration counters, 10 per round (IterationStats field order).
int orc_rounds(void* h, long long* out, int max_rounds) {
  const auto& r = static_cast<Handle*>(h)->sweg->rounds();
  const int count = std::min<int>(max_rounds, static_cast<int>(r.size()));
  for (int i = 0; i < count; ++i) {
    const auto& x = r[static_cast<size_t>(i)];
    long long* o = out + 10 * i;
    o[0] = x.iteration;
    o[1] = x.active_supernodes;
    o[2] = x.group_count;
    o[3] = static_cast<long long>(x.candidate_proxy_pairs_examined);
    o[4] = static_cast<long long>(x.candidate_pairs_after_topk);
    o[5] = static_cast<long long>(x.exact_gain_calls);
    o[6] = static_cast<long long>(x.exact_gain_positive_count);
    o[7] = static_cast<long long>(x.above_threshold_count);
    o[8] = static_cast<long long>(x.matching_selected_count);
    o[9] = static_cast<long long>(x.actual_merge_count);
  }
  return count;
}

// Divide for `iter`; writes the position->rep order (active count entries) and
// the groups as (start, length); returns the group count.
int orc_divide(void* h, int iter, int* order, int* starts, int* lengths) {
  return Guard([&] {
    auto& sw = *static_cast<Handle*>(h)->sweg;
    sw.Divide(iter);
    const auto& o = sw.GroupOrder();
    std::memcpy(order, o.data(), o.size() * sizeof(int));
    const auto& gr = sw.Groups();
    for (size_t i = 0; i < gr.size(); ++i) {
      starts[i] = gr[i].first;
      lengths[i] = gr[i].second;
    }
    return static_cast<int>(gr.size());
  });
}

int orc_active_count(void* h) {
  auto& sw = *static_cast<Handle*>(h)->sweg;
  int c = 0;
  for (int r = 0; r < sw.Q().N(); ++r) c += sw.Q().Active(r) ? 1 : 0;
  return c;
}

// Legacy candidate generation for one group of reps; returns the pair count.
int orc_prepare_group(void* h, const int* q, int g, int* a_idx, int* b_idx, int cap,
                      unsigned long long* examined) {
  return Guard([&] {
    std::vector<int> reps(q, q + g);
    uint64_t ex = 0;
    const auto pairs = static_cast<Handle*>(h)->sweg->PrepareGroup(reps, &ex);
    if (examined) *examined = ex;
    const int count = static_cast<int>(pairs.size());
    for (int i = 0; i < count && i < cap; ++i) {
      a_idx[i] = pairs[static_cast<size_t>(i)].a_idx;
      b_idx[i] = pairs[static_cast<size_t>(i)].b_idx;
    }
    return count;
  });
}

int orc_gain(void* h, int a, int b, long long* gain, long long* before) {
  return Guard([&] {
    const auto r = static_cast<Handle*>(h)->sweg->Q().Gain(a, b);
    *gain = r.gain;
    *before = r.before;
    return 0;
  });
}

int orc_full_gain(void* h, int a, int b, long long* gain) {
  return Guard([&] {
    auto* s = static_cast<Handle*>(h);
    *gain = oracle::FullMergeGain(s->graph, s->sweg->Labels(), a, b);
    return 0;
  });
}

int orc_merge(void* h, int keep, int remove) {
  return Guard([&] {
    static_cast<Handle*>(h)->sweg->UpdateS(keep, remove);
    return 0;
  });
}

int orc_merge_bulk(void* h, const int* keeps, const int* removes, int count) {
  return Guard([&] {
    auto& sw = *static_cast<Handle*>(h)->sweg;
    std::vector<std::pair<int, int>> pairs;
    for (int i = 0; i < count; ++i) pairs.emplace_back(keeps[i], removes[i]);
    oracle::Quotient copy = sw.Q();
    copy.MergeBulk(pairs);
    for (const auto& p : pairs) sw.UpdateS(p.first, p.second);
    return copy.SameAs(sw.Q()) ? 0 : 1;
  });
}

long long orc_nnz_cross(void* h) {
  const auto& q = static_cast<Handle*>(h)->sweg->Q();
  long long total = 0;
  for (int r = 0; r < q.N(); ++r)
    if (q.Active(r)) total += static_cast<long long>(q.RowOf(r).size());
  return total;
}

// Quotient export: row offsets (n+1), neighbours, counts, sizes, internal, active.
int orc_export_quotient(void* h, long long* offsets, int* nbr, long long* cnt, long long* sizes,
                        long long* internal, unsigned char* active) {
  const auto& q = static_cast<Handle*>(h)->sweg->Q();
  long long pos = 0;
  for (int r = 0; r < q.N(); ++r) {
    offsets[r] = pos;
    active[r] = q.Active(r) ? 1 : 0;
    sizes[r] = q.Active(r) ? q.Size(r) : 0;
    internal[r] = q.Active(r) ? q.Internal(r) : 0;
    if (!q.Active(r)) continue;
    for (const auto& [x, c] : q.RowOf(r)) {
      nbr[pos] = x;
      cnt[pos] = c;
      ++pos;
    }
  }
  offsets[q.N()] = pos;
  return 0;
}

// Encoding counts: P, Cp, Cm, loops in P; returns 0.
int orc_encode_counts(void* h, long long* out) {
  return Guard([&] {
    const oracle::Encoding e = static_cast<Handle*>(h)->sweg->Encode();
    long long loops = 0;
    for (const auto& p : e.P) loops += p.first == p.second ? 1 : 0;
    out[0] = static_cast<long long>(e.P.size());
    out[1] = static_cast<long long>(e.Cp.size());
    out[2] = static_cast<long long>(e.Cm.size());
    out[3] = loops;
    out[4] = static_cast<long long>(e.supernodes.size());
    return 0;
  });
}

// Writes G.txt, P.txt, Cp.txt and Cm.txt of the current partition into dir.
int orc_write_output(void* h, const char* dir) {
  return Guard([&] {
    oracle::Sweg::WriteOutput(dir, static_cast<Handle*>(h)->sweg->Encode());
    return 0;
  });
}

// Commit log of recorded blocks, flattened: for block i, counts[i] pairs.
int orc_commit_log(void* h, int* counts, int* keeps, int* removes, int max_blocks, int max_pairs) {
  const auto& log = static_cast<Handle*>(h)->sweg->commit_log();
  int pos = 0;
  const int blocks = std::min<int>(max_blocks, static_cast<int>(log.size()));
  for (int b = 0; b < blocks; ++b) {
    counts[b] = static_cast<int>(log[static_cast<size_t>(b)].size());
    for (const auto& p : log[static_cast<size_t>(b)]) {
      if (pos < max_pairs) {
        keeps[pos] = p.first;
        removes[pos] = p.second;
      }
      ++pos;
    }
  }
  return static_cast<int>(log.size());
}

}  // extern "C"
