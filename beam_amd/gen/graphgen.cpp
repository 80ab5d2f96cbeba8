// Deterministic synthetic graph generators for the five benchmark workloads.
//
// Output is the reference's input format (README.md:105-123; loader
// src/graph_io/graph_io.cpp:12-106): one "u v" line per undirected edge, each
// edge once, no self-loops, no duplicates, uint32 ids.  Lines are sorted by
// (u, v) with u < v.
//
//   graphgen <workload> <out.txt>     workload in {fa, am, yo, lj, u2, tiny<N>}
//   graphgen <workload> <out.bin>     the same edges in the binary edge-list
//                                     format the loader also reads: "BEAMEDG1",
//                                     uint64 edge count, then uint32 (u, v) pairs
//
// Choices that the workload descriptions leave open are fixed here and stated
// next to each generator (max-degree handling, community assignment, isolate
// relabelling).  Every random draw comes from xoshiro256** seeded through
// SplitMix64, so a run is a pure function of (workload, seed).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

uint64_t SplitMix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) w = SplitMix(x);
  }
  static uint64_t Rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t Next() {
    const uint64_t result = Rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = Rotl(s[3], 45);
    return result;
  }
  double Uniform() { return static_cast<double>(Next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t Below(uint64_t n) { return static_cast<uint64_t>(Uniform() * static_cast<double>(n)) % n; }
};

// Vose alias table over non-negative weights.
struct Alias {
  std::vector<double> prob;
  std::vector<uint32_t> alias;
  std::vector<uint32_t> ids;  // optional: sample returns ids[k] when non-empty

  void Build(const std::vector<double>& w) {
    const size_t n = w.size();
    prob.assign(n, 0.0);
    alias.assign(n, 0);
    double total = 0.0;
    for (double x : w) total += x;
    std::vector<double> scaled(n);
    std::vector<uint32_t> small, large;
    for (size_t i = 0; i < n; ++i) {
      scaled[i] = w[i] * static_cast<double>(n) / total;
      (scaled[i] < 1.0 ? small : large).push_back(static_cast<uint32_t>(i));
    }
    while (!small.empty() && !large.empty()) {
      const uint32_t s = small.back();
      small.pop_back();
      const uint32_t l = large.back();
      prob[s] = scaled[s];
      alias[s] = l;
      scaled[l] = (scaled[l] + scaled[s]) - 1.0;
      if (scaled[l] < 1.0) {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (uint32_t i : large) prob[i] = 1.0;
    for (uint32_t i : small) prob[i] = 1.0;
  }
  uint32_t Sample(Rng& rng) const {
    const uint64_t k = rng.Below(prob.size());
    const uint32_t pick = rng.Uniform() < prob[k] ? static_cast<uint32_t>(k) : alias[k];
    return ids.empty() ? pick : ids[pick];
  }
};

// Open-addressing set of undirected edge keys (lo << 32 | hi).
struct EdgeSet {
  std::vector<uint64_t> slots;
  uint64_t mask = 0;
  uint64_t count = 0;
  explicit EdgeSet(uint64_t expected) {
    uint64_t cap = 1;
    while (cap < expected * 2 + 16) cap <<= 1;
    slots.assign(cap, ~0ULL);
    mask = cap - 1;
  }
  static uint64_t Hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    return k ^ (k >> 33);
  }
  bool Insert(uint32_t u, uint32_t v) {
    const uint64_t lo = std::min(u, v), hi = std::max(u, v);
    const uint64_t key = (lo << 32) | hi;
    uint64_t h = Hash(key) & mask;
    while (true) {
      if (slots[h] == ~0ULL) {
        slots[h] = key;
        ++count;
        return true;
      }
      if (slots[h] == key) return false;
      h = (h + 1) & mask;
    }
  }
  std::vector<uint64_t> Sorted() const {
    std::vector<uint64_t> out;
    out.reserve(count);
    for (uint64_t k : slots)
      if (k != ~0ULL) out.push_back(k);
    std::sort(out.begin(), out.end());
    return out;
  }
};

// Chung-Lu weights w_i = (i + i0)^(-1/(gamma-1)); i0 >= 1 is chosen by
// bisection so that the largest expected degree 2m*w_0/sum(w) equals
// max_degree (an offset rather than a hard cap, so the tail stays a power law).
std::vector<double> ChungLuWeights(uint64_t n, uint64_t m, double gamma, double max_degree) {
  const double alpha = 1.0 / (gamma - 1.0);
  auto weights = [&](double i0) {
    std::vector<double> w(n);
    for (uint64_t i = 0; i < n; ++i) w[i] = std::pow(static_cast<double>(i) + i0, -alpha);
    return w;
  };
  auto max_expected = [&](double i0) {
    const std::vector<double> w = weights(i0);
    double total = 0.0;
    for (double x : w) total += x;
    return 2.0 * static_cast<double>(m) * w[0] / total;
  };
  double lo = 1.0, hi = 1.0;
  if (max_expected(1.0) <= max_degree) return weights(1.0);
  while (max_expected(hi) > max_degree) hi *= 2.0;
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    (max_expected(mid) > max_degree ? lo : hi) = mid;
  }
  return weights(hi);
}

// Draws from a continuous power law p(x) ~ x^-tau on [lo, hi].
double PowerLaw(Rng& rng, double tau, double lo, double hi) {
  const double a = 1.0 - tau;
  const double u = rng.Uniform();
  return std::pow(std::pow(lo, a) + u * (std::pow(hi, a) - std::pow(lo, a)), 1.0 / a);
}

void WriteEdgesBinary(const std::vector<uint64_t>& keys, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  const uint64_t m = keys.size();
  bool ok = std::fwrite("BEAMEDG1", 1, 8, f) == 8 && std::fwrite(&m, sizeof(m), 1, f) == 1;
  std::vector<uint32_t> buf;
  for (uint64_t i = 0; ok && i < m; i += 1 << 20) {
    const uint64_t e = std::min<uint64_t>(m, i + (1 << 20));
    buf.resize(2 * (e - i));
    for (uint64_t k = i; k < e; ++k) {
      buf[2 * (k - i)] = static_cast<uint32_t>(keys[k] >> 32);
      buf[2 * (k - i) + 1] = static_cast<uint32_t>(keys[k] & 0xffffffffU);
    }
    ok = std::fwrite(buf.data(), sizeof(uint32_t), buf.size(), f) == buf.size();
  }
  if (std::fclose(f) != 0 || !ok) throw std::runtime_error(std::string("cannot write ") + path);
}

void WriteEdges(const std::vector<uint64_t>& keys, const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  std::vector<char> buf(1 << 22);
  std::setvbuf(f, buf.data(), _IOFBF, buf.size());
  for (uint64_t k : keys) {
    std::fprintf(f, "%u %u\n", static_cast<uint32_t>(k >> 32), static_cast<uint32_t>(k & 0xffffffffU));
  }
  std::fclose(f);
}

// Communities for Chung-Lu with planted groups: node i belongs to community
// i % c, so each community holds an equal share of every weight decile.
std::vector<uint64_t> ChungLuCommunities(uint64_t n, uint64_t m, double gamma, double max_degree,
                                         int communities, double p_inside, uint64_t seed) {
  Rng rng(seed);
  const std::vector<double> w = ChungLuWeights(n, m, gamma, max_degree);
  Alias global;
  global.Build(w);
  std::vector<Alias> local(static_cast<size_t>(communities));
  for (int c = 0; c < communities; ++c) {
    std::vector<double> cw;
    for (uint64_t i = static_cast<uint64_t>(c); i < n; i += static_cast<uint64_t>(communities)) {
      cw.push_back(w[i]);
      local[static_cast<size_t>(c)].ids.push_back(static_cast<uint32_t>(i));
    }
    local[static_cast<size_t>(c)].Build(cw);
  }
  EdgeSet set(m);
  while (set.count < m) {
    const uint32_t u = global.Sample(rng);
    uint32_t v;
    if (communities > 0 && rng.Uniform() < p_inside) {
      v = local[u % static_cast<uint32_t>(communities)].Sample(rng);
    } else {
      v = global.Sample(rng);
    }
    if (u == v) continue;
    set.Insert(u, v);
  }
  return set.Sorted();
}

// LFR-style graph: power-law target degrees (exponent tau_deg, max dmax, mean
// 2m/n), power-law community sizes (exponent tau_com in [cmin, cmax]) filled
// from a seeded permutation, and a fraction mu of each node's edges leaving its
// community.  Edges are drawn Chung-Lu style from the target degrees: the first
// endpoint globally, the second inside the first endpoint's community with
// probability 1-mu, otherwise globally.
std::vector<uint64_t> LfrStyle(uint64_t n, uint64_t m, double tau_deg, double dmax, double tau_com,
                               double cmin, double cmax, double mu, uint64_t seed) {
  Rng rng(seed);
  const double mean = 2.0 * static_cast<double>(m) / static_cast<double>(n);
  // dmin so that the continuous power law on [dmin, dmax] has the target mean.
  auto pl_mean = [&](double lo) {
    const double a1 = 1.0 - tau_deg, a2 = 2.0 - tau_deg;
    return (a1 / a2) * (std::pow(dmax, a2) - std::pow(lo, a2)) / (std::pow(dmax, a1) - std::pow(lo, a1));
  };
  double lo = 0.5, hi = dmax;
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    (pl_mean(mid) < mean ? lo : hi) = mid;
  }
  const double dmin = 0.5 * (lo + hi);
  std::vector<double> deg(n);
  for (uint64_t i = 0; i < n; ++i) deg[i] = PowerLaw(rng, tau_deg, dmin, dmax);

  std::vector<uint32_t> perm(n);
  for (uint64_t i = 0; i < n; ++i) perm[i] = static_cast<uint32_t>(i);
  for (uint64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng.Below(i + 1)]);
  std::vector<uint32_t> community(n);
  std::vector<Alias> local;
  uint64_t filled = 0;
  while (filled < n) {
    uint64_t size = static_cast<uint64_t>(std::llround(PowerLaw(rng, tau_com, cmin, cmax)));
    size = std::min<uint64_t>(std::max<uint64_t>(size, 2), n - filled);
    Alias a;
    std::vector<double> cw;
    for (uint64_t k = 0; k < size; ++k) {
      const uint32_t v = perm[filled + k];
      community[v] = static_cast<uint32_t>(local.size());
      a.ids.push_back(v);
      cw.push_back(deg[v]);
    }
    a.Build(cw);
    local.push_back(std::move(a));
    filled += size;
  }
  Alias global;
  global.Build(deg);
  EdgeSet set(m);
  while (set.count < m) {
    const uint32_t u = global.Sample(rng);
    const uint32_t v = rng.Uniform() < 1.0 - mu ? local[community[u]].Sample(rng) : global.Sample(rng);
    if (u == v) continue;
    set.Insert(u, v);
  }
  return set.Sorted();
}

// Chung-Lu with overlapping communities: a fraction p_inside of the edges is
// drawn inside one of `communities` node sets (sizes power law tau_com on
// [cmin, cmax], members drawn uniformly, so sets overlap), the community chosen
// with probability proportional to its size; the rest is plain Chung-Lu.
std::vector<uint64_t> ChungLuOverlapping(uint64_t n, uint64_t m, double gamma, double max_degree,
                                         uint64_t communities, double tau_com, double cmin,
                                         double cmax, double p_inside, uint64_t seed) {
  Rng rng(seed);
  const std::vector<double> w = ChungLuWeights(n, m, gamma, max_degree);
  Alias global;
  global.Build(w);
  std::vector<Alias> local(communities);
  std::vector<double> com_weight(communities);
  for (uint64_t c = 0; c < communities; ++c) {
    const uint64_t size = static_cast<uint64_t>(std::llround(PowerLaw(rng, tau_com, cmin, cmax)));
    std::vector<double> cw;
    for (uint64_t k = 0; k < size; ++k) {
      const uint32_t v = static_cast<uint32_t>(rng.Below(n));
      local[c].ids.push_back(v);
      cw.push_back(w[v]);
    }
    local[c].Build(cw);
    com_weight[c] = static_cast<double>(size);
  }
  Alias pick_com;
  pick_com.Build(com_weight);
  EdgeSet set(m);
  while (set.count < m) {
    uint32_t u, v;
    if (rng.Uniform() < p_inside) {
      const Alias& a = local[pick_com.Sample(rng)];
      u = a.Sample(rng);
      v = a.Sample(rng);
    } else {
      u = global.Sample(rng);
      v = global.Sample(rng);
    }
    if (u == v) continue;
    set.Insert(u, v);
  }
  return set.Sorted();
}

unsigned Threads() { return std::max(1U, std::min(16U, std::thread::hardware_concurrency())); }

// Sorts v with T threads: slices sorted in parallel, then merged pairwise.
void ParallelSort(std::vector<uint64_t>& v) {
  const unsigned T = Threads();
  const size_t n = v.size();
  std::vector<size_t> cut(T + 1);
  for (unsigned t = 0; t <= T; ++t) cut[t] = n * t / T;
  std::vector<std::thread> ts;
  for (unsigned t = 0; t < T; ++t) ts.emplace_back([&, t] { std::sort(v.begin() + cut[t], v.begin() + cut[t + 1]); });
  for (auto& th : ts) th.join();
  for (unsigned w = 1; w < T; w *= 2) {
    ts.clear();
    for (unsigned t = 0; t + w < T; t += 2 * w) {
      const size_t lo = cut[t], mid = cut[t + w], hi = cut[std::min(T, t + 2 * w)];
      ts.emplace_back([&v, lo, mid, hi] { std::inplace_merge(v.begin() + lo, v.begin() + mid, v.begin() + hi); });
    }
    for (auto& th : ts) th.join();
  }
}

// R-MAT at the given scale with (a, b, c, d).  Edges are drawn in chunks of
// 2^20 draws; chunk k uses its own generator seeded from (seed, k), so the
// chunks can be drawn in parallel and the graph does not depend on the
// thread count.  Draws are symmetrised (lo < hi), self-loops dropped, and
// batches of chunks are drawn until at least `target_edges` distinct edges
// exist (the last batch is sized from the previous batches' distinct rate, so
// the count overshoots by well under one chunk's worth).  Isolated ids are
// then dropped by relabelling the remaining ids in ascending order to
// 0..n-1.  Progress goes to stderr.
std::vector<uint64_t> Rmat(int scale, double a, double b, double c, uint64_t target_edges,
                           uint64_t seed) {
  constexpr uint64_t kChunk = 1ULL << 20;
  const double ab = a + b, abc = a + b + c;
  const unsigned T = Threads();
  std::vector<uint64_t> uniq;
  uint64_t next_chunk = 0, drawn = 0;
  double rate = 1.0;  // distinct edges gained per draw in the last batch
  while (uniq.size() < target_edges) {
    const double need = static_cast<double>(target_edges - uniq.size()) / std::max(rate, 0.05);
    const uint64_t nchunks = std::max<uint64_t>(1, static_cast<uint64_t>(std::ceil(need / kChunk)));
    std::vector<uint64_t> batch(nchunks * kChunk);
    std::vector<uint64_t> filled(nchunks, 0);
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < T; ++t) {
      ts.emplace_back([&, t] {
        for (uint64_t k = t; k < nchunks; k += T) {
          Rng rng(seed ^ (0x9e3779b97f4a7c15ULL * (next_chunk + k + 1)));
          uint64_t* out = batch.data() + k * kChunk;
          uint64_t f = 0;
          for (uint64_t d = 0; d < kChunk; ++d) {
            uint32_t u = 0, v = 0;
            for (int bit = scale - 1; bit >= 0; --bit) {
              const double r = rng.Uniform();
              if (r < a) {
              } else if (r < ab) {
                v |= 1U << bit;
              } else if (r < abc) {
                u |= 1U << bit;
              } else {
                u |= 1U << bit;
                v |= 1U << bit;
              }
            }
            if (u == v) continue;
            out[f++] = (static_cast<uint64_t>(std::min(u, v)) << 32) | std::max(u, v);
          }
          filled[k] = f;
        }
      });
    }
    for (auto& th : ts) th.join();
    uint64_t w = 0;
    for (uint64_t k = 0; k < nchunks; ++k) {
      std::memmove(batch.data() + w, batch.data() + k * kChunk, filled[k] * sizeof(uint64_t));
      w += filled[k];
    }
    batch.resize(w);
    ParallelSort(batch);
    batch.erase(std::unique(batch.begin(), batch.end()), batch.end());
    const size_t before = uniq.size();
    std::vector<uint64_t> merged(uniq.size() + batch.size());
    merged.erase(std::set_union(uniq.begin(), uniq.end(), batch.begin(), batch.end(), merged.begin()), merged.end());
    uniq.swap(merged);
    next_chunk += nchunks;
    drawn += nchunks * kChunk;
    rate = static_cast<double>(uniq.size() - before) / static_cast<double>(nchunks * kChunk);
    std::fprintf(stderr, "rmat: %llu draws, %zu distinct edges\n", static_cast<unsigned long long>(drawn),
                 uniq.size());
  }
  std::vector<uint32_t> relabel(static_cast<size_t>(1) << scale, 0);
  for (uint64_t k : uniq) {
    relabel[k >> 32] = 1;
    relabel[k & 0xffffffffU] = 1;
  }
  uint32_t next = 0;
  for (auto& r : relabel) r = r ? next++ : ~0U;
  for (auto& k : uniq) {
    const uint64_t u = relabel[k >> 32], v = relabel[k & 0xffffffffU];
    k = (std::min(u, v) << 32) | std::max(u, v);
  }
  ParallelSort(uniq);
  std::fprintf(stderr, "rmat: %u nodes after dropping isolated ids\n", next);
  return uniq;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: graphgen <fa|am|yo|lj|u2|tinyN> <out.txt>\n");
    return 2;
  }
  const std::string which = argv[1];
  try {
    std::vector<uint64_t> edges;
    if (which == "fa") {
      // 4,039 nodes / 88,234 edges, gamma 2.5, max expected degree 1,045,
      // 10 planted communities, half of the second endpoints inside; seed 1.
      edges = ChungLuCommunities(4039, 88234, 2.5, 1045.0, 10, 0.5, 1);
    } else if (which == "am") {
      // 334,863 nodes / 925,872 edges, degree exponent 2.8, max degree 549,
      // community sizes exponent 1.5 in [10, 200], mixing 0.1; seed 2.
      edges = LfrStyle(334863, 925872, 2.8, 549.0, 1.5, 10.0, 200.0, 0.1, 2);
    } else if (which == "yo") {
      // 1,134,890 nodes / 2,987,624 edges, gamma 2.1, max expected degree 28,750; seed 3.
      edges = ChungLuCommunities(1134890, 2987624, 2.1, 28750.0, 0, 0.0, 3);
    } else if (which == "lj") {
      // 3,997,962 nodes / 34,681,189 edges, gamma 2.3, max expected degree
      // 15,000; 40% of edges inside 50k overlapping communities with sizes
      // exponent 1.6 in [5, 500]; seed 4.
      edges = ChungLuOverlapping(3997962, 34681189, 2.3, 15000.0, 50000, 1.6, 5.0, 500.0, 0.4, 4);
    } else if (which == "u2") {
      // R-MAT scale 24, (0.57, 0.19, 0.19, 0.05), 261.8M unique edges; seed 5.
      edges = Rmat(24, 0.57, 0.19, 0.19, 261800000ULL, 5);
    } else if (which.rfind("tiny", 0) == 0) {
      // Small Chung-Lu graphs for tests: n nodes, min(4n, n(n-1)/4) edges, gamma 2.3,
      // max expected degree n/4, 4 communities; seed = n.
      const uint64_t n = std::strtoull(which.c_str() + 4, nullptr, 10);
      if (n < 8) throw std::runtime_error("tinyN needs N >= 8");
      const uint64_t m = std::min<uint64_t>(4 * n, n * (n - 1) / 4);
      edges = ChungLuCommunities(n, m, 2.3, static_cast<double>(n) / 4.0, 4, 0.6, n);
    } else {
      throw std::runtime_error("unknown workload " + which);
    }
    const std::string out = argv[2];
    if (out.size() > 4 && out.compare(out.size() - 4, 4, ".bin") == 0) {
      WriteEdgesBinary(edges, argv[2]);
    } else {
      WriteEdges(edges, argv[2]);
    }
    std::fprintf(stderr, "%s: %zu edges -> %s\n", which.c_str(), edges.size(), argv[2]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "Error: %s\n", e.what());
    return 1;
  }
  return 0;
}
