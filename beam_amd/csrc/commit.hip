// Block commit on the device-resident quotient (g0 policy, sweg.cpp:2665-2769).
//
// The reference validates every selected pair of a block on the block-start
// state and then merges them one by one (QuotientGraph::Merge,
// quotient_graph.cpp:628-742).  The pairs are endpoint-disjoint, so the final
// state equals the order-free bulk merge (quotient_graph.cpp:759-856); it is
// built here in one launch whose work items are
//   unions    per pair: coalesce keep row + remove row through the block's
//             remove->keep map, drop the keep-remove entry into the internal
//             count, write the union (in place when it fits, else to freshly
//             bump-allocated pool space);
//   rewrites  per neighbour row of a remove (queued by the group kernel):
//             re-point remove entries to their keep, adding into an existing
//             keep entry (then the remove entry is deleted and the row
//             compacted).  Queued rows that are themselves endpoints of the
//             block are skipped: their unions already map every entry.
#include "engine.hpp"

namespace beam {

namespace {

constexpr int kThreads = 512;
constexpr int kUnionSlots = 8192;  // LDS hash slots for a union
constexpr int kSpecials = 4096;    // endpoint entries handled per rewritten row

__device__ __forceinline__ int MapRep(const CommitParams& p, int x) {
  return p.mark_dest[x] == p.epoch ? p.dest[x] : x;
}

// Locates pair k of the block: slots hold their selections contiguously.
__device__ void PairAt(const CommitParams& p, const uint32_t* slot_off, int nslots, uint32_t k, int* keep,
                       int* remove) {
  int lo = 0, hi = nslots;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (slot_off[mid] <= k) lo = mid; else hi = mid;
  }
  const uint32_t local = k - slot_off[lo];
  *keep = p.sel[2 * (lo * kMaxSelect + static_cast<int>(local))];
  *remove = p.sel[2 * (lo * kMaxSelect + static_cast<int>(local)) + 1];
}

struct CommitShared {
  uint32_t slot_off[kMaxSlotsPerBlock + 1];
  int32_t hkey[kUnionSlots];
  int32_t hcnt[kUnionSlots];
  int32_t hadd[kUnionSlots];
  int32_t sp_pos[kSpecials];
  unsigned char sp_keep[kSpecials];
  int32_t holes[kSpecials];
  int32_t lives[kSpecials];
  int distinct, write, nsp, del, nh, nl, fail;
  long long internal;
  unsigned long long dst, tbl;
};

// Union of one selected pair (QuotientGraph::Merge, quotient_graph.cpp:628-742,
// with every entry mapped through the block's remove->keep map).
__device__ bool UnionPair(const CommitParams& p, CommitShared& sh, int keep, int remove) {
  const int tid = static_cast<int>(threadIdx.x);
  const int lk = p.row_len[keep], lr = p.row_len[remove];
  const Entry* rk = p.pool + p.row_off[keep];
  const Entry* rr = p.pool + p.row_off[remove];
  uint32_t slots = 64;
  while (slots < 2u * static_cast<uint32_t>(lk + lr)) slots <<= 1;
  int32_t* tk = sh.hkey;
  int32_t* tc = sh.hcnt;
  if (tid == 0) {
    sh.distinct = 0;
    sh.write = 0;
    sh.internal = 0;
    sh.fail = 0;
    if (slots > static_cast<uint32_t>(kUnionSlots)) {
      const unsigned long long bytes = static_cast<unsigned long long>(slots) * 8;
      const unsigned long long base = atomicAdd(p.scratch_top, bytes);
      sh.tbl = base;
      if (base + bytes > p.scratch_cap) {
        sh.fail = 1;
        atomicOr(p.err, kErrScratch);
      }
    }
  }
  __syncthreads();
  if (sh.fail) return false;
  if (slots > static_cast<uint32_t>(kUnionSlots)) {
    tk = reinterpret_cast<int32_t*>(p.scratch + sh.tbl);
    tc = tk + slots;
  }
  const uint32_t mask = slots - 1;
  for (uint32_t s = tid; s < slots; s += kThreads) {
    tk[s] = -1;
    tc[s] = 0;
  }
  __syncthreads();
  for (int e = tid; e < lk + lr; e += kThreads) {
    const bool from_keep = e < lk;
    const Entry en = from_keep ? rk[e] : rr[e - lk];
    const int m = MapRep(p, en.nbr);
    if (m == keep) {
      // keep->remove count joins the internal edges once (quotient_graph.cpp:637-647).
      if (from_keep) atomicAdd(reinterpret_cast<unsigned long long*>(&sh.internal), static_cast<unsigned long long>(en.cnt));
      continue;
    }
    uint32_t h = HashInt(static_cast<uint32_t>(m)) & mask;
    while (true) {
      const int prev = atomicCAS(&tk[h], -1, m);
      if (prev == -1) {
        atomicAdd(&sh.distinct, 1);
        break;
      }
      if (prev == m) break;
      h = (h + 1) & mask;
    }
    atomicAdd(&tc[h], en.cnt);
  }
  __syncthreads();
  if (tid == 0) {
    const int n_new = sh.distinct;
    unsigned long long dst = p.row_off[keep];
    if (n_new > p.row_cap[keep]) {
      dst = atomicAdd(p.pool_top, static_cast<unsigned long long>(n_new));
      if (dst + static_cast<unsigned long long>(n_new) > p.pool_cap) {
        sh.fail = 1;
        atomicOr(p.err, kErrPool);
      }
    }
    sh.dst = dst;
  }
  __syncthreads();
  if (sh.fail) return false;
  Entry* out = p.pool + sh.dst;
  for (uint32_t s = tid; s < slots; s += kThreads) {
    const int key = tk[s];
    if (key < 0) continue;
    const int pos = atomicAdd(&sh.write, 1);
    Entry en;
    en.nbr = key;
    en.cnt = tc[s];
    out[pos] = en;
  }
  __syncthreads();
  if (tid == 0) {
    const int n_new = sh.distinct;
    if (sh.dst != p.row_off[keep]) {
      p.row_off[keep] = static_cast<uint32_t>(sh.dst);
      p.row_cap[keep] = n_new;
    }
    p.row_len[keep] = n_new;
    p.size[keep] += p.size[remove];
    p.internal[keep] += p.internal[remove] + static_cast<int32_t>(sh.internal);
    p.row_len[remove] = 0;
    p.size[remove] = 0;
    p.internal[remove] = 0;
    p.active[remove] = 0;
    atomicAdd(&p.ctr->merges, 1ULL);
  }
  __syncthreads();
  return true;
}

// Re-points one neighbour row x of this block's removes
// (quotient_graph.cpp:696-712: erase the remove entry and set the keep entry
// to the merged count).
__device__ bool RewriteRow(const CommitParams& p, CommitShared& sh, int x) {
  const int tid = static_cast<int>(threadIdx.x);
  const int lx = p.row_len[x];
  Entry* row = p.pool + p.row_off[x];
  if (tid == 0) {
    sh.nsp = 0;
    sh.del = 0;
    sh.nh = 0;
    sh.nl = 0;
    sh.fail = 0;
  }
  __syncthreads();
  for (int e = tid; e < lx; e += kThreads) {
    const int y = row[e].nbr;
    const bool is_keep = p.mark_keep[y] == p.epoch;
    if (is_keep || p.mark_dest[y] == p.epoch) {
      const int i = atomicAdd(&sh.nsp, 1);
      if (i < kSpecials) {
        sh.sp_pos[i] = e;
        sh.sp_keep[i] = is_keep ? 1 : 0;
      } else {
        sh.fail = 1;
      }
    }
  }
  __syncthreads();
  if (sh.fail) {
    if (tid == 0) atomicOr(p.err, kErrSpecials);
    return false;
  }
  const int nsp = sh.nsp;
  uint32_t slots = 64;
  while (slots < 2u * static_cast<uint32_t>(nsp)) slots <<= 1;
  const uint32_t mask = slots - 1;
  int32_t* hkey = sh.hkey;
  int32_t* hpos = sh.hcnt;
  int32_t* hadd = sh.hadd;
  for (uint32_t s = tid; s < slots; s += kThreads) {
    hkey[s] = -1;
    hadd[s] = 0;
  }
  __syncthreads();
  for (int i = tid; i < nsp; i += kThreads) {
    if (!sh.sp_keep[i]) continue;
    const int e = sh.sp_pos[i];
    const int y = row[e].nbr;
    uint32_t h = HashInt(static_cast<uint32_t>(y)) & mask;
    while (atomicCAS(&hkey[h], -1, y) != -1) h = (h + 1) & mask;
    hpos[h] = e;
  }
  __syncthreads();
  for (int i = tid; i < nsp; i += kThreads) {
    if (sh.sp_keep[i]) continue;
    const int e = sh.sp_pos[i];
    const int k = p.dest[row[e].nbr];
    uint32_t h = HashInt(static_cast<uint32_t>(k)) & mask;
    bool found = false;
    while (true) {
      const int key = hkey[h];
      if (key == k) {
        found = true;
        break;
      }
      if (key == -1) break;
      h = (h + 1) & mask;
    }
    if (found) {
      atomicAdd(&hadd[h], row[e].cnt);
      row[e].nbr = -1;
      atomicAdd(&sh.del, 1);
    } else {
      row[e].nbr = k;
    }
  }
  __syncthreads();
  for (uint32_t s = tid; s < slots; s += kThreads) {
    if (hkey[s] >= 0 && hadd[s] != 0) row[hpos[s]].cnt += hadd[s];
  }
  __syncthreads();
  const int del = sh.del;
  if (del > 0) {
    const int new_len = lx - del;
    for (int i = tid; i < nsp; i += kThreads) {
      const int e = sh.sp_pos[i];
      if (e < new_len && row[e].nbr < 0) sh.holes[atomicAdd(&sh.nh, 1)] = e;
    }
    for (int e = new_len + tid; e < lx; e += kThreads) {
      if (row[e].nbr >= 0) sh.lives[atomicAdd(&sh.nl, 1)] = e;
    }
    __syncthreads();
    for (int i = tid; i < sh.nh; i += kThreads) row[sh.holes[i]] = row[sh.lives[i]];
    __syncthreads();
    if (tid == 0) p.row_len[x] = new_len;
  }
  __syncthreads();
  return true;
}

}  // namespace

__global__ __launch_bounds__(kThreads) void CommitKernel(CommitParams p) {
  __shared__ CommitShared sh;
  const int tid = static_cast<int>(threadIdx.x);
  if (blockIdx.x == 0 && tid == 0) {
    *p.reset_top = 0;
    *p.naff_next = 0;
  }
  if (tid == 0) {
    uint32_t run = 0;
    for (int s = 0; s < p.nslots; ++s) {
      sh.slot_off[s] = run;
      run += static_cast<uint32_t>(p.sel_count[s]);
    }
    sh.slot_off[p.nslots] = run;
  }
  __syncthreads();
  const uint32_t npairs = sh.slot_off[p.nslots];
  const uint32_t naff = static_cast<uint32_t>(*p.naff);
  for (uint32_t item = blockIdx.x; item < npairs + naff; item += gridDim.x) {
    if (item < npairs) {
      int keep, remove;
      PairAt(p, sh.slot_off, p.nslots, item, &keep, &remove);
      if (!UnionPair(p, sh, keep, remove)) return;
    } else {
      const int x = p.aff[item - npairs];
      if (p.mark_dest[x] == p.epoch || p.mark_keep[x] == p.epoch) continue;
      if (!RewriteRow(p, sh, x)) return;
    }
  }
}

}  // namespace beam
