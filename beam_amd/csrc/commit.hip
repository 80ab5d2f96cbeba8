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
constexpr int kUnionSlots = 4096;  // LDS hash slots for a union
constexpr int kSpecials = 2048;    // endpoint entries handled per rewritten row

__device__ __forceinline__ int MapRep(const CommitParams& p, int x) {
  return p.mark_dest[x] == p.epoch ? p.dest[x] : x;
}

struct CommitShared {
  uint32_t slot_off[kMaxWindowSlots + 1];
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

// Exclusive prefix of the selection counts of `nslots` slots into off[0..nslots].
__device__ void SlotPrefix(const int32_t* sel_count, int nslots, uint32_t* off, uint32_t* wave_tot) {
  const int per = (nslots + kThreads - 1) / kThreads;
  const int begin = static_cast<int>(threadIdx.x) * per;
  const int end = min(nslots, begin + per);
  uint32_t local = 0;
  for (int s = begin; s < end; ++s) local += static_cast<uint32_t>(sel_count[s]);
  uint32_t total = 0;
  uint32_t run = BlockExclusiveScan(local, wave_tot, &total);
  for (int s = begin; s < end; ++s) {
    off[s] = run;
    run += static_cast<uint32_t>(sel_count[s]);
  }
  if (threadIdx.x == 0) off[nslots] = total;
  __syncthreads();
}

__device__ __forceinline__ void WindowPair(const int32_t* sel, const uint32_t* off, int nslots, uint32_t k,
                                           int* slot, int* keep, int* remove) {
  int lo = 0, hi = nslots;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= k) lo = mid; else hi = mid;
  }
  // Skip empty slots that share the offset.
  while (lo + 1 < nslots && off[lo + 1] <= k) ++lo;
  const uint32_t local = k - off[lo];
  *slot = lo;
  *keep = sel[2 * (lo * kMaxSelect + static_cast<int>(local))];
  *remove = sel[2 * (lo * kMaxSelect + static_cast<int>(local)) + 1];
}

}  // namespace

// Speculative window resolution.  Every group of the window was scored on the
// same state; block j of the window may use its selection only if no member
// of its groups is adjacent to a keep or remove selected by blocks 0..j-1 of
// the window (those are exactly the rows and neighbour sizes the merges
// change).  All workgroups record, per node, the first window block whose
// merges touch it; the last workgroup to finish finds the first block with a
// touched member, marks and queues the merges of the blocks before it, and
// adds their statistics to the iteration counters.
__global__ __launch_bounds__(kThreads) void ResolveKernel(ResolveParams p) {
  __shared__ uint32_t off[kMaxWindowSlots + 1];
  __shared__ uint32_t wave_tot[kThreads / kWave + 1];
  __shared__ int s_last, s_j;
  __shared__ unsigned long long s_acc[kSlotStats];
  __shared__ unsigned long long s_max;
  const int tid = static_cast<int>(threadIdx.x);
  SlotPrefix(p.sel_count, p.nslots, off, wave_tot);
  const uint32_t total = off[p.nslots];
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    int slot, keep, remove;
    WindowPair(p.sel, off, p.nslots, item, &slot, &keep, &remove);
    const unsigned long long val =
        (static_cast<unsigned long long>(p.epoch) << 16) | static_cast<unsigned long long>(0xffff - slot / p.block_size);
    const int lk = p.q.row_len[keep], lr = p.q.row_len[remove];
    const Entry* rk = p.q.pool + p.q.row_off[keep];
    const Entry* rr = p.q.pool + p.q.row_off[remove];
    for (int e = tid; e < lk + lr; e += kThreads) {
      const int x = e < lk ? rk[e].nbr : rr[e - lk].nbr;
      atomicMax(&p.touch[x], val);
    }
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) s_last = atomicAdd(p.done, 1U) == gridDim.x - 1 ? 1 : 0;
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  const int nblocks = (p.nslots + p.block_size - 1) / p.block_size;
  if (tid == 0) s_j = nblocks;
  if (tid < kSlotStats) s_acc[tid] = 0;
  if (tid == 0) s_max = 0;
  __syncthreads();
  for (int s = tid; s < p.nslots; s += kThreads) {
    const int rel = s / p.block_size;
    if (rel == 0) continue;
    const int gidx = p.first_group + s;
    const int g = p.glen[gidx];
    if (g < 2) continue;
    const int start = p.gstart[gidx];
    for (int i = 0; i < g; ++i) {
      const unsigned long long t = p.touch[p.order[start + i]];
      if (static_cast<long long>(t >> 16) == p.epoch && 0xffff - static_cast<int>(t & 0xffff) < rel) {
        atomicMin(&s_j, rel);
        break;
      }
    }
  }
  __syncthreads();
  const int j = s_j;
  const int cslots = min(p.nslots, j * p.block_size);
  const uint32_t npairs = off[cslots];
  // Merge marks of the committed pairs (Update_S / QuotientGraph::Merge run
  // on them, sweg.cpp:2745-2747).
  for (uint32_t item = tid; item < npairs; item += kThreads) {
    int slot, keep, remove;
    WindowPair(p.sel, off, cslots, item, &slot, &keep, &remove);
    p.mark_dest[remove] = static_cast<int32_t>(p.epoch);
    p.dest[remove] = keep;
    p.mark_keep[keep] = static_cast<int32_t>(p.epoch);
    p.parent[remove] = keep;
    p.merge_seq[remove] = static_cast<int32_t>(p.epoch);
  }
  // Neighbours of every committed remove are queued for re-pointing.
  for (uint32_t item = tid; item < npairs; item += kThreads) {
    int slot, keep, remove;
    WindowPair(p.sel, off, cslots, item, &slot, &keep, &remove);
    const Entry* rr = p.q.pool + p.q.row_off[remove];
    for (int e = 0; e < p.q.row_len[remove]; ++e) {
      const int x = rr[e].nbr;
      if (x == keep) continue;
      if (atomicExch(&p.mark_row[x], static_cast<int32_t>(p.epoch)) != static_cast<int32_t>(p.epoch)) {
        p.aff[atomicAdd(p.naff, 1)] = x;
      }
    }
  }
  // Statistics of the committed groups.
  unsigned long long acc[kSlotStats] = {};
  unsigned long long mx = 0;
  for (int s = tid; s < cslots; s += kThreads) {
    const long long* st = p.slot_stats + static_cast<long long>(s) * kSlotStats;
    const long long g = st[kStatGroup];
    if (g < 2) continue;
    acc[kStatGroup] += 1;
    acc[kStatCand] += static_cast<unsigned long long>(st[kStatCand]);
    acc[kStatProxy] += static_cast<unsigned long long>(st[kStatProxy]);
    acc[kStatPositive] += static_cast<unsigned long long>(st[kStatPositive]);
    acc[kStatPass] += static_cast<unsigned long long>(st[kStatPass]);
    acc[kStatOverlap] += static_cast<unsigned long long>(st[kStatOverlap]);
    acc[kStatChosen] += static_cast<unsigned long long>(st[kStatChosen]);
    acc[kStatGain] += static_cast<unsigned long long>(st[kStatGain]);
    acc[kStatWork] += static_cast<unsigned long long>(st[kStatWork]);
    acc[kStatEntries] += static_cast<unsigned long long>(st[kStatEntries]);
    acc[kSlotStats - 1] += static_cast<unsigned long long>(g * (g - 1) / 2);
    acc[kSlotStats - 2] += static_cast<unsigned long long>(g);
    mx = max(mx, static_cast<unsigned long long>(g));
  }
  for (int k = 0; k < kSlotStats; ++k) {
    if (acc[k]) atomicAdd(&s_acc[k], acc[k]);
  }
  if (mx) atomicMax(&s_max, mx);
  __syncthreads();
  if (tid == 0) {
    Counters* c = p.ctr;
    c->group_count += s_acc[kStatGroup];
    c->group_max_size = max(c->group_max_size, s_max);
    c->raw_pair_count += s_acc[kSlotStats - 1];
    c->candidate_pairs_after_prune += s_acc[kStatCand];
    c->proxy_pairs_examined += s_acc[kStatProxy];
    c->exact_gain_calls += s_acc[kStatCand];
    c->positive_gain_count += s_acc[kStatPositive];
    c->above_threshold += s_acc[kStatPass];
    c->rejected_by_threshold += s_acc[kStatCand] - s_acc[kStatPass];
    c->rejected_by_overlap += s_acc[kStatOverlap];
    c->selected += s_acc[kStatChosen];
    c->total_local_gain += static_cast<long long>(s_acc[kStatGain]);
    c->validation_row_entry_work += s_acc[kStatWork];
    c->member_rows += s_acc[kSlotStats - 2];
    c->member_entries += s_acc[kStatEntries];
    *p.commit_slots = cslots;
    *p.commit_blocks = j;
    *p.done = 0;
  }
}

__global__ __launch_bounds__(kThreads) void CommitKernel(CommitParams p) {
  __shared__ CommitShared sh;
  __shared__ uint32_t wave_tot[kThreads / kWave + 1];
  const int tid = static_cast<int>(threadIdx.x);
  if (blockIdx.x == 0 && tid == 0) {
    *p.reset_top = 0;
    *p.naff_next = 0;
  }
  const int nslots = *p.nslots;
  SlotPrefix(p.sel_count, nslots, sh.slot_off, wave_tot);
  const uint32_t npairs = sh.slot_off[nslots];
  const uint32_t naff = static_cast<uint32_t>(*p.naff);
  for (uint32_t item = blockIdx.x; item < npairs + naff; item += gridDim.x) {
    if (item < npairs) {
      int slot, keep, remove;
      WindowPair(p.sel, sh.slot_off, nslots, item, &slot, &keep, &remove);
      if (!UnionPair(p, sh, keep, remove)) return;
    } else {
      const int x = p.aff[item - npairs];
      if (p.mark_dest[x] == p.epoch || p.mark_keep[x] == p.epoch) continue;
      if (!RewriteRow(p, sh, x)) return;
    }
  }
}

}  // namespace beam
