// Exact MAGS-compatible merge gain of one pair, computed by one wave.
#pragma once

#include "engine.hpp"

namespace beam {

constexpr int kScoreSlots = 512;  // per-wave LDS hash slots for the gain join

__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Exact MAGS-compatible merge gain of (a, b) on the current quotient, computed
// by one wave (quotient_graph.cpp:340-462).  The reference's merge-join sums
// separate into per-row streams:
//   before = sum_{x in A} BC(sa,sx,ea) + sum_{x in B} BC(sb,sx,eb) + internal terms
//   after  = sum_{x in A} BC(sw,sx,ea) + sum_{x in B} BC(sw,sx,eb)
//          + sum_{x in A and B} [BC(sw,sx,ea+eb) - BC(sw,sx,ea) - BC(sw,sx,eb)] + internal
// so both rows are streamed once (coalesced, no lookups) and only the common
// neighbours need a join: the shorter row is hashed in LDS (in chunks of
// kScoreSlots/2 when longer) and the longer row probes it.
__device__ inline void ScorePairWave(const QuotientView& q, int a, int b, int32_t* tkey, int32_t* tcnt,
                              int* err, long long* out_gain, long long* out_before) {
  const int lane = LaneId();
  const int la = q.row_len[a], lb = q.row_len[b];
  const bool a_short = la <= lb;
  const int s_rep = a_short ? a : b, l_rep = a_short ? b : a;
  const int ls = a_short ? la : lb, ll = a_short ? lb : la;
  const Entry* rs = q.pool + q.row_off[s_rep];
  const Entry* rl = q.pool + q.row_off[l_rep];
  const int64_t ss = q.size[s_rep], sl = q.size[l_rep];
  const int64_t sw = ss + sl;
  int64_t before = 0, after = 0, eab_s = 0, eab_l = 0;
  const int chunk = kScoreSlots / 2;
  for (int c0 = 0; c0 == 0 || c0 < ls; c0 += chunk) {
    const int cn = min(chunk, ls - c0);
    int slots = 64;
    while (slots < 2 * cn) slots <<= 1;
    const uint32_t mask = static_cast<uint32_t>(slots - 1);
    for (int k = lane; k < slots; k += kWave) tkey[k] = -1;
    WaveSync();
    for (int k = lane; k < cn; k += kWave) {
      const Entry e = rs[c0 + k];
      uint32_t h = HashInt(static_cast<uint32_t>(e.nbr)) & mask;
      while (true) {
        const int prev = atomicCAS(&tkey[h], -1, e.nbr);
        if (prev == -1) {
          tcnt[h] = e.cnt;
          break;
        }
        h = (h + 1) & mask;
      }
      if (e.nbr == l_rep) {
        eab_s = e.cnt;
      } else {
        const int64_t sx = q.size[e.nbr];
        before += BlockCost(ss, sx, e.cnt, false, err);
        after += BlockCost(sw, sx, e.cnt, false, err);
      }
    }
    WaveSync();
    for (int k = lane; k < ll; k += kWave) {
      const Entry e = rl[k];
      if (e.nbr == s_rep) {
        if (c0 == 0) eab_l = e.cnt;
        continue;
      }
      int64_t sx = -1;
      if (c0 == 0) {
        sx = q.size[e.nbr];
        before += BlockCost(sl, sx, e.cnt, false, err);
        after += BlockCost(sw, sx, e.cnt, false, err);
      }
      if (cn > 0) {
        uint32_t h = HashInt(static_cast<uint32_t>(e.nbr)) & mask;
        int found = -1;
        while (true) {
          const int key = tkey[h];
          if (key == e.nbr) {
            found = tcnt[h];
            break;
          }
          if (key == -1) break;
          h = (h + 1) & mask;
        }
        if (found >= 0) {
          if (sx < 0) sx = q.size[e.nbr];
          after += BlockCost(sw, sx, static_cast<int64_t>(found) + e.cnt, false, err) -
                   BlockCost(sw, sx, found, false, err) - BlockCost(sw, sx, e.cnt, false, err);
        }
      }
    }
    WaveSync();
  }
  before = WaveSum64(before);
  after = WaveSum64(after);
  eab_s = WaveSum64(eab_s);
  eab_l = WaveSum64(eab_l);
  if (lane == 0) {
    if (eab_s != eab_l) atomicOr(err, kErrAsymmetric);
    const int64_t sa = q.size[a], sb = q.size[b];
    const int64_t ia = q.internal[a], ib = q.internal[b];
    before += BlockCost(sa, sa, ia, true, err) + BlockCost(sb, sb, ib, true, err) +
              BlockCost(sa, sb, eab_s, false, err);
    after += BlockCost(sw, sw, ia + ib + eab_s, true, err);
    *out_gain = before - after;
    *out_before = before;
  }
}

}  // namespace beam
