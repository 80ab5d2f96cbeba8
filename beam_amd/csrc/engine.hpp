// Device-resident BEAM merge engine: state layout and kernel interfaces.
//
// HBM layout (all arrays indexed by supernode representative unless noted):
//   pool      Entry[pool_cap]   row storage; row r occupies
//                               pool[row_off[r] .. row_off[r] + row_len[r])
//                               with row_cap[r] >= row_len[r] slots reserved.
//                               Rows are unsorted: every consumer on the path
//                               only observes each row's multiset of
//                               (neighbour, count).
//   row_off   uint32, row_len / row_cap int32
//   size      int32   supernode sizes (quotient_graph.hpp:197)
//   internal  int32   internal edge counts (quotient_graph.hpp:198)
//   active    uint8
//   parent    int32   keep that absorbed r (r itself while active); node
//                     labels are resolved through it once per iteration
//   label     int32[n] node -> representative (S_, sweg.hpp:690)
//   csr       original graph (row_ptr int64[n+1], col_idx int32[arcs]) for
//             Divide's min-hash shingles (sweg.cpp:754-764)
#pragma once

#include "common.hpp"

namespace beam {

constexpr int kMaxSelect = 256;          // selected pairs per group (<= 512 / 2)
constexpr int kMaxSlotsPerBlock = 1024;  // groups per block (--group-batch-size)

struct QuotientView {
  const Entry* pool;
  const uint32_t* row_off;
  const int32_t* row_len;
  const int32_t* size;
  const int32_t* internal;
};

struct MergeParams {
  QuotientView q;
  const int32_t* order;   // position -> rep for this iteration
  const int32_t* gstart;  // group start positions
  const int32_t* glen;    // group lengths
  int ngroups;
  int first_group;        // first group of this block
  int top_k;
  double theta;           // 1.0 / (iter + 1), computed on the host
  int epoch;              // unique per block; tags dest / keep marks
  int commit;             // record selections as merges of this block
  int32_t* sel;           // [slot][kMaxSelect][2] selected (keep, remove)
  int32_t* sel_count;     // [slot]
  int32_t* mark_dest;     // epoch tag: rep is a remove of this block
  int32_t* dest;          // keep for a tagged remove
  int32_t* mark_keep;     // epoch tag: rep is a keep of this block
  int32_t* parent;
  int32_t* mark_row;      // epoch tag: neighbour row already queued
  int32_t* aff;           // queued neighbour rows of the block's removes
  int32_t* naff;
  unsigned char* scratch;
  unsigned long long scratch_cap;
  unsigned long long* scratch_top;
  unsigned long long* reset_top;  // the commit's scratch top, zeroed here
  Counters* ctr;
  int* err;
  // Optional phase trace: 16 values per group at trace[16 * group]
  // (g, E, bucket pairs, candidates, survivors, then wall-clock stamps).
  long long* trace;
  // Optional taps for tests: candidate pairs and (gain, before) per slot.
  int32_t* dbg_pairs;
  long long* dbg_gain;
  int32_t* dbg_count;
  int dbg_cap;
};

struct CommitParams {
  Entry* pool;
  uint32_t* row_off;
  int32_t* row_len;
  int32_t* row_cap;
  int32_t* size;
  int32_t* internal;
  unsigned char* active;
  const int32_t* sel;
  const int32_t* sel_count;
  int nslots;
  int epoch;
  const int32_t* mark_dest;
  const int32_t* dest;
  const int32_t* mark_keep;
  int32_t* aff;                // queued neighbour rows
  int32_t* naff;               // queue length for this block
  int32_t* naff_next;          // queue length of the next block (zeroed here)
  unsigned long long* pool_top;
  unsigned long long pool_cap;
  unsigned char* scratch;      // union hash tables that exceed LDS
  unsigned long long scratch_cap;
  unsigned long long* scratch_top;
  unsigned long long* reset_top;  // zeroed by this launch (the other pool's top)
  Counters* ctr;
  int* err;
};

__global__ void MergeGroupsKernel(MergeParams p);
__global__ void CommitKernel(CommitParams p);

}  // namespace beam
