// Internal interface of the onesweep radix sort shared with the LBVH builder.
#pragma once

#include "bajo_common.h"

namespace bj {

constexpr int kSortBlock = 256;                 // 4 waves of 64
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kSortKpt = 16;                    // keys per thread
constexpr int kSortTile = kSortBlock * kSortKpt;  // 4096 keys per tile
constexpr int kHistBlock = 256;
constexpr uint32_t kLookbackAgg = 0x40000000u;
constexpr uint32_t kLookbackInc = 0x80000000u;
constexpr uint32_t kLookbackMask = 0x3FFFFFFFu;
constexpr uint64_t kSortMaxKeys = (1ull << 30) - 1;
constexpr uint32_t kLookbackSpinLimit = 1u << 24;

// Workspace layout of one sort.  All offsets are 256-byte aligned.
struct SortPlan {
  uint64_t n = 0;
  int key_bytes = 8;
  int has_values = 0;
  int begin_bit = 0, end_bit = 64, passes = 0;
  unsigned tiles = 0;
  size_t off_alt_keys = 0, off_alt_vals = 0, off_hist = 0, off_lookback = 0, off_tickets = 0, total = 0;
};

SortPlan make_sort_plan(uint64_t n, int key_bytes, int has_values, int begin_bit, int end_bit);

// Runs the digit-binning passes.  When `hist_ready` is false the plan's
// counters are cleared and the digit histogram is computed first; when true
// the caller has already cleared the tickets, zeroed the look-back table and
// filled the histogram (the LBVH Morton kernel does this in its own pass).
// Sorted output always lands back in `keys` / `values`.
int32_t onesweep_run(bajo_ctx* ctx, const SortPlan& plan, void* keys, uint32_t* values, void* ws, bool hist_ready);

}  // namespace bj
