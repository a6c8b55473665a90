// sort.hpp — host launchers of the LSD radix sort (sort.hip).
#pragma once
#include <algorithm>

#include "common.hpp"

namespace ysim {

constexpr uint32_t SORT_MAX_PASSES = 8;       // 8-bit digits cover 64-bit keys in 8 passes
constexpr uint32_t SORT_HIST_WORDS = 12288;   // max over digit widths of passes * 2^digit_bits

// A sort configuration: digit width (8 or 11 bits) and keys per thread (8 or
// 16; tile = 256 * items).  Every configuration yields the same permutation.
struct SortPlan {
  uint32_t n = 0, key_bits = 0, passes = 0, tile = 0, num_blocks = 0;
  uint32_t digit_bits = 8, digits = 256, items = 8;
  bool wide = false;  // uint64 keys
  size_t hist_words = 0, status_words = 0, counter_words = 0;
  // hybrid: one onesweep pass on the 11-bit digit at hist_shift (the top key
  // bits), then per-bucket LSD passes (8-bit digits) in one workgroup each
  bool hybrid = false;
  uint32_t hist_shift = 0;     // shift of the first histogrammed digit
  uint32_t local_passes = 0;   // hybrid: 8-bit passes per bucket
};

bool sort_config_ok(uint32_t digit_bits, uint32_t items);
SortPlan sort_plan(uint32_t n, uint32_t key_bits, uint32_t digit_bits = 8, uint32_t items = 8);
SortPlan sort_plan_hybrid(uint32_t n, uint32_t key_bits, uint32_t items = 8);
// Runs the hybrid sort; the result is in a.
// run_flag (optional): every kernel does nothing while *run_flag == 0.
int launch_sort_hybrid(hipStream_t s, const SortPlan& p, void* keys_a, uint32_t* vals_a, void* keys_b,
                       uint32_t* vals_b, const uint32_t* hist, uint64_t* status, uint32_t* counters,
                       const uint32_t* epoch, uint32_t* err, const uint32_t* run_flag = nullptr);
// Incremental sort, after launch_morton_inc: the half-shifted window pass and
// the check of every adjacent pair, which sets *need_full if the keys are not
// yet in (key, value) order.
// bcnt / edges: boundary arrival counters (zeroed once) and per-window end
// records, incremental_windows(n) entries each (edges: 3 words per window).
uint32_t incremental_windows(uint32_t n);
int launch_incremental_fix(hipStream_t s, bool wide, void* keys, uint32_t* vals, uint32_t n, uint32_t* need_full,
                           uint32_t* bcnt, uint64_t* edges);
// Largest status footprint (64-bit words) over all configurations for n keys.
size_t sort_status_words_max(uint32_t n, uint32_t key_bits);
// Look-back status words are 64-bit: (epoch << 32) | flag << 30 | count.  A word
// belongs to the sort whose epoch equals the device word *epoch, so the status
// array is never cleared between sorts; launch_sort_reset clears it and sets
// the epoch to 1 (for scratch memory of unknown contents).
int launch_sort_reset(hipStream_t s, const SortPlan& p, uint32_t* hist, uint64_t* status,
                      uint32_t* counters, uint32_t* epoch, bool zero_hist);
int launch_sort_histogram(hipStream_t s, const SortPlan& p, const void* keys, uint32_t* hist);
// Runs p.passes passes ping-ponging a <-> b; the result is in a when passes is
// even, else in b.
int launch_sort_passes(hipStream_t s, const SortPlan& p, void* keys_a, uint32_t* vals_a,
                       void* keys_b, uint32_t* vals_b, const uint32_t* hist, uint64_t* status,
                       uint32_t* counters, const uint32_t* epoch, uint32_t* err);

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n);
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v);

}  // namespace ysim
