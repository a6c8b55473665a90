// sort.hpp — host launchers of the LSD radix sort (sort.hip).
#pragma once
#include <algorithm>

#include "common.hpp"

namespace ysim {

constexpr uint32_t SORT_ITEMS = 8;        // keys per thread per tile (tile = 2048)
constexpr uint32_t SORT_MAX_PASSES = 8;   // 8-bit digits, keys up to 64 bits

struct SortPlan {
  uint32_t n = 0, key_bits = 0, passes = 0, tile = 0, num_blocks = 0;
  bool wide = false;  // uint64 keys
  size_t hist_words = 0, status_words = 0, counter_words = 0;
};

SortPlan sort_plan(uint32_t n, uint32_t key_bits);
// Zero the look-back state (and optionally the histogram) with kernels (graph-safe).
int launch_sort_reset(hipStream_t s, const SortPlan& p, uint32_t* hist, uint32_t* status,
                      uint32_t* counters, bool zero_hist);
int launch_sort_histogram(hipStream_t s, const SortPlan& p, const void* keys, uint32_t* hist);
// Runs p.passes passes ping-ponging a <-> b; the result is in a when passes is
// even, else in b.
int launch_sort_passes(hipStream_t s, const SortPlan& p, void* keys_a, uint32_t* vals_a,
                       void* keys_b, uint32_t* vals_b, const uint32_t* hist, uint32_t* status,
                       uint32_t* counters, uint32_t* err);

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n);

}  // namespace ysim
