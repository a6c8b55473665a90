// incsort.hpp — device helpers of the incremental sort: windows of keys kept
// in LDS and sorted by (key, facet id) with a bitonic network.
//
// Between substeps the Morton order of a moving cloth changes very little (a
// few positions, for a small fraction of triangles).  The incremental build
// computes the keys in the previous substep's sorted order, sorts aligned
// windows of INC_MORTON_WINDOW keys and then windows of INC_WINDOW keys
// shifted by INC_SHIFT in LDS (every boundary of the first lies at least
// INC_SHIFT keys inside one of the second), and checks every adjacent pair;
// only if that check fails does the full sort run.  Because a stable sort by key is
// exactly the order by (key, facet id) — facet ids are distinct, or keys carry
// the mesh segment — the result equals the full sort's whenever the check
// passes.
#pragma once
#include "common.hpp"

namespace ysim {

constexpr uint32_t INC_WINDOW = 2048;         // keys per shifted window (power of two)
constexpr uint32_t INC_MORTON_WINDOW = 1024;  // keys per window of the Morton kernel
constexpr uint32_t INC_SHIFT = 512;           // shifted window w starts at w INC_WINDOW - INC_SHIFT
static_assert(INC_WINDOW % INC_MORTON_WINDOW == 0 && INC_SHIFT % INC_MORTON_WINDOW != 0 &&
                  INC_SHIFT <= INC_MORTON_WINDOW / 2 && INC_SHIFT < INC_WINDOW,
              "every Morton window boundary lies INC_SHIFT or more keys inside a shifted window");

template <typename K>
__device__ __forceinline__ bool kv_greater(K a, uint32_t fa, K b, uint32_t fb) {
  return a > b || (a == b && fa > fb);
}

// Sorts sk/sv[0, W) ascending by (key, value) with NT threads; entries past
// the live count must hold the maximum (key, value).  Nothing is done when the
// window is already sorted; a nearly sorted window (the common case: keys move
// a few places per substep) is finished by up to ODD_EVEN_ROUNDS rounds of
// odd-even transposition, stopping at the first round without a swap; anything
// else goes through a bitonic network.  Returns whether the window was permuted.
constexpr int ODD_EVEN_ROUNDS = 16;

template <typename K, uint32_t W, uint32_t NT = 256>
__device__ bool window_sort(K* sk, uint32_t* sv, uint32_t cnt) {
  const uint32_t tid = threadIdx.x;
  int unsorted = 0;
  for (uint32_t i = tid; i + 1 < cnt; i += NT)
    if (kv_greater(sk[i], sv[i], sk[i + 1], sv[i + 1])) unsorted = 1;
  if (!__syncthreads_or(unsorted)) return false;
  bool done = false;
  for (int round = 0; round < ODD_EVEN_ROUNDS && !done; ++round) {
    int swapped = 0;
    for (uint32_t phase = 0; phase < 2; ++phase) {
      for (uint32_t t = tid; 2 * t + phase + 1 < cnt; t += NT) {
        const uint32_t i = 2 * t + phase;
        const K a = sk[i], b = sk[i + 1];
        const uint32_t fa = sv[i], fb = sv[i + 1];
        if (kv_greater(a, fa, b, fb)) {
          sk[i] = b; sk[i + 1] = a;
          sv[i] = fb; sv[i + 1] = fa;
          swapped = 1;
        }
      }
      __syncthreads();
    }
    // a round without any swap means every adjacent pair is ordered
    done = !__syncthreads_or(swapped);
  }
  if (done) return true;
  for (uint32_t k = 2; k <= W; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = tid; t < W / 2; t += NT) {
        const uint32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        const uint32_t l = i + j;
        const bool asc = (i & k) == 0;
        const K a = sk[i], b = sk[l];
        const uint32_t fa = sv[i], fb = sv[l];
        if (kv_greater(a, fa, b, fb) == asc) {
          sk[i] = b; sk[l] = a;
          sv[i] = fb; sv[l] = fa;
        }
      }
      __syncthreads();
    }
  return true;
}

}  // namespace ysim
