// sort.hip — stable LSD radix sort of (key, value) pairs for gfx950.
//
// Replaces RadixSorter<METAL>::sort (include/sim/spatial_hash.hpp:48-70) and
// radixCountBlock/ScanOffset/Scatter_8Bits (src/metal/radixsort.metal:21-112):
// same result (stable by key, ties keep input order) with a different
// structure.  All digit histograms come from one read of the keys (fused into
// the Morton kernel on the scene path); each 8-bit digit pass is then ONE
// kernel: a tile is ranked with wave64 ballot matching (stable within the tile),
// tiles obtain their global digit offsets by decoupled look-back over the
// predecessors' published counts, and keys are staged in LDS in digit order so
// the global scatter writes contiguous runs.
#include "sort.hpp"

namespace ysim {

namespace {

constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_INC = 2u << 30;
constexpr uint32_t COUNT_MASK = (1u << 30) - 1;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift) {
  return (uint32_t)(key >> shift) & 255u;
}

// Exclusive scan of one value per thread over a 256-thread block (4 waves).
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_wsum[w] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (int i = 0; i < w; ++i) pre += s_wsum[i];
  __syncthreads();
  return pre + inc - v;
}

}  // namespace

// Digit histograms of every pass from one read of the keys (used by the stage
// entry point; the scene path fuses the same accumulation into Morton).
template <typename K>
__global__ __launch_bounds__(256) void k_sort_histogram(const K* __restrict__ keys, uint32_t n,
                                                        uint32_t passes, uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[SORT_MAX_PASSES * 256];
  for (uint32_t i = threadIdx.x; i < passes * 256; i += 256) s_h[i] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    K k = keys[i];
    for (uint32_t p = 0; p < passes; ++p) atomicAdd(&s_h[p * 256 + digit_of(k, 8 * p)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * 256; i += 256)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// One LSD pass.  Block = 4 waves; wave w owns a contiguous run of 64*ITEMS keys
// of the tile, processed item-row by item-row so ranks follow input order.
template <typename K, int ITEMS>
__global__ __launch_bounds__(256) void k_onesweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                  K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                  uint32_t n, uint32_t shift,
                                                  const uint32_t* __restrict__ hist,
                                                  uint32_t* __restrict__ status,
                                                  uint32_t* __restrict__ counter,
                                                  uint32_t* __restrict__ err) {
  constexpr int WAVES = 4, BLOCK = 256, TILE = BLOCK * ITEMS;
  __shared__ uint32_t s_wc[WAVES][256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_lbase[256];
  __shared__ uint32_t s_wsum[WAVES];
  __shared__ K s_keys[TILE];
  __shared__ uint32_t s_vals[TILE];
  __shared__ uint32_t s_bid;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) s_bid = atomicAdd(counter, 1u);
  for (int i = tid; i < WAVES * 256; i += BLOCK) (&s_wc[0][0])[i] = 0;
  __syncthreads();
  const uint32_t bid = s_bid;
  const uint32_t tile_base = bid * TILE;
  const uint32_t wbase = tile_base + (uint32_t)w * 64u * ITEMS;

  K key[ITEMS];
  uint32_t val[ITEMS], rank[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    uint32_t idx = wbase + (uint32_t)k * 64u + lane;
    bool ok = idx < n;
    key[k] = ok ? kin[idx] : (K)0;
    val[k] = ok ? vin[idx] : 0u;
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    uint32_t idx = wbase + (uint32_t)k * 64u + lane;
    bool ok = idx < n;
    uint32_t d = digit_of(key[k], shift);
    uint64_t match = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    uint32_t lt = (uint32_t)__popcll(match & lt_mask);
    uint32_t base = ok ? s_wc[w][d] : 0u;
    rank[k] = base + lt;
    if (ok && lt == 0) s_wc[w][d] = base + (uint32_t)__popcll(match);
  }
  __syncthreads();

  // per-digit totals, exclusive wave offsets, block-local digit offsets
  const uint32_t d = (uint32_t)tid;
  uint32_t total = 0;
#pragma unroll
  for (int ww = 0; ww < WAVES; ++ww) {
    uint32_t c = s_wc[ww][d];
    s_wc[ww][d] = total;
    total += c;
  }
  uint32_t lbase = block_excl_scan256(total, s_wsum);
  uint32_t gexcl = block_excl_scan256(hist[d], s_wsum);
  s_lbase[d] = lbase;

  // decoupled look-back for digit d
  uint32_t* st = status + (size_t)bid * 256 + d;
  st_agent_u32(st, (bid == 0 ? FLAG_INC : FLAG_AGG) | total);
  uint32_t prefix = 0;
  if (bid > 0) {
    int j = (int)bid - 1;
    uint32_t spins = 0;
    while (j >= 0) {
      uint32_t s = ld_agent_u32(status + (size_t)j * 256 + d);
      uint32_t flag = s & ~COUNT_MASK;
      if (flag == 0) {
        if (++spins > (1u << 22)) {  // wedge guard: a predecessor never published
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      prefix += s & COUNT_MASK;
      if (flag == FLAG_INC) break;
      --j;
    }
    st_agent_u32(st, FLAG_INC | (prefix + total));
  }
  s_gbase[d] = gexcl + prefix;
  __syncthreads();

  // stage in LDS in digit order, then write contiguous runs
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    uint32_t idx = wbase + (uint32_t)k * 64u + lane;
    if (idx < n) {
      uint32_t dg = digit_of(key[k], shift);
      uint32_t lp = s_lbase[dg] + s_wc[w][dg] + rank[k];
      s_keys[lp] = key[k];
      s_vals[lp] = val[k];
    }
  }
  __syncthreads();
  const uint32_t cnt = (n > tile_base) ? min((uint32_t)TILE, n - tile_base) : 0u;
  for (uint32_t i = tid; i < cnt; i += BLOCK) {
    K kk = s_keys[i];
    uint32_t dg = digit_of(kk, shift);
    uint32_t dst = s_gbase[dg] + (i - s_lbase[dg]);
    kout[dst] = kk;
    vout[dst] = s_vals[i];
  }
}

SortPlan sort_plan(uint32_t n, uint32_t key_bits) {
  SortPlan p;
  p.n = n;
  p.key_bits = key_bits;
  p.wide = key_bits > 32;
  p.passes = (key_bits + 7) / 8;
  p.tile = 256 * SORT_ITEMS;
  p.num_blocks = n ? ceil_div(n, p.tile) : 0;
  p.hist_words = SORT_MAX_PASSES * 256;
  p.status_words = (size_t)p.passes * p.num_blocks * 256;
  p.counter_words = SORT_MAX_PASSES;
  return p;
}

int launch_sort_reset(hipStream_t s, const SortPlan& p, uint32_t* hist, uint32_t* status,
                      uint32_t* counters, bool zero_hist) {
  if (zero_hist) {
    hipLaunchKernelGGL(k_zero_u32, dim3(8), dim3(256), 0, s, hist, (uint32_t)p.hist_words);
    YSIM_CHECK_LAUNCH();
  }
  if (p.status_words) {
    uint32_t blocks = std::min<uint32_t>(1024, ceil_div(p.status_words, 256));
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks), dim3(256), 0, s, status, (uint32_t)p.status_words);
    YSIM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, counters, (uint32_t)p.counter_words);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

int launch_sort_histogram(hipStream_t s, const SortPlan& p, const void* keys, uint32_t* hist) {
  if (!p.n) return YSIM_OK;
  uint32_t blocks = std::min<uint32_t>(512, ceil_div(p.n, 256 * 8));
  if (p.wide)
    hipLaunchKernelGGL(k_sort_histogram<uint64_t>, dim3(blocks), dim3(256), 0, s,
                       (const uint64_t*)keys, p.n, p.passes, hist);
  else
    hipLaunchKernelGGL(k_sort_histogram<uint32_t>, dim3(blocks), dim3(256), 0, s,
                       (const uint32_t*)keys, p.n, p.passes, hist);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

int launch_sort_passes(hipStream_t s, const SortPlan& p, void* keys_a, uint32_t* vals_a,
                       void* keys_b, uint32_t* vals_b, const uint32_t* hist, uint32_t* status,
                       uint32_t* counters, uint32_t* err) {
  if (!p.n) return YSIM_OK;
  for (uint32_t pass = 0; pass < p.passes; ++pass) {
    const bool even = (pass % 2) == 0;
    void* ki = even ? keys_a : keys_b;
    void* ko = even ? keys_b : keys_a;
    uint32_t* vi = even ? vals_a : vals_b;
    uint32_t* vo = even ? vals_b : vals_a;
    uint32_t* st = status + (size_t)pass * p.num_blocks * 256;
    if (p.wide)
      hipLaunchKernelGGL((k_onesweep<uint64_t, SORT_ITEMS>), dim3(p.num_blocks), dim3(256), 0, s,
                         (const uint64_t*)ki, vi, (uint64_t*)ko, vo, p.n, 8 * pass,
                         hist + pass * 256, st, counters + pass, err);
    else
      hipLaunchKernelGGL((k_onesweep<uint32_t, SORT_ITEMS>), dim3(p.num_blocks), dim3(256), 0, s,
                         (const uint32_t*)ki, vi, (uint32_t*)ko, vo, p.n, 8 * pass,
                         hist + pass * 256, st, counters + pass, err);
    YSIM_CHECK_LAUNCH();
  }
  return YSIM_OK;
}

}  // namespace ysim
