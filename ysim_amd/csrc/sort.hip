// sort.hip — stable LSD radix sort of (key, value) pairs for gfx950.
//
// Replaces RadixSorter<METAL>::sort (include/sim/spatial_hash.hpp:48-70) and
// radixCountBlock/ScanOffset/Scatter_8Bits (src/metal/radixsort.metal:21-112):
// same result (stable by key, ties keep input order) with a different
// structure.  All digit histograms come from one read of the keys (fused into
// the Morton kernel on the scene path); each digit pass is then ONE kernel: a
// tile is ranked with wave64 ballot matching (stable within the tile), tiles
// obtain their global digit offsets by decoupled look-back over the
// predecessors' published counts, and keys are staged in LDS in digit order so
// the global scatter writes contiguous runs.  Digit width (8 or 11 bits) and
// tile size (256 x 8 or 256 x 16 keys) are template parameters chosen per
// scene by the autotuner; the permutation does not depend on them.
#include "sort.hpp"

namespace ysim {

namespace {

constexpr uint64_t FLAG_AGG = 1ull << 30;
constexpr uint64_t FLAG_INC = 2ull << 30;
constexpr uint64_t COUNT_MASK = (1ull << 30) - 1;
constexpr uint64_t FLAG_MASK = 3ull << 30;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift, uint32_t mask) {
  return (uint32_t)(key >> shift) & mask;
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
template <typename K, int DB>
__global__ __launch_bounds__(256) void k_sort_histogram(const K* __restrict__ keys, uint32_t n,
                                                        uint32_t passes, uint32_t* __restrict__ hist) {
  constexpr uint32_t DIGITS = 1u << DB;
  __shared__ uint32_t s_h[SORT_HIST_WORDS];
  for (uint32_t i = threadIdx.x; i < passes * DIGITS; i += 256) s_h[i] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    K k = keys[i];
    for (uint32_t p = 0; p < passes; ++p) atomicAdd(&s_h[p * DIGITS + digit_of(k, DB * p, DIGITS - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * DIGITS; i += 256)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_sort_epoch_one(uint32_t* epoch) { *epoch = 1u; }

// One LSD pass.  Block = 4 waves; wave w owns a contiguous run of 64*ITEMS keys
// of the tile, processed item-row by item-row so ranks follow input order.
// Thread t owns digits [t*DPT, (t+1)*DPT) for the per-digit work.
template <typename K, int DB, int ITEMS>
__global__ __launch_bounds__(256) void k_onesweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                  K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                  uint32_t n, uint32_t shift,
                                                  const uint32_t* __restrict__ hist,
                                                  uint64_t* __restrict__ status,
                                                  uint32_t* __restrict__ counter,
                                                  const uint32_t* __restrict__ epoch_p,
                                                  uint32_t* __restrict__ err) {
  constexpr int WAVES = 4, BLOCK = 256, TILE = BLOCK * ITEMS;
  constexpr uint32_t DIGITS = 1u << DB, DMASK = DIGITS - 1;
  constexpr int DPT = DIGITS / BLOCK;
  __shared__ uint32_t s_wc[WAVES][DIGITS];
  __shared__ uint32_t s_gbase[DIGITS];
  __shared__ uint32_t s_lbase[DIGITS];
  __shared__ uint32_t s_wsum[WAVES];
  __shared__ K s_keys[TILE];
  __shared__ uint32_t s_vals[TILE];
  __shared__ uint32_t s_bid;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) s_bid = atomicAdd(counter, 1u);
  for (int i = tid; i < WAVES * (int)DIGITS; i += BLOCK) (&s_wc[0][0])[i] = 0;
  __syncthreads();
  const uint32_t bid = s_bid;
  const uint64_t ep = (uint64_t)(*epoch_p) << 32;
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
    uint32_t d = digit_of(key[k], shift, DMASK);
    uint64_t match = __ballot(ok);
#pragma unroll
    for (int b = 0; b < DB; ++b) {
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

  // per-digit totals and exclusive wave offsets; block-local and global digit
  // bases by scans over the thread-owned digit runs
  uint32_t tot[DPT], gh[DPT];
  uint32_t tsum = 0, hsum = 0;
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const uint32_t d = (uint32_t)tid * DPT + q;
    uint32_t total = 0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) {
      uint32_t c = s_wc[ww][d];
      s_wc[ww][d] = total;
      total += c;
    }
    tot[q] = total;
    tsum += total;
    gh[q] = hist[d];
    hsum += gh[q];
  }
  uint32_t lrun = block_excl_scan256(tsum, s_wsum);
  uint32_t grun = block_excl_scan256(hsum, s_wsum);

  // decoupled look-back for the thread's digits (walked jointly)
  uint64_t* row = status + (size_t)bid * DIGITS + (size_t)tid * DPT;
#pragma unroll
  for (int q = 0; q < DPT; ++q) st_agent_u64(row + q, ep | (bid == 0 ? FLAG_INC : FLAG_AGG) | tot[q]);
  uint32_t prefix[DPT];
#pragma unroll
  for (int q = 0; q < DPT; ++q) prefix[q] = 0;
  if (bid > 0) {
    int jq[DPT];
#pragma unroll
    for (int q = 0; q < DPT; ++q) jq[q] = (int)bid - 1;
    uint32_t pending = (DPT >= 32) ? ~0u : ((1u << DPT) - 1u);
    uint32_t spins = 0;
    while (pending) {
      bool progressed = false;
#pragma unroll
      for (int q = 0; q < DPT; ++q) {
        if (!((pending >> q) & 1u)) continue;
        const uint64_t s = ld_agent_u64(status + (size_t)jq[q] * DIGITS + (size_t)tid * DPT + q);
        if ((s & ~(COUNT_MASK | FLAG_MASK)) != ep || (s & FLAG_MASK) == 0) continue;
        progressed = true;
        prefix[q] += (uint32_t)(s & COUNT_MASK);
        if ((s & FLAG_MASK) == FLAG_INC || --jq[q] < 0) pending &= ~(1u << q);
      }
      if (!progressed) {
        if (++spins > (1u << 22)) {  // wedge guard: a predecessor never published
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
#pragma unroll
    for (int q = 0; q < DPT; ++q) st_agent_u64(row + q, ep | FLAG_INC | (prefix[q] + tot[q]));
  }
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const uint32_t d = (uint32_t)tid * DPT + q;
    s_lbase[d] = lrun;
    s_gbase[d] = grun + prefix[q];
    lrun += tot[q];
    grun += gh[q];
  }
  __syncthreads();

  // stage in LDS in digit order, then write contiguous runs
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    uint32_t idx = wbase + (uint32_t)k * 64u + lane;
    if (idx < n) {
      uint32_t dg = digit_of(key[k], shift, DMASK);
      uint32_t lp = s_lbase[dg] + s_wc[w][dg] + rank[k];
      s_keys[lp] = key[k];
      s_vals[lp] = val[k];
    }
  }
  __syncthreads();
  const uint32_t cnt = (n > tile_base) ? min((uint32_t)TILE, n - tile_base) : 0u;
  for (uint32_t i = tid; i < cnt; i += BLOCK) {
    K kk = s_keys[i];
    uint32_t dg = digit_of(kk, shift, DMASK);
    uint32_t dst = s_gbase[dg] + (i - s_lbase[dg]);
    kout[dst] = kk;
    vout[dst] = s_vals[i];
  }
}

bool sort_config_ok(uint32_t digit_bits, uint32_t items) {
  return (digit_bits == 8 || digit_bits == 11) && (items == 8 || items == 16);
}

SortPlan sort_plan(uint32_t n, uint32_t key_bits, uint32_t digit_bits, uint32_t items) {
  SortPlan p;
  p.n = n;
  p.key_bits = key_bits;
  p.wide = key_bits > 32;
  p.digit_bits = digit_bits;
  p.digits = 1u << digit_bits;
  p.items = items;
  p.passes = (key_bits + digit_bits - 1) / digit_bits;
  p.tile = 256 * items;
  p.num_blocks = n ? ceil_div(n, p.tile) : 0;
  p.hist_words = (size_t)p.passes * p.digits;
  p.status_words = (size_t)p.passes * p.num_blocks * p.digits;
  p.counter_words = SORT_MAX_PASSES;
  return p;
}

size_t sort_status_words_max(uint32_t n, uint32_t key_bits) {
  size_t m = 0;
  for (uint32_t db : {8u, 11u})
    for (uint32_t it : {8u, 16u}) m = std::max(m, sort_plan(n, key_bits, db, it).status_words);
  return m;
}

int launch_sort_reset(hipStream_t s, const SortPlan& p, uint32_t* hist, uint64_t* status,
                      uint32_t* counters, uint32_t* epoch, bool zero_hist) {
  if (zero_hist) {
    hipLaunchKernelGGL(k_zero_u32, dim3(16), dim3(256), 0, s, hist, (uint32_t)p.hist_words);
    YSIM_CHECK_LAUNCH();
  }
  if (p.status_words) {
    const uint32_t words = (uint32_t)(2 * p.status_words);
    uint32_t blocks = std::min<uint32_t>(1024, ceil_div(words, 256));
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks), dim3(256), 0, s, (uint32_t*)status, words);
    YSIM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, counters, (uint32_t)p.counter_words);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sort_epoch_one, dim3(1), dim3(1), 0, s, epoch);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

int launch_sort_histogram(hipStream_t s, const SortPlan& p, const void* keys, uint32_t* hist) {
  if (!p.n) return YSIM_OK;
  uint32_t blocks = std::min<uint32_t>(512, ceil_div(p.n, 256 * 8));
  if (p.digit_bits == 8) {
    if (p.wide)
      hipLaunchKernelGGL((k_sort_histogram<uint64_t, 8>), dim3(blocks), dim3(256), 0, s, (const uint64_t*)keys,
                         p.n, p.passes, hist);
    else
      hipLaunchKernelGGL((k_sort_histogram<uint32_t, 8>), dim3(blocks), dim3(256), 0, s, (const uint32_t*)keys,
                         p.n, p.passes, hist);
  } else {
    if (p.wide)
      hipLaunchKernelGGL((k_sort_histogram<uint64_t, 11>), dim3(blocks), dim3(256), 0, s, (const uint64_t*)keys,
                         p.n, p.passes, hist);
    else
      hipLaunchKernelGGL((k_sort_histogram<uint32_t, 11>), dim3(blocks), dim3(256), 0, s, (const uint32_t*)keys,
                         p.n, p.passes, hist);
  }
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

namespace {
template <typename K, int DB, int ITEMS>
void launch_pass(hipStream_t s, const SortPlan& p, const void* ki, const uint32_t* vi, void* ko, uint32_t* vo,
                 uint32_t shift, const uint32_t* hist, uint64_t* st, uint32_t* counter, const uint32_t* epoch,
                 uint32_t* err) {
  hipLaunchKernelGGL((k_onesweep<K, DB, ITEMS>), dim3(p.num_blocks), dim3(256), 0, s, (const K*)ki, vi, (K*)ko,
                     vo, p.n, shift, hist, st, counter, epoch, err);
}
}  // namespace

int launch_sort_passes(hipStream_t s, const SortPlan& p, void* keys_a, uint32_t* vals_a,
                       void* keys_b, uint32_t* vals_b, const uint32_t* hist, uint64_t* status,
                       uint32_t* counters, const uint32_t* epoch, uint32_t* err) {
  if (!p.n) return YSIM_OK;
  if (!sort_config_ok(p.digit_bits, p.items)) return YSIM_ERR_INVALID;
  for (uint32_t pass = 0; pass < p.passes; ++pass) {
    const bool even = (pass % 2) == 0;
    void* ki = even ? keys_a : keys_b;
    void* ko = even ? keys_b : keys_a;
    uint32_t* vi = even ? vals_a : vals_b;
    uint32_t* vo = even ? vals_b : vals_a;
    uint64_t* st = status + (size_t)pass * p.num_blocks * p.digits;
    const uint32_t shift = p.digit_bits * pass;
    const uint32_t* h = hist + (size_t)pass * p.digits;
    uint32_t* c = counters + pass;
#define YSIM_SORT_PASS(KT, DBV, IT) launch_pass<KT, DBV, IT>(s, p, ki, vi, ko, vo, shift, h, st, c, epoch, err)
    if (p.wide) {
      if (p.digit_bits == 8) { if (p.items == 8) YSIM_SORT_PASS(uint64_t, 8, 8); else YSIM_SORT_PASS(uint64_t, 8, 16); }
      else { if (p.items == 8) YSIM_SORT_PASS(uint64_t, 11, 8); else YSIM_SORT_PASS(uint64_t, 11, 16); }
    } else {
      if (p.digit_bits == 8) { if (p.items == 8) YSIM_SORT_PASS(uint32_t, 8, 8); else YSIM_SORT_PASS(uint32_t, 8, 16); }
      else { if (p.items == 8) YSIM_SORT_PASS(uint32_t, 11, 8); else YSIM_SORT_PASS(uint32_t, 11, 16); }
    }
#undef YSIM_SORT_PASS
    YSIM_CHECK_LAUNCH();
  }
  return YSIM_OK;
}

}  // namespace ysim
