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
#include "incsort.hpp"
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
__global__ __launch_bounds__(256) void k_sort_histogram(const K* __restrict__ keys, uint32_t n, uint32_t passes,
                                                        uint32_t shift0, uint32_t* __restrict__ hist) {
  constexpr uint32_t DIGITS = 1u << DB;
  __shared__ uint32_t s_h[SORT_HIST_WORDS];
  for (uint32_t i = threadIdx.x; i < passes * DIGITS; i += 256) s_h[i] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    K k = keys[i];
    for (uint32_t p = 0; p < passes; ++p) atomicAdd(&s_h[p * DIGITS + digit_of(k, shift0 + DB * p, DIGITS - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * DIGITS; i += 256)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
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
                                                  uint32_t* __restrict__ err,
                                                  const uint32_t* __restrict__ run_flag) {
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

  if (run_flag && *run_flag == 0u) return;
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

// ------------------------------------------------------------------ hybrid
// MSD-then-local sort: one onesweep pass on the top 11 key bits partitions the
// keys into 2048 buckets (stably), then k_bucket_sort gives each bucket one
// workgroup that sorts it by the remaining bits with stable 8-bit LSD passes.
// A bucket of up to 256 * LITEMS keys stays in LDS for all of its passes; a
// larger one runs the same passes over global memory in chunks of that size,
// ping-ponging between the two buffers inside its own range.  Sorting each
// bucket stably by the low bits after a stable partition by the high bits is
// the same permutation as the full LSD sort.

namespace {
// Stable within-chunk ranks of one wave's rows: rank[k] = number of earlier
// keys of the wave with the same 8-bit digit; wc[] counts per digit.
template <typename K, int ITEMS>
__device__ __forceinline__ void rank_rows8(const K (&key)[ITEMS], uint32_t base_idx, uint32_t limit, uint32_t shift,
                                           uint32_t* wc, uint32_t (&rank)[ITEMS]) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t idx = base_idx + (uint32_t)k * 64u + lane;
    const bool ok = idx < limit;
    const uint32_t d = digit_of(key[k], shift, 255u);
    uint64_t match = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    const uint32_t lt = (uint32_t)__popcll(match & lt_mask);
    const uint32_t base = ok ? wc[d] : 0u;
    rank[k] = base + lt;
    if (ok && lt == 0) wc[d] = base + (uint32_t)__popcll(match);
  }
}
}  // namespace

template <typename K, int LITEMS>
__global__ __launch_bounds__(256) void k_bucket_sort(K* __restrict__ ka, uint32_t* __restrict__ va,
                                                     K* __restrict__ kb, uint32_t* __restrict__ vb,
                                                     const uint32_t* __restrict__ msd_hist, uint32_t local_passes,
                                                     const uint32_t* __restrict__ run_flag) {
  constexpr int WAVES = 4, BLOCK = 256, CAP = BLOCK * LITEMS;
  __shared__ uint32_t s_wc[WAVES][256];
  __shared__ uint32_t s_lbase[256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_tot[256];
  __shared__ uint32_t s_wsum[WAVES];
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_vals[CAP];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (run_flag && *run_flag == 0u) return;
  const uint32_t bucket = blockIdx.x;
  const uint32_t S = msd_hist[bucket];
  if (S == 0) return;
  // bucket start: sum of the preceding bucket sizes
  uint32_t part = 0;
  for (uint32_t i = tid; i < bucket; i += BLOCK) part += msd_hist[i];
  const uint32_t start = block_excl_scan256(part, s_wsum) + part;  // inclusive for the last thread
  __shared__ uint32_t s_start;
  if (tid == BLOCK - 1) s_start = start;
  __syncthreads();
  const uint32_t st = s_start;
  const uint32_t wbase = (uint32_t)w * 64u * LITEMS;

  // per-digit totals of the staged rows, exclusive wave offsets in s_wc,
  // chunk-local digit bases in s_lbase, totals in s_tot
  auto digit_bases = [&]() {
    const uint32_t d = (uint32_t)tid;
    uint32_t total = 0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) {
      const uint32_t c = s_wc[ww][d];
      s_wc[ww][d] = total;
      total += c;
    }
    s_tot[d] = total;
    s_lbase[d] = block_excl_scan256(total, s_wsum);
  };

  if (S <= (uint32_t)CAP) {
    // the bucket (the MSD pass left it in b) lives in LDS for every pass
    for (uint32_t i = tid; i < S; i += BLOCK) {
      s_keys[i] = kb[st + i];
      s_vals[i] = vb[st + i];
    }
    __syncthreads();
    for (uint32_t pass = 0; pass < local_passes; ++pass) {
      K key[LITEMS];
      uint32_t val[LITEMS], rank[LITEMS];
#pragma unroll
      for (int k = 0; k < LITEMS; ++k) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + lane;
        key[k] = idx < S ? s_keys[idx] : (K)0;
        val[k] = idx < S ? s_vals[idx] : 0u;
      }
      for (int i = tid; i < WAVES * 256; i += BLOCK) (&s_wc[0][0])[i] = 0;
      __syncthreads();
      rank_rows8<K, LITEMS>(key, wbase, S, 8u * pass, s_wc[w], rank);
      __syncthreads();
      digit_bases();
      __syncthreads();
#pragma unroll
      for (int k = 0; k < LITEMS; ++k) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + lane;
        if (idx < S) {
          const uint32_t dg = digit_of(key[k], 8u * pass, 255u);
          const uint32_t lp = s_lbase[dg] + s_wc[w][dg] + rank[k];
          s_keys[lp] = key[k];
          s_vals[lp] = val[k];
        }
      }
      __syncthreads();
    }
    for (uint32_t i = tid; i < S; i += BLOCK) {
      ka[st + i] = s_keys[i];
      va[st + i] = s_vals[i];
    }
    return;
  }

  // large bucket: chunked passes over global memory, b -> a -> b ...
  K* ksrc = kb;
  uint32_t* vsrc = vb;
  K* kdst = ka;
  uint32_t* vdst = va;
  for (uint32_t pass = 0; pass < local_passes; ++pass) {
    const uint32_t shift = 8u * pass;
    s_tot[tid] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < S; i += BLOCK) atomicAdd(&s_tot[digit_of(ksrc[st + i], shift, 255u)], 1u);
    __syncthreads();
    const uint32_t cnt = s_tot[tid];
    s_gbase[tid] = st + block_excl_scan256(cnt, s_wsum);
    __syncthreads();
    for (uint32_t c0 = 0; c0 < S; c0 += CAP) {
      const uint32_t cs = min((uint32_t)CAP, S - c0);
      K key[LITEMS];
      uint32_t val[LITEMS], rank[LITEMS];
#pragma unroll
      for (int k = 0; k < LITEMS; ++k) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + lane;
        key[k] = idx < cs ? ksrc[st + c0 + idx] : (K)0;
        val[k] = idx < cs ? vsrc[st + c0 + idx] : 0u;
      }
      for (int i = tid; i < WAVES * 256; i += BLOCK) (&s_wc[0][0])[i] = 0;
      __syncthreads();
      rank_rows8<K, LITEMS>(key, wbase, cs, shift, s_wc[w], rank);
      __syncthreads();
      digit_bases();
      __syncthreads();
#pragma unroll
      for (int k = 0; k < LITEMS; ++k) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + lane;
        if (idx < cs) {
          const uint32_t dg = digit_of(key[k], shift, 255u);
          const uint32_t lp = s_lbase[dg] + s_wc[w][dg] + rank[k];
          s_keys[lp] = key[k];
          s_vals[lp] = val[k];
        }
      }
      __syncthreads();
      for (uint32_t i = tid; i < cs; i += BLOCK) {
        const K kk = s_keys[i];
        const uint32_t dg = digit_of(kk, shift, 255u);
        const uint32_t dst = s_gbase[dg] + (i - s_lbase[dg]);
        kdst[dst] = kk;
        vdst[dst] = s_vals[i];
      }
      __syncthreads();
      s_gbase[tid] += s_tot[tid];
      __syncthreads();
    }
    K* kt = ksrc; ksrc = kdst; kdst = kt;
    uint32_t* vt = vsrc; vsrc = vdst; vdst = vt;
  }
  if (ksrc != ka) {  // result must end in a
    for (uint32_t i = tid; i < S; i += BLOCK) {
      ka[st + i] = ksrc[st + i];
      va[st + i] = vsrc[st + i];
    }
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

SortPlan sort_plan_hybrid(uint32_t n, uint32_t key_bits, uint32_t items) {
  SortPlan p = sort_plan(n, key_bits, 11, items);
  p.hybrid = true;
  p.passes = 1;
  p.hist_shift = key_bits > 11 ? key_bits - 11 : 0;
  p.local_passes = key_bits > 11 ? (key_bits - 11 + 7) / 8 : 0;
  p.hist_words = p.digits;
  p.status_words = (size_t)p.num_blocks * p.digits;
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
                         p.n, p.passes, p.hist_shift, hist);
    else
      hipLaunchKernelGGL((k_sort_histogram<uint32_t, 8>), dim3(blocks), dim3(256), 0, s, (const uint32_t*)keys,
                         p.n, p.passes, p.hist_shift, hist);
  } else {
    if (p.wide)
      hipLaunchKernelGGL((k_sort_histogram<uint64_t, 11>), dim3(blocks), dim3(256), 0, s, (const uint64_t*)keys,
                         p.n, p.passes, p.hist_shift, hist);
    else
      hipLaunchKernelGGL((k_sort_histogram<uint32_t, 11>), dim3(blocks), dim3(256), 0, s, (const uint32_t*)keys,
                         p.n, p.passes, p.hist_shift, hist);
  }
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

namespace {
template <typename K, int DB, int ITEMS>
void launch_pass(hipStream_t s, const SortPlan& p, const void* ki, const uint32_t* vi, void* ko, uint32_t* vo,
                 uint32_t shift, const uint32_t* hist, uint64_t* st, uint32_t* counter, const uint32_t* epoch,
                 uint32_t* err, const uint32_t* run_flag = nullptr) {
  hipLaunchKernelGGL((k_onesweep<K, DB, ITEMS>), dim3(p.num_blocks), dim3(256), 0, s, (const K*)ki, vi, (K*)ko,
                     vo, p.n, shift, hist, st, counter, epoch, err, run_flag);
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

int launch_sort_hybrid(hipStream_t s, const SortPlan& p, void* keys_a, uint32_t* vals_a, void* keys_b,
                       uint32_t* vals_b, const uint32_t* hist, uint64_t* status, uint32_t* counters,
                       const uint32_t* epoch, uint32_t* err, const uint32_t* run_flag) {
  if (!p.n) return YSIM_OK;
  if (!p.hybrid || !sort_config_ok(11, p.items)) return YSIM_ERR_INVALID;
  const uint32_t sh = p.hist_shift;
  if (p.wide) {
    if (p.items == 8)
      launch_pass<uint64_t, 11, 8>(s, p, keys_a, vals_a, keys_b, vals_b, sh, hist, status, counters, epoch, err, run_flag);
    else
      launch_pass<uint64_t, 11, 16>(s, p, keys_a, vals_a, keys_b, vals_b, sh, hist, status, counters, epoch, err, run_flag);
  } else {
    if (p.items == 8)
      launch_pass<uint32_t, 11, 8>(s, p, keys_a, vals_a, keys_b, vals_b, sh, hist, status, counters, epoch, err, run_flag);
    else
      launch_pass<uint32_t, 11, 16>(s, p, keys_a, vals_a, keys_b, vals_b, sh, hist, status, counters, epoch, err, run_flag);
  }
  YSIM_CHECK_LAUNCH();
  if (p.wide)
    hipLaunchKernelGGL((k_bucket_sort<uint64_t, 16>), dim3(2048), dim3(256), 0, s, (uint64_t*)keys_a, vals_a,
                       (uint64_t*)keys_b, vals_b, hist, p.local_passes, run_flag);
  else
    hipLaunchKernelGGL((k_bucket_sort<uint32_t, 16>), dim3(2048), dim3(256), 0, s, (uint32_t*)keys_a, vals_a,
                       (uint32_t*)keys_b, vals_b, hist, p.local_passes, run_flag);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}


// ------------------------------------------------------------------ incremental
// Second window pass of the incremental sort: windows straddling the Morton
// kernel's window boundaries (block b covers [b W - INC_SHIFT, (b + 1) W -
// INC_SHIFT), W = INC_WINDOW), sorted in place,
// then the exact check.  Every adjacent pair inside a window is ordered once
// the window is sorted, so only the pairs across the shifted windows'
// boundaries remain: each block publishes its window's first and last (key,
// facet) and arrives at its two boundaries (counters advance by two per pass,
// so they need no clearing); the second arrival compares the pair and
// requests the full sort if it is not strictly increasing.
// 1024 threads per window: two keys per thread keep enough loads in flight
// for one window per CU.
constexpr uint32_t WS_THREADS = 1024;
template <typename K>
__global__ __launch_bounds__(WS_THREADS) void k_window_sort_shifted(K* __restrict__ keys, uint32_t* __restrict__ vals,
                                                             uint32_t n, uint32_t* __restrict__ need_full,
                                                             uint32_t* __restrict__ bcnt, uint64_t* edges) {
  __shared__ K sk[INC_WINDOW];
  __shared__ uint32_t sv[INC_WINDOW];
  const uint32_t w = blockIdx.x;
  const int64_t s0 = (int64_t)w * INC_WINDOW - INC_SHIFT;
  const uint32_t lo = (uint32_t)(s0 < 0 ? 0 : s0);
  const uint32_t hi = (uint32_t)min<int64_t>(s0 + INC_WINDOW, (int64_t)n);
  if (lo >= hi) return;
  const uint32_t cnt = hi - lo;
  for (uint32_t i = threadIdx.x; i < INC_WINDOW; i += WS_THREADS) {
    sk[i] = i < cnt ? keys[lo + i] : (K)~(K)0;
    sv[i] = i < cnt ? vals[lo + i] : ~0u;
  }
  __syncthreads();
  if (window_sort<K, INC_WINDOW, WS_THREADS>(sk, sv, cnt))
    for (uint32_t i = threadIdx.x; i < cnt; i += WS_THREADS) {
      keys[lo + i] = sk[i];
      vals[lo + i] = sv[i];
    }
  if (threadIdx.x != 0) return;
  const uint32_t nw = (n + INC_SHIFT + INC_WINDOW - 1) / INC_WINDOW;  // non-empty windows
  uint64_t* e = edges + 3 * (size_t)w;
  st_agent_u64(e + 0, (uint64_t)sk[0]);
  st_agent_u64(e + 1, (uint64_t)sk[cnt - 1]);
  st_agent_u64(e + 2, (uint64_t)sv[0] | ((uint64_t)sv[cnt - 1] << 32));
  wait_all_mem();
  // both arrivals go out together, then the second arrivals' neighbour loads
  // (three round trips after the window instead of up to five)
  const bool okl = w >= 1 && w < nw, okr = w + 1 < nw;  // boundaries w - 1 and w
  const uint32_t oldl = okl ? __hip_atomic_fetch_add(bcnt + (w - 1), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
  const uint32_t oldr = okr ? __hip_atomic_fetch_add(bcnt + w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
  const bool chkl = okl && (oldl & 1u), chkr = okr && (oldr & 1u);  // second arrivals check
  uint64_t lk = 0, lv = 0, rk = 0, rv = 0;
  if (chkl) {  // window w - 1's last (key, facet)
    const uint64_t* o = edges + 3 * (size_t)(w - 1);
    lk = ld_agent_u64(o + 1);
    lv = ld_agent_u64(o + 2);
  }
  if (chkr) {  // window w + 1's first (key, facet)
    const uint64_t* o = edges + 3 * (size_t)(w + 1);
    rk = ld_agent_u64(o + 0);
    rv = ld_agent_u64(o + 2);
  }
  if (chkl && !kv_greater(sk[0], sv[0], (K)lk, (uint32_t)(lv >> 32))) atomicOr(need_full, 1u);
  if (chkr && !kv_greater((K)rk, (uint32_t)rv, sk[cnt - 1], sv[cnt - 1])) atomicOr(need_full, 1u);
}

uint32_t incremental_windows(uint32_t n) { return n / INC_WINDOW + 2; }

int launch_incremental_fix(hipStream_t s, bool wide, void* keys, uint32_t* vals, uint32_t n, uint32_t* need_full,
                           uint32_t* bcnt, uint64_t* edges) {
  if (!n) return YSIM_OK;
  const uint32_t wblocks = ceil_div(n, INC_WINDOW) + 1;
  if (wide)
    hipLaunchKernelGGL(k_window_sort_shifted<uint64_t>, dim3(wblocks), dim3(WS_THREADS), 0, s, (uint64_t*)keys, vals, n,
                       need_full, bcnt, edges);
  else
    hipLaunchKernelGGL(k_window_sort_shifted<uint32_t>, dim3(wblocks), dim3(WS_THREADS), 0, s, (uint32_t*)keys, vals, n,
                       need_full, bcnt, edges);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

}  // namespace ysim
