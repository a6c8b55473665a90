// lbvh.hip — LBVH construction and refit for gfx950.
//
//   k_box       S0  scene box per mesh (BVH::build CPU loop, bvh.hpp:1229-1231)
//   k_morton    S1  Morton keys (fillMortons_Tri, bvh.metal:42-63) + sort histograms
//   k_hierarchy S3  Karras hierarchy (buildTree_Tri, bvh.metal:197-258, 352-393)
//   k_refit     S4+S5 leaf boxes (buildLeaf/buildSweptLeaf_Tri, bvh.metal:260-351)
//                   + atomic bottom-up union (bottomUpBoxes, bvh.metal:464-586)
//
// All kernels are segmented (one segment per mesh) and write each mesh's tree
// into its own node window with mesh-local child indices, so every tree is
// identical to a standalone per-mesh tree of the reference.
#include <algorithm>
#include <cmath>

#include "incsort.hpp"
#include "kernels.hpp"

namespace ysim {

namespace {

constexpr uint32_t BOX_ITEMS = 8;
constexpr uint32_t MORTON_ITEMS = 8;

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ void flush_box(uint32_t* acc, uint32_t seg, f3 lo, f3 hi) {
  uint32_t* a = acc + seg * 6;
  atomicMin(a + 0, fkey(lo.x));
  atomicMin(a + 1, fkey(lo.y));
  atomicMin(a + 2, fkey(lo.z));
  atomicMax(a + 3, fkey(hi.x));
  atomicMax(a + 4, fkey(hi.y));
  atomicMax(a + 5, fkey(hi.z));
}

}  // namespace

// ------------------------------------------------------------------ S0 box
__global__ __launch_bounds__(256) void k_box(SegTable vs, MeshTables mt, const float* __restrict__ x,
                                             uint32_t n, uint32_t* __restrict__ acc) {
  __shared__ float s_red[6][4];
  const uint32_t start = blockIdx.x * (256 * BOX_ITEMS);
  if (start >= n) return;
  const uint32_t end = min(start + 256 * BOX_ITEMS, n);
  const uint32_t s0 = seg_of(vs, start), s1 = seg_of(vs, end - 1);
  const float INF = __uint_as_float(0x7f800000u);
  if (s0 == s1) {
    const uint32_t gbase = mt.voff[vs.mesh[s0]] - vs.off[s0];
    f3 lo = {INF, INF, INF}, hi = {-INF, -INF, -INF};
#pragma unroll
    for (uint32_t i = 0; i < BOX_ITEMS; ++i) {
      uint32_t j = start + threadIdx.x + 256 * i;
      if (j < end) {
        f3 p = ld3(x, gbase + j);
        lo = vmin(lo, p);
        hi = vmax(hi, p);
      }
    }
    float r[6] = {wave_min(lo.x), wave_min(lo.y), wave_min(lo.z), wave_max(hi.x), wave_max(hi.y), wave_max(hi.z)};
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
      for (int k = 0; k < 6; ++k) s_red[k][w] = r[k];
    __syncthreads();
    if (threadIdx.x == 0) {
      f3 L = {s_red[0][0], s_red[1][0], s_red[2][0]}, H = {s_red[3][0], s_red[4][0], s_red[5][0]};
      for (int ww = 1; ww < 4; ++ww) {
        L = vmin(L, f3{s_red[0][ww], s_red[1][ww], s_red[2][ww]});
        H = vmax(H, f3{s_red[3][ww], s_red[4][ww], s_red[5][ww]});
      }
      flush_box(acc, s0, L, H);
    }
    return;
  }
  // block spans several meshes: per-thread running accumulator, flushed on change
  uint32_t cur = 0xffffffffu;
  f3 lo = {INF, INF, INF}, hi = {-INF, -INF, -INF};
  for (uint32_t i = 0; i < BOX_ITEMS; ++i) {
    uint32_t j = start + threadIdx.x + 256 * i;
    if (j >= end) break;
    uint32_t sg = seg_of(vs, j);
    if (sg != cur) {
      if (cur != 0xffffffffu) flush_box(acc, cur, lo, hi);
      cur = sg;
      lo = f3{INF, INF, INF};
      hi = f3{-INF, -INF, -INF};
    }
    f3 p = ld3(x, mt.voff[vs.mesh[sg]] + (j - vs.off[sg]));
    lo = vmin(lo, p);
    hi = vmax(hi, p);
  }
  if (cur != 0xffffffffu) flush_box(acc, cur, lo, hi);
}

// One mesh: 16 K vertices per 1024-thread block, so a 262 K-vertex mesh makes
// 16 blocks and 96 atomics on the six box words (same-address atomics from
// many workgroups serialize at the memory side).
constexpr uint32_t BOX1_THREADS = 1024, BOX1_ITEMS = 16;
__global__ __launch_bounds__(BOX1_THREADS) void k_box1(MeshTables mt, uint32_t mesh, const float* __restrict__ x,
                                                      uint32_t n, uint32_t* __restrict__ acc) {
  __shared__ float s_red[6][BOX1_THREADS / 64];
  const uint32_t start = blockIdx.x * (BOX1_THREADS * BOX1_ITEMS);
  const uint32_t end = min(start + BOX1_THREADS * BOX1_ITEMS, n);
  const uint32_t gbase = mt.voff[mesh];
  const float INF = __uint_as_float(0x7f800000u);
  f3 lo = {INF, INF, INF}, hi = {-INF, -INF, -INF};
#pragma unroll
  for (uint32_t i = 0; i < BOX1_ITEMS; ++i) {
    const uint32_t j = start + threadIdx.x + BOX1_THREADS * i;
    if (j < end) {
      const f3 p = ld3(x, gbase + j);
      lo = vmin(lo, p);
      hi = vmax(hi, p);
    }
  }
  const float r[6] = {wave_min(lo.x), wave_min(lo.y), wave_min(lo.z), wave_max(hi.x), wave_max(hi.y), wave_max(hi.z)};
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) s_red[k][w] = r[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const uint32_t k = threadIdx.x;
    float v = s_red[k][0];
    for (uint32_t ww = 1; ww < BOX1_THREADS / 64; ++ww) v = k < 3 ? fminf(v, s_red[k][ww]) : fmaxf(v, s_red[k][ww]);
    if (k < 3) atomicMin(acc + k, fkey(v));
    else atomicMax(acc + k, fkey(v));
  }
}

// One mesh over every CU: 1024 vertices per 256-thread block, the blocks'
// boxes reduced by last arrivals (grid_box_reduce).
constexpr uint32_t BOXP_THREADS = 256, BOXP_ITEMS = 4;
__global__ __launch_bounds__(BOXP_THREADS) void k_box_parts(MeshTables mt, uint32_t mesh, const float* __restrict__ x,
                                                            uint32_t n, uint32_t* parts, uint32_t* __restrict__ acc) {
  const uint32_t start = blockIdx.x * (BOXP_THREADS * BOXP_ITEMS);
  const uint32_t end = min(start + BOXP_THREADS * BOXP_ITEMS, n);
  const uint32_t gbase = mt.voff[mesh];
  const float INF = __uint_as_float(0x7f800000u);
  f3 lo = {INF, INF, INF}, hi = {-INF, -INF, -INF};
#pragma unroll
  for (uint32_t i = 0; i < BOXP_ITEMS; ++i) {
    const uint32_t j = start + threadIdx.x + BOXP_THREADS * i;
    if (j < end) {
      const f3 p = ld3(x, gbase + j);
      lo = vmin(lo, p);
      hi = vmax(hi, p);
    }
  }
  grid_box_reduce<BOXP_THREADS>(lo, hi, parts, acc);
}

uint32_t box_parts_words(uint32_t nverts) {
  return grid_box_parts_words(ceil_div(nverts, BOXP_THREADS * BOXP_ITEMS));
}

int launch_box(hipStream_t s, SegTable vsegs, MeshTables mt, const float* x, uint32_t nitems,
               uint32_t* box_acc, uint32_t single_mesh, uint32_t* parts) {
  if (!nitems) return YSIM_OK;
  if (single_mesh != 0xffffffffu && parts && grid_box_fits(ceil_div(nitems, BOXP_THREADS * BOXP_ITEMS))) {
    hipLaunchKernelGGL(k_box_parts, dim3(ceil_div(nitems, BOXP_THREADS * BOXP_ITEMS)), dim3(BOXP_THREADS), 0, s, mt,
                       single_mesh, x, nitems, parts, box_acc);
    YSIM_CHECK_LAUNCH();
    return YSIM_OK;
  }
  if (single_mesh != 0xffffffffu)
    hipLaunchKernelGGL(k_box1, dim3(ceil_div(nitems, BOX1_THREADS * BOX1_ITEMS)), dim3(BOX1_THREADS), 0, s, mt,
                       single_mesh, x, nitems, box_acc);
  else
    hipLaunchKernelGGL(k_box, dim3(ceil_div(nitems, 256 * BOX_ITEMS)), dim3(256), 0, s, vsegs, mt, x, nitems,
                       box_acc);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// ------------------------------------------------------------------ S1 Morton
// NH histograms of DB-bit digits at shifts hist_shift + DB * p (p < passes <= NH).
template <typename K, int DB, int NH>
__global__ __launch_bounds__(256) void k_morton(SegTable ts, MeshTables mt, const float* __restrict__ x,
                                                const uint32_t* __restrict__ facets,
                                                const uint32_t* __restrict__ acc, uint32_t code_bits,
                                                K* __restrict__ keys, uint32_t* __restrict__ vals,
                                                uint32_t n, uint32_t passes, uint32_t hist_shift,
                                                uint32_t* __restrict__ hist,
                                                uint32_t* __restrict__ counters, uint32_t* __restrict__ epoch,
                                                const uint32_t* __restrict__ run_flag,
                                                uint32_t* __restrict__ cut_reset, int32_t* __restrict__ cut_fill) {
  constexpr uint32_t DIGITS = 1u << DB;
  constexpr uint32_t HW = NH * DIGITS;
  __shared__ uint32_t s_h[HW];
  if (run_flag && *run_flag == 0u) return;  // the incremental sort succeeded: nothing to do
  if (cut_reset && blockIdx.x == 0) {
    for (uint32_t t = threadIdx.x; t < ts.nseg; t += 256) cut_reset[ts.mesh[t]] = 0u;
    if (cut_fill && ts.nseg == 1)
      for (uint32_t k = threadIdx.x; k < CUT_CAP; k += 256) cut_fill[(size_t)ts.mesh[0] * CUT_CAP + k] = -1;
  }
  for (uint32_t i = threadIdx.x; i < passes * DIGITS; i += 256) s_h[i] = 0;
  // the next sort's block counters start at zero and its look-back words carry
  // a fresh epoch (the sort passes run after this kernel completes)
  if (blockIdx.x == 0 && threadIdx.x < SORT_MAX_PASSES && counters) counters[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0 && epoch) epoch[0] = epoch[0] + 1u;
  __syncthreads();
  const bool wide = code_bits > 32;
  const uint32_t start = blockIdx.x * (256 * MORTON_ITEMS);
  for (uint32_t i = 0; i < MORTON_ITEMS; ++i) {
    uint32_t j = start + threadIdx.x + 256 * i;
    if (j >= n) break;
    uint32_t sg = seg_of(ts, j);
    uint32_t m = ts.mesh[sg];
    uint32_t fid = j - ts.off[sg];
    const uint32_t* f = facets + 3 * (size_t)(mt.foff[m] + fid);
    uint32_t vb = mt.voff[m];
    f3 x0 = ld3(x, vb + f[0]), x1 = ld3(x, vb + f[1]), x2 = ld3(x, vb + f[2]);
    const uint32_t* a = acc + sg * 6;
    f3 bmin = {funkey(a[0]), funkey(a[1]), funkey(a[2])};
    f3 bmax = {funkey(a[3]), funkey(a[4]), funkey(a[5])};
    uint64_t code = morton_of(x0, x1, x2, bmin, bmax, wide);
    K key = (ts.nseg > 1) ? (K)(((uint64_t)sg << code_bits) | code) : (K)code;
    keys[j] = key;
    vals[j] = fid;
    for (uint32_t p = 0; p < passes; ++p)
      atomicAdd(&s_h[p * DIGITS + (uint32_t)((key >> (hist_shift + DB * p)) & (DIGITS - 1))], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * DIGITS; i += 256)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

int launch_morton(hipStream_t s, SegTable tsegs, MeshTables mt, const float* x,
                  const uint32_t* facets, const uint32_t* box_acc, uint32_t code_bits, bool key64,
                  void* keys, uint32_t* vals, uint32_t nitems, const SortPlan& plan, uint32_t* hist,
                  uint32_t* counters, uint32_t* epoch, const uint32_t* run_flag, uint32_t* cut_reset,
                  int32_t* cut_fill) {
  if (!nitems) return YSIM_OK;
  uint32_t blocks = ceil_div(nitems, 256 * MORTON_ITEMS);
#define YSIM_MORTON(KT, DBV, NHV)                                                                             \
  hipLaunchKernelGGL((k_morton<KT, DBV, NHV>), dim3(blocks), dim3(256), 0, s, tsegs, mt, x, facets, box_acc, \
                     code_bits, (KT*)keys, vals, nitems, plan.passes, plan.hist_shift, hist, counters, epoch, \
                     run_flag, cut_reset, cut_fill)
  if (plan.hybrid) {
    if (key64) YSIM_MORTON(uint64_t, 11, 1); else YSIM_MORTON(uint32_t, 11, 1);
  } else if (plan.digit_bits == 11) {
    if (key64) YSIM_MORTON(uint64_t, 11, 6); else YSIM_MORTON(uint32_t, 11, 6);
  } else {
    if (key64) YSIM_MORTON(uint64_t, 8, 8); else YSIM_MORTON(uint32_t, 8, 8);
  }
#undef YSIM_MORTON
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// S1 of the incremental build: Morton keys in the previous substep's sorted
// order (perm[j] = facet at sorted position j; the positions of a mesh stay in
// its segment because keys carry the segment above the code bits), each block's
// INC_MORTON_WINDOW positions then sorted in LDS by (key, facet).  Clears
// need_full for the check that follows.
// 512 threads per 1024-key window, two blocks per CU: one block's code
// arithmetic and sort overlap the other's chain of dependent loads (order ->
// facet -> positions); one 1024-thread block per CU ran those phases in
// lockstep.
constexpr uint32_t INC_THREADS = 512;

template <typename K>
__global__ __launch_bounds__(INC_THREADS) void k_morton_inc(SegTable ts, MeshTables mt, const float* __restrict__ x,
                                                    const uint32_t* __restrict__ facets, uint32_t* acc,
                                                    uint32_t code_bits, const uint32_t* __restrict__ perm,
                                                    K* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t n,
                                                    uint32_t* __restrict__ need_full,
                                                    const uint32_t* __restrict__ box_parts, uint32_t nparts,
                                                    uint32_t* __restrict__ cut_reset, int32_t* __restrict__ cut_fill) {
  __shared__ K sk[INC_MORTON_WINDOW];
  __shared__ uint32_t sv[INC_MORTON_WINDOW];
  __shared__ uint32_t s_part[INC_THREADS];
  __shared__ uint32_t s_box[6];
  if (blockIdx.x == 0 && threadIdx.x == 0) *need_full = 0u;
  if (cut_reset && blockIdx.x == 0) {
    for (uint32_t t = threadIdx.x; t < ts.nseg; t += INC_THREADS) cut_reset[ts.mesh[t]] = 0u;
    if (cut_fill && ts.nseg == 1)
      for (uint32_t k = threadIdx.x; k < CUT_CAP; k += INC_THREADS) cut_fill[(size_t)ts.mesh[0] * CUT_CAP + k] = -1;
  }
  // Two items per thread.  Their facet ids, then facets, then positions load
  // first; the box records (one segment: the cloth step's per-block boxes)
  // load with the facet ids and are reduced while the positions are in
  // flight, so the box costs no round of its own.
  static_assert(INC_MORTON_WINDOW == 2 * INC_THREADS, "two items per thread");
  const bool wide = code_bits > 32;
  const uint32_t start = blockIdx.x * INC_MORTON_WINDOW;
  const uint32_t cnt = min(INC_MORTON_WINDOW, n - start);
  uint32_t fid[2], sgi[2];
  f3 x0[2], x1[2], x2[2];
  SegInfo si[2];
  // items past the window's count (last window only) load the last item's
  // data, so the loads need no branches and issue together
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint32_t i = min(threadIdx.x + k * INC_THREADS, cnt - 1);
    si[k] = seg_info(ts, mt, start + i);
    fid[k] = perm[start + i];
  }
  uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  if (nparts)  // records of 24 bytes (8-byte aligned): three 8-byte loads each, 4 records in flight
#pragma unroll 4
    for (uint32_t r = threadIdx.x; r < nparts; r += INC_THREADS) {
      const uint2* q = reinterpret_cast<const uint2*>(box_parts + 6 * (size_t)r);
      const uint2 u0 = q[0], u1 = q[1], u2 = q[2];
      v[0] = min(v[0], u0.x);
      v[1] = min(v[1], u0.y);
      v[2] = min(v[2], u1.x);
      v[3] = max(v[3], u1.y);
      v[4] = max(v[4], u2.x);
      v[5] = max(v[5], u2.y);
    }
  pu3 f[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) f[k] = *reinterpret_cast<const pu3*>(facets + 3 * (size_t)(si[k].foff + fid[k]));
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint32_t vb = si[k].voff;
    x0[k] = ld3(x, vb + f[k].a);
    x1[k] = ld3(x, vb + f[k].b);
    x2[k] = ld3(x, vb + f[k].c);
    sgi[k] = si[k].sg;
  }
  if (nparts) {
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = k < 3 ? wave_min_u32(v[k]) : wave_max_u32(v[k]);
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    constexpr uint32_t NW = INC_THREADS / 64;
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 6; ++k) s_part[k * NW + wv] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
      const uint32_t k = threadIdx.x;
      uint32_t o = s_part[k * NW];
      for (uint32_t ww = 1; ww < NW; ++ww) o = k < 3 ? min(o, s_part[k * NW + ww]) : max(o, s_part[k * NW + ww]);
      s_box[threadIdx.x] = o;
      if (blockIdx.x == 0) acc[threadIdx.x] = o;
    }
    __syncthreads();
  }
  const uint32_t* box = nparts ? s_box : acc;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint32_t i = threadIdx.x + k * INC_THREADS;
    if (i >= cnt) {
      sk[i] = (K)~(K)0;
      sv[i] = ~0u;
      continue;
    }
    const uint32_t sg = sgi[k];
    const uint32_t* a = box + sg * 6;
    const f3 bmin = {funkey(a[0]), funkey(a[1]), funkey(a[2])};
    const f3 bmax = {funkey(a[3]), funkey(a[4]), funkey(a[5])};
    const uint64_t code = morton_of(x0[k], x1[k], x2[k], bmin, bmax, wide);
    sk[i] = (ts.nseg > 1) ? (K)(((uint64_t)sg << code_bits) | code) : (K)code;
    sv[i] = fid[k];
  }
  __syncthreads();
  window_sort<K, INC_MORTON_WINDOW, INC_THREADS>(sk, sv, cnt);
  for (uint32_t i = threadIdx.x; i < cnt; i += INC_THREADS) {
    keys[start + i] = sk[i];
    vals[start + i] = sv[i];
  }
}

int launch_morton_inc(hipStream_t s, SegTable tsegs, MeshTables mt, const float* x, const uint32_t* facets,
                      uint32_t* box_acc, uint32_t code_bits, bool key64, const uint32_t* perm, void* keys,
                      uint32_t* vals, uint32_t nitems, uint32_t* need_full, const uint32_t* box_parts,
                      uint32_t nparts, uint32_t* cut_reset, int32_t* cut_fill) {
  if (!nitems) return YSIM_OK;
  if (nparts && (!box_parts || tsegs.nseg != 1)) return YSIM_ERR_INVALID;
  const uint32_t blocks = ceil_div(nitems, INC_MORTON_WINDOW);
  if (key64)
    hipLaunchKernelGGL(k_morton_inc<uint64_t>, dim3(blocks), dim3(INC_THREADS), 0, s, tsegs, mt, x, facets, box_acc,
                       code_bits, perm, (uint64_t*)keys, vals, nitems, need_full, box_parts, nparts, cut_reset,
                       cut_fill);
  else
    hipLaunchKernelGGL(k_morton_inc<uint32_t>, dim3(blocks), dim3(INC_THREADS), 0, s, tsegs, mt, x, facets, box_acc,
                       code_bits, perm, (uint32_t*)keys, vals, nitems, need_full, box_parts, nparts, cut_reset,
                       cut_fill);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// ------------------------------------------------------------------ S3 Karras
namespace {
template <typename K>
struct KarrasSeg {
  const K* keys;
  const uint32_t* idx;
  int n;
  uint64_t mask;
  bool wide;
  __device__ __forceinline__ uint64_t code(int i) const { return (uint64_t)keys[i] & mask; }
  __device__ __forceinline__ int clzw(uint64_t v) const {
    return wide ? __clzll((long long)v) : __clz((int)(uint32_t)v);
  }
  // delta, bvh.metal:226-238
  __device__ __forceinline__ int delta(int i, int j) const {
    if (j < 0 || j >= n) return -1;
    uint64_t ci = code(i), cj = code(j);
    if (ci != cj) return clzw(ci ^ cj);
    return (wide ? 64 : 32) + __clz((int)(idx[i] ^ idx[j]));
  }
  // delta from codes already loaded (ci = code(i), cj = code(j) when j is in range)
  __device__ __forceinline__ int delta_c(int i, int j, uint64_t ci, uint64_t cj) const {
    if (j < 0 || j >= n) return -1;
    if (ci != cj) return clzw(ci ^ cj);
    return (wide ? 64 : 32) + __clz((int)(idx[i] ^ idx[j]));
  }
  __device__ __forceinline__ uint64_t code_at(int j) const { return code(j < 0 ? 0 : (j >= n ? n - 1 : j)); }
};
}  // namespace

template <typename K>
__global__ __launch_bounds__(256) void k_hierarchy(SegTable ts, MeshTables mt, const K* __restrict__ keys,
                                                   const uint32_t* __restrict__ vals, uint32_t code_bits,
                                                   ysim_bvh_node* __restrict__ tree,
                                                   int32_t* __restrict__ parent,
                                                   uint32_t* __restrict__ visit, uint32_t n,
                                                   uint32_t* __restrict__ box_acc, uint32_t nbox,
                                                   uint32_t* __restrict__ hist, uint32_t hist_words,
                                                   uint32_t* __restrict__ perm_out, RefitLinks rl,
                                                   HierarchySkip skip) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  // reset accumulators for the next build (their readers ran in earlier kernels)
  if (j < nbox * 6) box_acc[j] = ((j % 6) < 3) ? 0xffffffffu : 0u;
  if (j < hist_words) hist[j] = 0;
  // lanes past the items stay for the wave's cooperative searches below
  const bool live = j < n;
  if (__ballot(live) == 0) {
    if (j == 0 && skip.flag && *skip.flag && skip.note)
      __hip_atomic_store(skip.note, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  const SegInfo si = seg_info(ts, mt, live ? j : n - 1);
  const uint32_t m = si.m;
  const uint32_t s = si.s;
  const int nn = (int)si.n;
  const int i = (int)((live ? j : n - 1) - s);
  const uint32_t base = si.noff;
  ysim_bvh_node* T = tree + base;
  int32_t* P = parent + base;
  const bool wide = code_bits > 32;
  const uint64_t mask = (code_bits >= 64) ? ~0ull : ((1ull << code_bits) - 1);
  KarrasSeg<K> K_{keys + s, vals + s, nn, mask, wide};
  // Internal node i (i < nn - 1; the other lanes only help the wave's
  // cooperative searches below).
  const bool node = live && i < nn - 1;
  // The item's facet, the node's hint and its neighbours' keys load before
  // the deferral flag is read (every address is valid either way), so a
  // replayed build pays one round trip for all of them; nothing is stored
  // before the flag is known.
  const uint32_t idx_i = K_.idx[i];
  const int2 h = (rl.hint && node) ? rl.hint[base + i] : make_int2(i, i);
  const uint64_t ci = K_.code(i), cl = K_.code_at(i - 1), cr = K_.code_at(i + 1);
  if (skip.flag && *skip.flag) {
    if (j == 0 && skip.note) __hip_atomic_store(skip.note, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  if (live) {
    // the sorted facet order seeds the next incremental build
    if (perm_out) perm_out[j] = idx_i;
    // leaf slot nn-1+i: child_a = -1, child_b = facet id (the LDS refit writes
    // the node when rl.fid is set)
    if (rl.fid) {
      rl.fid[base + nn - 1 + i] = idx_i;
    } else {
      T[nn - 1 + i].child_a = -1;
      T[nn - 1 + i].child_b = (int32_t)idx_i;
    }
    if (i == 0) P[0] = -1;
  }
  if (node) visit[base + i] = 0;
  // determineRange, bvh.metal:240-258.  delta(i, i + k d) does not increase
  // with k (sorted (code, index) keys), so the range end is the last k whose
  // delta exceeds dmin.  The doubling search stops at k = 1024 (HSEQ); nodes
  // of longer ranges (about one in 500) are searched by the whole wave, 64
  // probes per round, which finds the same boundary in ~log64 dependent
  // rounds instead of ~2 log2.  On the 512 x 512 fold: 29.2 -> 26.2 us; a
  // threshold of 64 sends ~16 nodes of each wave through the cooperative
  // rounds one after another (28.9 us) and fetches 3.6x the keys.
  //
  // With rl.hint, the node's range end and split of the last build are
  // checked first; the searches run only where they fail.  Both answers are
  // defined by monotone boundaries, so a hint (jj, s) is the answer exactly
  // when delta(i, jj) > dmin >= delta(i, jj + d) and, with lo, hi the ends
  // and common = delta(lo, hi) = delta(i, jj), s = lo or delta(lo, s) >
  // common, and s + 1 = hi or delta(lo, s + 1) <= common.  Every probe
  // address comes from the hint, so the four probe keys load together with
  // the node's neighbours.  Between substeps of a moving cloth the sorted
  // order barely changes: on the 512 x 512 fold ~99.99% of the nodes keep
  // their range and split, and only nodes of a few leaves change.  A wrong
  // or stale hint costs the search, never a different tree.
  constexpr int HSEQ = 1024;
  int d = 1, dmin = -1, l = 0;
  bool coop = false, hinted = false;
  int lo = 0, hi = 0;
  uint32_t split = 0;
  if (node) {
    const int hd = h.x > i ? 1 : -1;
    const uint64_t cj = K_.code_at(h.x), cjn = K_.code_at(h.x + hd), cs = K_.code_at(h.y), cs1 = K_.code_at(h.y + 1);
    const int dr = K_.delta_c(i, i + 1, ci, cr), dl = K_.delta_c(i, i - 1, ci, cl);
    d = (dr - dl >= 0) ? 1 : -1;
    dmin = d > 0 ? dl : dr;  // delta(i, i - d)
    if (rl.hint && h.x != i && hd == d && h.x >= 0 && h.x < nn) {
      const int common = K_.delta_c(i, h.x, ci, cj);
      const int hlo = min(i, h.x), hhi = max(i, h.x);
      const uint64_t clo = d > 0 ? ci : cj;
      if (common > dmin && K_.delta_c(i, h.x + d, ci, cjn) <= dmin && h.y >= hlo && h.y < hhi &&
          (h.y == hlo || K_.delta_c(hlo, h.y, clo, cs) > common) &&
          (h.y + 1 == hhi || K_.delta_c(hlo, h.y + 1, clo, cs1) <= common)) {
        hinted = true;
        lo = hlo;
        hi = hhi;
        split = (uint32_t)h.y;
      }
    }
    if (!hinted) {
      int lmax = 2;
      while (lmax <= HSEQ && K_.delta(i, i + lmax * d) > dmin) lmax <<= 1;
      coop = lmax > HSEQ;
      if (!coop)
        for (int t = lmax >> 1; t >= 1; t >>= 1)
          if (K_.delta(i, i + (l + t) * d) > dmin) l += t;
    }
  }
  // findSplit, bvh.metal:197-224, measuring the common prefix with delta()
  // (code, then index tie-break) in every case; the reference's midpoint
  // shortcut for equal end codes contradicts that tie-break for runs of >= 3
  // equal codes.  Identical split whenever the end codes differ.  The split
  // is the last position x in [lo, hi) with delta(lo, x) > delta(lo, hi)
  // (lo when none): again a monotone boundary.
  if (node && !coop && !hinted) {
    const int jj = i + l * d;
    lo = d < 0 ? jj : i;
    hi = d < 0 ? i : jj;
    const int common = K_.delta(lo, hi);
    uint32_t sp = (uint32_t)lo, step = (uint32_t)(hi - lo);
    do {
      step = (step + 1) >> 1;
      uint32_t ns = sp + step;
      if (ns < (uint32_t)hi && K_.delta(lo, (int)ns) > common) sp = ns;
    } while (step > 1);
    split = sp;
  }
  // Cooperative searches, one long-range node of the wave at a time (lanes
  // of other segments help with the node's own segment table).  Each
  // round probes 64 evenly spaced positions of the open interval; the
  // ballot of a monotone predicate is a prefix, whose length narrows it.
  const uint32_t lane = __lane_id();
  auto boundary = [&](int a, int b, auto&& pred) {  // last x in [a, b] with pred(x); pred(a) holds
    while (b > a) {
      const int stride = (b - a + 63) / 64;
      const int x = a + ((int)lane + 1) * stride;
      const uint64_t hit = __ballot(x <= b && pred(x));
      const int c = (int)__popcll(hit);
      b = min(b, a + (c + 1) * stride - 1);
      a += c * stride;
    }
    return a;
  };
  uint64_t todo = __ballot(coop);
  while (todo) {
    const int src = (int)__builtin_ctzll(todo);
    todo &= todo - 1;
    const int ni = __shfl(i, src, 64), nd = __shfl(d, src, 64), ndm = __shfl(dmin, src, 64);
    const uint32_t ns0 = __shfl(s, src, 64);  // the node's segment (lanes may lie in others)
    const int nnn = __shfl(nn, src, 64);
    const KarrasSeg<K> KN{keys + ns0, vals + ns0, nnn, mask, wide};
    const int kmax = nd > 0 ? nnn - 1 - ni : ni;  // delta is -1 past the segment
    const int nl = boundary(HSEQ, kmax, [&](int k) { return KN.delta(ni, ni + k * nd) > ndm; });
    const int jj = ni + nl * nd;
    const int nlo = nd < 0 ? jj : ni, nhi = nd < 0 ? ni : jj;
    const int common = KN.delta(nlo, nhi);
    const int nsp = boundary(nlo, nhi - 1, [&](int x) { return KN.delta(nlo, x) > common; });
    if ((int)lane == src) {
      lo = nlo;
      hi = nhi;
      split = (uint32_t)nsp;
    }
  }
  if (rl.hint && node && !hinted) rl.hint[base + i] = make_int2(d > 0 ? hi : lo, (int)split);
  int ca = ((int)split == lo) ? (int)split + nn - 1 : (int)split;
  int cb = ((int)split + 1 == hi) ? (int)split + nn : (int)split + 1;
  if (rl.span_rb) {  // nodes whose leaf range crosses a refit block boundary
    const uint32_t rb = rl.span_rb;
    const bool sp = node && (s + (uint32_t)lo) / rb != (s + (uint32_t)hi) / rb;
    const uint64_t bm = __ballot(sp);
    if (lane == 0) rl.span_mask[j >> 6] = bm;  // lane 0 is live: lanes are ascending positions
    if (sp) rl.span_rec[j] = make_int4(lo, hi, ca, cb);
  }
  if (!node) return;
  if (!rl.fid) {
    T[i].child_a = ca;
    T[i].child_b = cb;
  }
  // a node whose range and split passed the check has the children it had
  // at the last build, whose parent and `up` links are already these values
  if (!hinted) {
    P[ca] = i;
    P[cb] = i;
    if (rl.up) {  // this node is the high end of its range when d < 0
      const uint32_t size4 = ((uint32_t)(hi - lo + 1) << 2) | (d < 0 ? 2u : 0u);
      rl.up[base + ca] = size4;       // left child
      rl.up[base + cb] = size4 | 1u;  // right child
    }
  }
  if (rl.cut) {  // children of at most S leaves under a node of more
    const uint32_t S = max(1u, ((uint32_t)nn + 127u) / 128u);
    if ((uint32_t)(hi - lo + 1) > S) {
      const uint32_t sa = split - (uint32_t)lo + 1, sb = (uint32_t)hi - split;
      int32_t* CL = rl.cut + (size_t)m * CUT_CAP;
      if (sa <= S) {
        const uint32_t k = atomicAdd(rl.cut_cnt + m, 1u);
        if (k < CUT_CAP) CL[k] = ca;
      }
      if (sb <= S) {
        const uint32_t k = atomicAdd(rl.cut_cnt + m, 1u);
        if (k < CUT_CAP) CL[k] = cb;
      }
    }
  }
}

int launch_hierarchy(hipStream_t s, SegTable tsegs, MeshTables mt, const void* keys, bool key64,
                     const uint32_t* vals, uint32_t code_bits, ysim_bvh_node* tree, int32_t* parent,
                     uint32_t* visit, uint32_t nitems, uint32_t* box_acc, uint32_t nbox, uint32_t* hist,
                     uint32_t hist_words, uint32_t* perm_out, RefitLinks rl, HierarchySkip skip) {
  uint32_t work = std::max<uint32_t>(nitems, std::max<uint32_t>(nbox * 6, hist_words));
  if (!work) return YSIM_OK;
  uint32_t blocks = ceil_div(work, 256);
  if (key64)
    hipLaunchKernelGGL(k_hierarchy<uint64_t>, dim3(blocks), dim3(256), 0, s, tsegs, mt,
                       (const uint64_t*)keys, vals, code_bits, tree, parent, visit, nitems, box_acc, nbox,
                       hist, hist_words, perm_out, rl, skip);
  else
    hipLaunchKernelGGL(k_hierarchy<uint32_t>, dim3(blocks), dim3(256), 0, s, tsegs, mt,
                       (const uint32_t*)keys, vals, code_bits, tree, parent, visit, nitems, box_acc, nbox,
                       hist, hist_words, perm_out, rl, skip);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// ------------------------------------------------------------------ S3' agglomerative build
// Apetrei's single-pass hierarchy (agglomerativeBuild_Tri + agglomerativeSwapRoot,
// bvh.metal:884-1011, 1135-1199), fused here with the leaf boxes of the
// substep's refit (swept when requested), so a rebuild needs no separate refit.
// One thread per sorted leaf climbs: its node [L, R] merges toward the side
// whose neighbouring key pair differs less (code XOR, facet-index XOR on
// equal codes); the parent is internal node R (merging right: this node is
// its child_a and records L) or L - 1 (merging left: child_b, records R).  The
// second arrival at a parent (visit counters advance by two per pass, as in
// the refit) reads the sibling's link, box and range end and continues with
// the union in (child_a, child_b) order.  The thread that completes [0, n-1]
// holds the root at slot k and swaps it into slot 0, retargeting every link
// that named either slot; nothing else of its segment is still running.
// Links, ranges and boxes that another workgroup reads are published as in
// the refit (device-coherent stores, then the arrival RMW; FORMAL: acq_rel).
// Also does the hierarchy kernel's duties: the sorted facet order for the
// next incremental build and the box / histogram resets.
template <typename K, bool FORMAL>
__global__ __launch_bounds__(256) void k_agglo(SegTable ts, MeshTables mt, const K* __restrict__ keys,
                                               const uint32_t* __restrict__ vals, uint32_t code_bits,
                                               const float* __restrict__ x, const float* __restrict__ v, float dt,
                                               int swept, const uint32_t* __restrict__ facets, ysim_bvh_node* tree,
                                               int32_t* parent, uint32_t* visit, int2* range, uint32_t n,
                                               uint32_t* __restrict__ box_acc, uint32_t nbox,
                                               uint32_t* __restrict__ hist, uint32_t hist_words,
                                               uint32_t* __restrict__ perm_out, uint32_t* err) {
  const uint32_t j = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (j < nbox * 6) box_acc[j] = ((j % 6) < 3) ? 0xffffffffu : 0u;
  if (j < hist_words) hist[j] = 0;
  if (j >= n) return;
  const uint32_t sg = seg_of(ts, j);
  const uint32_t m = ts.mesh[sg];
  const uint32_t s = ts.off[sg];
  const int nn = (int)(ts.off[sg + 1] - s);
  const int i = (int)(j - s);
  const uint32_t base = mt.noff[m];
  ysim_bvh_node* T = tree + base;
  int32_t* P = parent + base;
  int2* RG = range + base;
  uint32_t* C = visit + base;
  const K* key = keys + s;
  const uint32_t* idx = vals + s;
  const uint64_t mask = (code_bits >= 64) ? ~0ull : ((1ull << code_bits) - 1);
  const uint32_t fid = idx[i];
  if (perm_out) perm_out[j] = fid;

  const uint32_t vb = mt.voff[m];
  f3 lo, hi;
  tri_box(x, v, dt, swept, facets + 3 * (size_t)(mt.foff[m] + fid), vb, lo, hi);
  auto store_box = [&](int node, f3 a, f3 b) {
    float* q = reinterpret_cast<float*>(T + node);
    if (FORMAL) {
      q[0] = a.x; q[1] = a.y; q[2] = a.z; q[4] = b.x; q[5] = b.y; q[6] = b.z;
    } else {
      st_agent_u64(reinterpret_cast<uint64_t*>(q + 0), pack2(a.x, a.y));
      st_agent(q + 2, a.z);
      st_agent_u64(reinterpret_cast<uint64_t*>(q + 4), pack2(b.x, b.y));
      st_agent(q + 6, b.z);
    }
  };
  auto st_i = [&](int32_t* q, int32_t val) {
    if (FORMAL) *q = val;
    else st_agent_u32(reinterpret_cast<uint32_t*>(q), (uint32_t)val);
  };
  auto ld_i = [&](const int32_t* q) -> int32_t {
    if (FORMAL) return *q;
    return (int32_t)ld_agent_u32(reinterpret_cast<const uint32_t*>(q));
  };
  int cur = nn - 1 + i;
  T[cur].child_a = -1;
  T[cur].child_b = (int32_t)fid;
  store_box(cur, lo, hi);
  if (nn <= 1) {
    P[0] = -1;
    return;
  }
  int L = i, R = i;
  for (int step = 0; step < 2 * nn; ++step) {
    if (L == 0 && R == nn - 1) {
      // agglomerativeSwapRoot: every other node of this segment is final
      const int k = cur;
      if (k == 0) {
        float* q0 = reinterpret_cast<float*>(T);
        q0[0] = lo.x; q0[1] = lo.y; q0[2] = lo.z; q0[4] = hi.x; q0[5] = hi.y; q0[6] = hi.z;
      } else {
        const int rca = ld_i(&T[k].child_a), rcb = ld_i(&T[k].child_b);
        const int oa = ld_i(&T[0].child_a), ob = ld_i(&T[0].child_b);
        const int p0 = ld_i(&P[0]);
        float ob_box[6];
        {
          const float* q = reinterpret_cast<const float*>(T);
          if (FORMAL) {
            ob_box[0] = q[0]; ob_box[1] = q[1]; ob_box[2] = q[2]; ob_box[3] = q[4]; ob_box[4] = q[5]; ob_box[5] = q[6];
          } else {
            const uint64_t a01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 0));
            const uint64_t b01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 4));
            ob_box[0] = lo_f(a01); ob_box[1] = hi_f(a01); ob_box[2] = ld_agent(q + 2);
            ob_box[3] = lo_f(b01); ob_box[4] = hi_f(b01); ob_box[5] = ld_agent(q + 6);
          }
        }
        const int ra = (rca == 0) ? k : rca, rb = (rcb == 0) ? k : rcb;
        const int nodes = 2 * nn - 1;
        if (oa < 0 || oa >= nodes || ob < 0 || ob >= nodes || ra < 0 || ra >= nodes || rb < 0 || rb >= nodes ||
            p0 < 0 || p0 >= nn - 1) {
          atomicOr(err, 4u);
          return;
        }
        float* q0 = reinterpret_cast<float*>(T);
        q0[0] = lo.x; q0[1] = lo.y; q0[2] = lo.z; q0[4] = hi.x; q0[5] = hi.y; q0[6] = hi.z;
        T[0].child_a = ra;
        T[0].child_b = rb;
        float* qk = reinterpret_cast<float*>(T + k);
        qk[0] = ob_box[0]; qk[1] = ob_box[1]; qk[2] = ob_box[2]; qk[4] = ob_box[3]; qk[5] = ob_box[4]; qk[6] = ob_box[5];
        T[k].child_a = oa;
        T[k].child_b = ob;
        P[oa] = k;
        P[ob] = k;
        P[ra] = 0;
        P[rb] = 0;
        if (p0 != k) {
          if (ld_i(&T[p0].child_a) == 0) T[p0].child_a = k;
          else T[p0].child_b = k;
          P[k] = p0;
        }
      }
      P[0] = -1;
      return;
    }
    bool right;
    if (L == 0) {
      right = true;
    } else if (R == nn - 1) {
      right = false;
    } else {
      const uint64_t dr = ((uint64_t)key[R] & mask) ^ ((uint64_t)key[R + 1] & mask);
      const uint64_t dl = ((uint64_t)key[L - 1] & mask) ^ ((uint64_t)key[L] & mask);
      right = (dr != dl) ? (dr < dl) : ((idx[R] ^ idx[R + 1]) < (idx[L - 1] ^ idx[L]));
    }
    const int p = right ? R : L - 1;
    if (p < 0 || p >= nn - 1) {
      atomicOr(err, 4u);
      return;
    }
    if (right) {
      st_i(&T[p].child_a, cur);
      st_i(&RG[p].x, L);
    } else {
      st_i(&T[p].child_b, cur);
      st_i(&RG[p].y, R);
    }
    st_i(&P[cur], p);
    uint32_t old;
    if (FORMAL) {
      old = __hip_atomic_fetch_add(C + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      wait_all_mem();
      old = __hip_atomic_fetch_add(C + p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if ((old & 1u) == 0u) return;  // first arrival
    const int sib = right ? ld_i(&T[p].child_b) : ld_i(&T[p].child_a);
    if (sib < 0 || sib >= 2 * nn - 1) {
      atomicOr(err, 4u);
      return;
    }
    if (right) R = ld_i(&RG[p].y);
    else L = ld_i(&RG[p].x);
    const float* q = reinterpret_cast<const float*>(T + sib);
    f3 smin, smax;
    if (FORMAL) {
      smin = f3{q[0], q[1], q[2]};
      smax = f3{q[4], q[5], q[6]};
    } else {
      const uint64_t m01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 0));
      const uint64_t x01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 4));
      smin = f3{lo_f(m01), hi_f(m01), ld_agent(q + 2)};
      smax = f3{lo_f(x01), hi_f(x01), ld_agent(q + 6)};
    }
    if (right) {  // this node is child_a
      lo = vmin(lo, smin);
      hi = vmax(hi, smax);
    } else {
      lo = vmin(smin, lo);
      hi = vmax(smax, hi);
    }
    if (L < 0 || R >= nn || L > R) {
      atomicOr(err, 4u);
      return;
    }
    cur = p;
    if (!(L == 0 && R == nn - 1)) store_box(cur, lo, hi);
  }
  atomicOr(err, 4u);
}

int launch_agglo(hipStream_t s, SegTable tsegs, MeshTables mt, const void* keys, bool key64, const uint32_t* vals,
                 uint32_t code_bits, const float* x, const float* v, float dt, bool swept, const uint32_t* facets,
                 ysim_bvh_node* tree, int32_t* parent, uint32_t* visit, int2* range, uint32_t nitems,
                 uint32_t* box_acc, uint32_t nbox, uint32_t* hist, uint32_t hist_words, uint32_t* perm_out,
                 uint32_t* err, bool formal) {
  uint32_t work = std::max<uint32_t>(nitems, std::max<uint32_t>(nbox * 6, hist_words));
  if (!work) return YSIM_OK;
  const uint32_t blocks = ceil_div(work, 256);
#define YSIM_A(KT, F)                                                                                          \
  hipLaunchKernelGGL((k_agglo<KT, F>), dim3(blocks), dim3(256), 0, s, tsegs, mt, (const KT*)keys, vals,         \
                     code_bits, x, v ? v : x, dt, (int)swept, facets, tree, parent, visit, range, nitems, box_acc, \
                     nbox, hist, hist_words, perm_out, err)
  if (key64) {
    if (formal) YSIM_A(uint64_t, true); else YSIM_A(uint64_t, false);
  } else {
    if (formal) YSIM_A(uint32_t, true); else YSIM_A(uint32_t, false);
  }
#undef YSIM_A
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// ------------------------------------------------------------------ S4+S5 refit
// One thread per leaf.  The leaf box is the hull of the triangle's vertices
// (and of x + v*dt when swept).  Each thread then climbs toward the root; at a
// parent the first arrival stops and the second (odd counter value: counters
// advance by two per refit, so they need no clearing between refits) unions
// its own box with the sibling's and continues.  Publication of a box to the
// sibling thread on another XCD:
//   formal: acq_rel device-scope RMW (L2 writeback / invalidate), plain boxes;
//   fast:   boxes written and read with device-coherent (sc1) accesses and the
//           writer waits for its stores to complete before the RMW.
template <bool FORMAL>
__global__ __launch_bounds__(256) void k_refit(SegTable ts, MeshTables mt, const float* __restrict__ x,
                                               const float* __restrict__ v, float dt, int swept,
                                               const uint32_t* __restrict__ facets,
                                               ysim_bvh_node* tree, const int32_t* __restrict__ parent,
                                               uint32_t* visit, uint32_t n, uint32_t* err) {
  const uint32_t j = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (j >= n) return;
  const uint32_t sg = seg_of(ts, j);
  const uint32_t m = ts.mesh[sg];
  const uint32_t s = ts.off[sg];
  const int nn = (int)(ts.off[sg + 1] - s);
  const int i = (int)(j - s);
  const uint32_t base = mt.noff[m];
  ysim_bvh_node* T = tree + base;
  const int32_t* P = parent + base;
  uint32_t* C = visit + base;

  int child = nn - 1 + i;
  const uint32_t fid = (uint32_t)T[child].child_b;
  const uint32_t vb = mt.voff[m];
  f3 lo, hi;
  tri_box(x, v, dt, swept, facets + 3 * (size_t)(mt.foff[m] + fid), vb, lo, hi);
  auto store_box = [&](int node, f3 a, f3 b) {
    float* q = reinterpret_cast<float*>(T + node);
    if (FORMAL) {
      q[0] = a.x; q[1] = a.y; q[2] = a.z; q[4] = b.x; q[5] = b.y; q[6] = b.z;
    } else {
      // nodes are 32-byte aligned: (min.x, min.y) and (max.x, max.y) are 8-byte words
      st_agent_u64(reinterpret_cast<uint64_t*>(q + 0), pack2(a.x, a.y));
      st_agent(q + 2, a.z);
      st_agent_u64(reinterpret_cast<uint64_t*>(q + 4), pack2(b.x, b.y));
      st_agent(q + 6, b.z);
    }
  };
  store_box(child, lo, hi);
  if (nn <= 1) return;
  const int nnodes = 2 * nn - 1;
  for (int step = 0; step < 2 * nn; ++step) {
    int p = P[child];
    if (p < 0 || p >= nn - 1) {  // index guard (corrupted topology)
      atomicOr(err, 4u);
      return;
    }
    uint32_t old;
    if (FORMAL) {
      old = __hip_atomic_fetch_add(C + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      wait_all_mem();
      old = __hip_atomic_fetch_add(C + p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if ((old & 1u) == 0u) return;  // first arrival: the sibling's thread combines
    const int ca = T[p].child_a, cb = T[p].child_b;
    if (ca < 0 || ca >= nnodes || cb < 0 || cb >= nnodes) {
      atomicOr(err, 4u);
      return;
    }
    const int sib = (ca == child) ? cb : ca;
    const float* q = reinterpret_cast<const float*>(T + sib);
    f3 smin, smax;
    if (FORMAL) {
      smin = f3{q[0], q[1], q[2]};
      smax = f3{q[4], q[5], q[6]};
    } else {
      const uint64_t m01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 0));
      const uint64_t x01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 4));
      smin = f3{lo_f(m01), hi_f(m01), ld_agent(q + 2)};
      smax = f3{lo_f(x01), hi_f(x01), ld_agent(q + 6)};
    }
    // union in (child_a, child_b) order, bvh.metal:567-573
    if (sib == cb) {
      lo = vmin(lo, smin);
      hi = vmax(hi, smax);
    } else {
      lo = vmin(smin, lo);
      hi = vmax(smax, hi);
    }
    store_box(p, lo, hi);
    child = p;
    if (p == 0) return;
  }
  atomicOr(err, 4u);
}

// Same refit with the lower levels resolved inside the workgroup.  A block
// owns RB (256, 512 or 1024; chosen by the autotuner) consecutive sorted leaves.
// Each thread tracks the leaf range of its current node; the parent's range
// follows from the parent's leaf count and the side, and the parent's index
// from which end of that range it is, all in `up` (written by the hierarchy
// kernel), and so does the sibling: a Karras node's children are the nodes at
// its split and split + 1.  So the parent array is not read.  The block first
// copies the `up` words of the internal nodes indexed inside it to LDS.  An
// internal node whose range lies inside the block is then combined through LDS
// with no device memory access (the children's boxes and an arrival counter
// per node; a Karras internal node is an endpoint of its own range, so its
// index falls inside the block and addresses the LDS slot).  Unions keep
// (child_a, child_b) operand order; the boxes are identical to k_refit's.
// Every node is written whole, box and child links: a leaf's links are
// (-1, facet id from fids), an internal node's are its two children, known to
// the thread that combines it (its own node and the sibling from the ranges).
//
// The nodes spanning blocks (their parent is an "escape" node's parent) are
// finished in one of two ways:
//  TOP = false: the device-wide second-arrival protocol of k_refit per level,
//    with the next level's links loaded while the arrival resolves (chains of
//    up to ~23 dependent device-scope round trips near block boundaries);
//  TOP = true (spanning nodes recorded by the hierarchy kernel for RB): the
//    block's escape nodes partition the leaves of its meshes that span blocks.
//    A spanning node's range is [lo, hi] with lo in block a and hi in block b
//    (a < b): its leftmost descendant escape (down the child_a side) starts at
//    lo in block a, its rightmost (child_b side) ends at hi in block b, and
//    the blocks in between lie inside the range.  So its box is the union of
//    three parts: the leaves [lo, end of a] (the escapes of block a from lo
//    on), whole blocks a+1 .. b-1, and the leaves [start of b, hi].  Each
//    child_a escape writes the first kind at part[lo], each child_b escape
//    the last kind at part[hi] (escapes are disjoint, so no position gets
//    both), the block writes its box to blk, and k_refit_span combines them
//    after this kernel: no dependent chain across workgroups.  A float min or
//    max is exact, so unions in another order give the same bits for finite
//    boxes without signed zeros of both signs on an axis (positions built by
//    additions never become -0).
template <bool FORMAL, int RB, bool TOP>
__global__ __launch_bounds__(RB) void k_refit_lds(SegTable ts, MeshTables mt, const float* __restrict__ x,
                                                  const float* __restrict__ v, float dt, int swept,
                                                  const uint32_t* __restrict__ facets, ysim_bvh_node* tree,
                                                  const uint32_t* __restrict__ up_links,
                                                  const uint32_t* __restrict__ fids, uint32_t* visit,
                                                  uint32_t n, uint32_t* err, RefitSpan sp) {
  __shared__ uint32_t s_cnt[RB];
  __shared__ float s_box[RB][2][6];
  __shared__ uint32_t s_up[RB];  // `up` of the internal node at the slot (of the thread's mesh)
  __shared__ uint32_t s_nesc;
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t b0 = blk * RB;
  s_cnt[threadIdx.x] = 0;
  if (TOP && threadIdx.x == 0) s_nesc = 0;
  const uint32_t j = b0 + threadIdx.x;
  const bool valid = j < n;
  const SegInfo si = seg_info(ts, mt, valid ? j : n - 1);
  const uint32_t s = si.s;
  const int nn = (int)si.n;
  const int i = (int)((valid ? j : n - 1) - s);
  const uint32_t base = si.noff;
  ysim_bvh_node* T = tree + base;
  const uint32_t* UP = up_links + base;
  uint32_t* C = visit + base;
  // the block's internal links load while the leaf's box is gathered; they
  // reach LDS after it
  const bool has_link = valid && i < nn - 1;
  uint32_t link = 0;
  if (has_link) link = UP[i];
  int child = nn - 1 + i;
  uint32_t fid = 0, upnext = 0;
  f3 lo = {0.f, 0.f, 0.f}, hi = lo;
  if (valid) {
    fid = fids[base + child];
    upnext = UP[child];
    tri_box(x, v, dt, swept, facets + 3 * (size_t)(si.foff + fid), si.voff, lo, hi);
  }
  if (has_link) s_up[threadIdx.x] = link;
  __syncthreads();
  if (!TOP && !valid) return;
  int la = -1, lb = (int)fid;  // child links of the node `child`
  // A node whose box another workgroup reads during this kernel (its parent is
  // combined through the device-wide protocol) is written with device-coherent
  // stores in the fast variant; every other node with two plain 16-byte
  // stores (the traversal reads them after the kernel ends).
  auto store_node = [&](int node, f3 a, f3 b, int ca, int cb, bool coherent) {
    float* q = reinterpret_cast<float*>(T + node);
    if (FORMAL || !coherent) {
      float4* q4 = reinterpret_cast<float4*>(q);
      q4[0] = make_float4(a.x, a.y, a.z, __int_as_float(ca));
      q4[1] = make_float4(b.x, b.y, b.z, __int_as_float(cb));
    } else {
      uint64_t* w = reinterpret_cast<uint64_t*>(q);
      st_agent_u64(w + 0, pack2(a.x, a.y));
      st_agent_u64(w + 1, pack2(a.z, __int_as_float(ca)));
      st_agent_u64(w + 2, pack2(b.x, b.y));
      st_agent_u64(w + 3, pack2(b.z, __int_as_float(cb)));
    }
  };
  uint32_t clo = (uint32_t)i, chi = (uint32_t)i;  // leaf range of the current node
  uint32_t side = 0;                              // of the current node (0: child_a)
  bool dev = false;                               // `child` is an escape: its parent spans blocks
  if (valid && nn <= 1) {
    store_node(child, lo, hi, la, lb, false);
  } else if (valid) {
    // Local climb: LDS arrivals only, one exit test per outcome and plain
    // node stores (the branch-light loop keeps the per-level instruction
    // count low; with 16 waves of a block per CU the loop is issue-bound).
    for (int step = 0;; ++step) {
      const uint32_t up = upnext;
      side = up & 1u;  // 0: child_a (left part of p's range)
      const uint32_t psize = up >> 2;
      const uint32_t plo = side ? chi + 1 - psize : clo, phi = side ? chi : clo + psize - 1;
      const int p = (int)((up & 2u) ? phi : plo);
      // sibling: child_b at split + 1 for a child_a, child_a at split for a
      // child_b; a leaf when that position ends p's range
      const uint32_t q = side ? clo - 1 : chi + 1, qe = side ? plo : phi;
      const int sib = (int)q + (q == qe ? nn - 1 : 0);
      const uint32_t slot = s + (uint32_t)p - b0;  // in [0, RB) for a local p: an endpoint of its range
      // (an inconsistent parent size would wrap plo or push phi past the
      // segment: a corrupted topology, flagged, never used as a range)
      const bool bad = (p < 0) | (p >= nn - 1) | (sib < 0) | (sib >= 2 * nn - 1) | (step >= 2 * nn) |
                       (plo > phi) | (phi >= (uint32_t)nn);
      const bool local = !bad & (s + plo >= b0) & (s + phi < b0 + RB);
      if (!local || slot >= (uint32_t)RB) {
        dev = !local && !bad;  // else a corrupted topology
        if (!dev) atomicOr(err, 4u);
        break;
      }
      store_node(child, lo, hi, la, lb, false);
      float* mine = s_box[slot][side];
      mine[0] = lo.x; mine[1] = lo.y; mine[2] = lo.z; mine[3] = hi.x; mine[4] = hi.y; mine[5] = hi.z;
      const uint32_t old = __hip_atomic_fetch_add(&s_cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0u) break;  // first arrival: the sibling's thread combines
      const float* o = s_box[slot][side ^ 1u];
      const f3 omin = {o[0], o[1], o[2]}, omax = {o[3], o[4], o[5]};
      // union in (child_a, child_b) order, bvh.metal:567-573
      lo = side ? vmin(omin, lo) : vmin(lo, omin);
      hi = side ? vmax(omax, hi) : vmax(hi, omax);
      la = side ? sib : child;
      lb = side ? child : sib;
      child = p;
      clo = plo;
      chi = phi;
      upnext = s_up[slot];
      if (p == 0) {
        store_node(0, lo, hi, la, lb, false);
        break;
      }
    }
  }
  if constexpr (TOP) {
    // Escapes: stored whole (plain stores; read after the kernel), listed in
    // LDS (s_box is free once every local node is combined), then each writes
    // its partial union and thread 0 the block's box.
    __syncthreads();
    if (dev) {
      store_node(child, lo, hi, la, lb, false);
      const uint32_t k = atomicAdd(&s_nesc, 1u);  // < RB: escapes are disjoint leaf ranges of the block
      float* r = &s_box[k][0][0];
      r[0] = lo.x; r[1] = lo.y; r[2] = lo.z; r[3] = hi.x; r[4] = hi.y; r[5] = hi.z;
      r[6] = __uint_as_float(s + clo); r[7] = __uint_as_float(s + chi);  // launch positions
    }
    __syncthreads();
    const uint32_t ne = s_nesc;
    if (dev) {
      // child_a escape: the escapes of the block right of it (its parent
      // starts at clo and ends past the block, so its mesh covers the rest of
      // the block); child_b: those left of it.
      f3 ulo = lo, uhi = hi;
      for (uint32_t e = 0; e < ne; ++e) {
        const float* r = &s_box[e][0][0];
        const uint32_t elo = __float_as_uint(r[6]), ehi = __float_as_uint(r[7]);
        if (side ? ehi < s + clo : elo > s + chi) {
          ulo = vmin(ulo, f3{r[0], r[1], r[2]});
          uhi = vmax(uhi, f3{r[3], r[4], r[5]});
        }
      }
      float4* q = sp.part + 2 * (size_t)(s + (side ? chi : clo));
      q[0] = make_float4(ulo.x, ulo.y, ulo.z, 0.f);
      q[1] = make_float4(uhi.x, uhi.y, uhi.z, 0.f);
    }
    if (threadIdx.x == 0 && ne > 0) {
      // the block's box: read only for blocks inside a spanning range, whose
      // leaves all lie in that one mesh
      const float* r = &s_box[0][0][0];
      f3 ulo = {r[0], r[1], r[2]}, uhi = {r[3], r[4], r[5]};
      for (uint32_t e = 1; e < ne; ++e) {
        r = &s_box[e][0][0];
        ulo = vmin(ulo, f3{r[0], r[1], r[2]});
        uhi = vmax(uhi, f3{r[3], r[4], r[5]});
      }
      sp.blk[2 * blk] = make_float4(ulo.x, ulo.y, ulo.z, 0.f);
      sp.blk[2 * blk + 1] = make_float4(uhi.x, uhi.y, uhi.z, 0.f);
    }
    return;
  } else {
    if (!dev) return;
    // Device-wide climb from `child`, whose parent spans blocks: the node goes
    // out with device-coherent stores (another workgroup reads it as a
    // sibling), then the second-arrival protocol of k_refit per level, with the
    // next level's links loaded while the arrival resolves.
    store_node(child, lo, hi, la, lb, true);
    for (int step = 0; step < 2 * nn; ++step) {
      const uint32_t up = upnext;
      side = up & 1u;
      const uint32_t psize = up >> 2;
      const uint32_t plo = side ? chi + 1 - psize : clo, phi = side ? chi : clo + psize - 1;
      const int p = (int)((up & 2u) ? phi : plo);
      if (plo > phi || p < 0 || p >= nn - 1) break;
      const uint32_t sq = side ? clo - 1 : chi + 1, sqe = side ? plo : phi;
      const int sib = (int)sq + (sq == sqe ? nn - 1 : 0);
      if (sib < 0 || sib >= 2 * nn - 1) break;
      upnext = UP[p];
      uint32_t old;
      if (FORMAL) {
        old = __hip_atomic_fetch_add(C + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        wait_all_mem();
        old = __hip_atomic_fetch_add(C + p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if ((old & 1u) == 0u) return;
      const float* q = reinterpret_cast<const float*>(T + sib);
      f3 smin, smax;
      if (FORMAL) {
        smin = f3{q[0], q[1], q[2]};
        smax = f3{q[4], q[5], q[6]};
      } else {
        const uint64_t m01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 0));
        const uint64_t x01 = ld_agent_u64(reinterpret_cast<const uint64_t*>(q + 4));
        smin = f3{lo_f(m01), hi_f(m01), ld_agent(q + 2)};
        smax = f3{lo_f(x01), hi_f(x01), ld_agent(q + 6)};
      }
      if (side == 0) {
        lo = vmin(lo, smin);
        hi = vmax(hi, smax);
      } else {
        lo = vmin(smin, lo);
        hi = vmax(smax, hi);
      }
      la = side ? sib : child;
      lb = side ? child : sib;
      child = p;
      clo = plo;
      chi = phi;
      // the root is read by no other workgroup; every other spanning node is
      store_node(p, lo, hi, la, lb, p != 0);
      if (p == 0) return;
    }
    atomicOr(err, 4u);
  }
}

// Second kernel of the TOP refit: one wave per word of the span mask, one
// lane per spanning node of the word.  Box = part[lo] (leaves lo .. end of
// block a) U blk[a+1 .. b-1] U part[hi] (start of block b .. hi), as ordered
// keys (exact min/max in any order).  A lane takes up to SPAN_LANE_BLOCKS
// middle blocks itself; the few nodes with more (near the root) are reduced
// by the whole wave, one after another.  The node is written whole with the
// child links the hierarchy kernel recorded.
constexpr uint32_t SPAN_LANE_BLOCKS = 4;
template <int RB>
__global__ __launch_bounds__(256) void k_refit_span(SegTable ts, MeshTables mt, ysim_bvh_node* tree, RefitSpan sp,
                                                    uint32_t nwords) {
  const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nwords) return;
  const uint64_t bits = sp.mask[w];
  if (!bits) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t j = w * 64 + lane;
  const bool mine = (bits >> lane) & 1u;
  uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  auto take = [](uint32_t* t, const float4* q) {
    const float4 l = q[0], h = q[1];
    t[0] = min(t[0], fkey(l.x)); t[1] = min(t[1], fkey(l.y)); t[2] = min(t[2], fkey(l.z));
    t[3] = max(t[3], fkey(h.x)); t[4] = max(t[4], fkey(h.y)); t[5] = max(t[5], fkey(h.z));
  };
  int4 r = make_int4(0, 0, 0, 0);
  uint32_t s = 0, base = 0, a = 0, b = 0;
  if (mine) {
    r = sp.rec[j];
    const SegInfo si = seg_info(ts, mt, j);
    s = si.s;
    base = si.noff;
    const uint32_t plo = s + (uint32_t)r.x, phi = s + (uint32_t)r.y;
    a = plo / RB;
    b = phi / RB;
    take(k, sp.part + 2 * (size_t)plo);
    take(k, sp.part + 2 * (size_t)phi);
    if (b - a - 1 <= SPAN_LANE_BLOCKS)
      for (uint32_t c = a + 1; c < b; ++c) take(k, sp.blk + 2 * (size_t)c);
  }
  uint64_t big = __ballot(mine && b - a - 1 > SPAN_LANE_BLOCKS);
  while (big) {
    const int src = __builtin_ctzll(big);
    big &= big - 1;
    const uint32_t ba = __shfl(a, src, 64), bb = __shfl(b, src, 64);
    uint32_t t[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (uint32_t c = ba + 1 + lane; c < bb; c += 64) take(t, sp.blk + 2 * (size_t)c);
#pragma unroll
    for (int e = 0; e < 6; ++e) t[e] = e < 3 ? wave_min_u32(t[e]) : wave_max_u32(t[e]);
    if ((int)lane == src) {
      k[0] = min(k[0], t[0]); k[1] = min(k[1], t[1]); k[2] = min(k[2], t[2]);
      k[3] = max(k[3], t[3]); k[4] = max(k[4], t[4]); k[5] = max(k[5], t[5]);
    }
  }
  if (mine) {
    float4* q4 = reinterpret_cast<float4*>(tree + base + (j - s));
    q4[0] = make_float4(funkey(k[0]), funkey(k[1]), funkey(k[2]), __int_as_float(r.z));
    q4[1] = make_float4(funkey(k[3]), funkey(k[4]), funkey(k[5]), __int_as_float(r.w));
  }
}

bool refit_block_ok(uint32_t b) { return b == 256 || b == 512 || b == 1024; }

int launch_refit(hipStream_t s, SegTable tsegs, MeshTables mt, const float* x, const float* v, float dt,
                 bool swept, const uint32_t* facets, ysim_bvh_node* tree, const int32_t* parent,
                 uint32_t* visit, uint32_t nitems, uint32_t* err, bool formal, const RefitLinks* rl,
                 uint32_t block, const RefitSpan* span) {
  if (!nitems) return YSIM_OK;
  if (rl && rl->up) {
    if (!refit_block_ok(block) || !rl->fid) return YSIM_ERR_INVALID;
    const bool top = span && span->mask && !formal;
    const RefitSpan sp = top ? *span : RefitSpan{};
#define YSIM_R(F, B, TP)                                                                                     \
  hipLaunchKernelGGL((k_refit_lds<F, B, TP>), dim3(ceil_div(nitems, B)), dim3(B), 0, s, tsegs, mt, x, v, dt, \
                     (int)swept, facets, tree, rl->up, rl->fid, visit, nitems, err, sp)
    if (formal) {
      if (block == 256) YSIM_R(true, 256, false); else if (block == 512) YSIM_R(true, 512, false); else YSIM_R(true, 1024, false);
    } else if (!top) {
      if (block == 256) YSIM_R(false, 256, false); else if (block == 512) YSIM_R(false, 512, false); else YSIM_R(false, 1024, false);
    } else {
      if (block == 256) YSIM_R(false, 256, true); else if (block == 512) YSIM_R(false, 512, true); else YSIM_R(false, 1024, true);
    }
#undef YSIM_R
    YSIM_CHECK_LAUNCH();
    if (top) {
      const uint32_t nwords = ceil_div(nitems, 64);
#define YSIM_S(B) \
  hipLaunchKernelGGL((k_refit_span<B>), dim3(ceil_div(nwords, 4)), dim3(256), 0, s, tsegs, mt, tree, sp, nwords)
      if (block == 256) YSIM_S(256); else if (block == 512) YSIM_S(512); else YSIM_S(1024);
#undef YSIM_S
      YSIM_CHECK_LAUNCH();
    }
    return YSIM_OK;
  }
  if (formal)
    hipLaunchKernelGGL(k_refit<true>, dim3(ceil_div(nitems, 256)), dim3(256), 0, s, tsegs, mt, x, v, dt,
                       (int)swept, facets, tree, parent, visit, nitems, err);
  else
    hipLaunchKernelGGL(k_refit<false>, dim3(ceil_div(nitems, 256)), dim3(256), 0, s, tsegs, mt, x, v, dt,
                       (int)swept, facets, tree, parent, visit, nitems, err);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

}  // namespace ysim
