// common.hpp — device helpers shared by every kernel of the gfx950 cloth
// library, plus the host-side context and error plumbing.
//
// Arithmetic contract: every translation unit is compiled with
// -ffp-contract=off and without fast-math; IEEE division and sqrt are the HIP
// defaults.  The FP32 results are therefore bit-identical to a strict-IEEE CPU
// evaluation of the same operation sequence.  Each helper states the reference
// kernel it restates.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ysim_hip.h"

// Helpers marked YSIM_HD are pure functions usable on the device and, for
// CPU checks of the kernel logic, on the host.
#define YSIM_HD __host__ __device__ __forceinline__

static_assert(sizeof(ysim_aabb4) == 32, "AABB4 layout");
static_assert(sizeof(ysim_bvh_node) == 32, "BVHNode layout");
static_assert(sizeof(ysim_broad_collision) == 32, "BroadCollision layout");
static_assert(sizeof(ysim_narrow_collision) == 48, "NarrowCollision layout");
static_assert(sizeof(ysim_fastgrid_params) == 44, "FastGridClothBehaviorParams layout");

namespace ysim {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define YSIM_HIP_TRY(expr)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::ysim::fail(YSIM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define YSIM_CHECK_LAUNCH() YSIM_HIP_TRY(hipGetLastError())

// ------------------------------------------------------------------ context
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 256;
  bool formal_refit = false;  // YSIM_CTX_FORMAL_REFIT
  // scratch reused by stage entry points (sort, counters)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  int ensure_scratch(size_t bytes);
  // device address of this device's static scratch array (the allocation-free
  // sort), or null when its lookup failed (noalloc_err says why)
  void* noalloc = nullptr;
  int noalloc_err = 0;
};

// ------------------------------------------------------------------ vectors
struct f3 {
  float x, y, z;
};
struct __attribute__((packed, aligned(4))) pf3 {
  float x, y, z;
};

__device__ __forceinline__ f3 ld3(const float* a, uint32_t i) {
  pf3 v = reinterpret_cast<const pf3*>(a)[i];
  return {v.x, v.y, v.z};
}
__device__ __forceinline__ void st3(float* a, uint32_t i, f3 v) {
  reinterpret_cast<pf3*>(a)[i] = pf3{v.x, v.y, v.z};
}
YSIM_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
YSIM_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
YSIM_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
YSIM_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
YSIM_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
YSIM_HD f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
YSIM_HD f3 vmin(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
YSIM_HD f3 vmax(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

// ------------------------------------------------------------------ ordered float keys
// Monotone float <-> uint32 map: a < b (with -0 < +0) iff key(a) < key(b), so
// atomicMin/atomicMax on keys give the exact float min/max in any order.
__device__ __forceinline__ uint32_t fkey(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

// ------------------------------------------------------------------ Morton
// expandBits, src/metal/bvh.metal:26-32.
YSIM_HD uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
// 21 -> 63 bit spread (the 63-bit Morton codes).
YSIM_HD uint64_t expand21(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x001f00000000ffffull;
  v = (v | (v << 16)) & 0x001f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
// mortonCode + fillMortons_Tri normalisation, src/metal/bvh.metal:34-63.
YSIM_HD uint64_t morton_of(f3 x0, f3 x1, f3 x2, f3 bmin, f3 bmax, bool wide) {
  float cen[3] = {((x0.x + x1.x) + x2.x) / 3.0f, ((x0.y + x1.y) + x2.y) / 3.0f,
                  ((x0.z + x1.z) + x2.z) / 3.0f};
  float lo[3] = {bmin.x, bmin.y, bmin.z}, hi[3] = {bmax.x, bmax.y, bmax.z};
  uint64_t e[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float width = fmaxf(hi[k] - lo[k], 1e-8f);
    float c = (cen[k] - lo[k]) / width;
    if (!wide) {
      float p = fminf(fmaxf(c * 1024.0f, 0.0f), 1023.0f);
      e[k] = expand10((uint32_t)p);
    } else {
      float p = fminf(fmaxf(c * 2097152.0f, 0.0f), 2097151.0f);
      e[k] = expand21((uint64_t)(uint32_t)p);
    }
  }
  if (!wide) return e[0] * 4 + e[1] * 2 + e[2];
  return (e[0] << 2) | (e[1] << 1) | e[2];
}

// ------------------------------------------------------------------ narrow
// pointInTriangleBary, src/metal/bruteforce.metal:9-28.
YSIM_HD bool point_in_tri(f3 p, f3 v0, f3 v1) {
  float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(p, v0), d21 = dot(p, v1);
  float denom = d00 * d11 - d01 * d01;
  if (fabsf(denom) < 1e-12f) return false;
  float b = (d11 * d20 - d01 * d21) / denom;
  float c = (d00 * d21 - d01 * d20) / denom;
  float a = 1.0f - b - c;
  return a >= 0.0f && b >= 0.0f && c >= 0.0f;
}

// Square root of a squared length whose argument is 0, at least 2^-96 or not
// finite, or whose root is only compared with a bound above 2^-48 when the
// argument is smaller (its callers: narrow_test and the contact projection
// after a guard of 1e-12, calc_spring, which compares the root with 1e-7).
// On the device: the hardware root,
// then the neighbour below or above it whose square brackets the argument
// (two fma residuals) -- the compiler's correctly rounded square root without
// its scaling of arguments below 2^-96 and its return of +-0 / +inf as given,
// which leave every such argument's result unchanged.
YSIM_HD float sqrt_len(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u);
  const float su = __uint_as_float(__float_as_uint(s) + 1u);
  float r = s;
  if (fmaf(-sd, s, x) <= 0.0f) r = sd;
  if (fmaf(-su, s, x) > 0.0f) r = su;
  return r;
#else
  return sqrtf(x);
#endif
}

// narrow_pt_tri geometry, src/metal/bruteforce.metal:114-158: swept segment
// [qprev, qcur] against the triangle plane; n oriented toward qprev's side;
// fires when crossed or d_cur < band; contact point projected into the plane
// must lie inside (barycentric).  rsqrt is restated as 1/sqrt (IEEE).
YSIM_HD bool narrow_test(f3 qcur, f3 qprev, f3 t0, f3 t1, f3 t2, float band,
                                            f3& n_out, float& d_out) {
  f3 v0 = t1 - t0, v1 = t2 - t0;
  f3 n = cross(v0, v1);
  float nlen2 = dot(n, n);
  if (nlen2 < 1e-12f) return false;
  float inv = 1.0f / sqrt_len(nlen2);
  n = n * inv;
  float dprev = dot(n, qprev - t0);
  float dcur = dot(n, qcur - t0);
  float dref = (dprev != 0.0f) ? dprev : dcur;
  if (dref < 0.0f) {
    n = f3{-n.x, -n.y, -n.z};
    dprev = -dprev;
    dcur = -dcur;
  }
  bool crossed = dcur < 0.0f;
  bool inmargin = dcur < band;
  if (!crossed && !inmargin) return false;
  f3 cp;
  if (crossed && dprev > 0.0f) {
    float t = dprev / (dprev - dcur);
    cp = qprev + (qcur - qprev) * t;
  } else {
    cp = qcur;
  }
  f3 pc = cp - t0;
  float l = dot(n, pc);
  f3 inplane = pc - n * l;
  if (!point_in_tri(inplane, v0, v1)) return false;
  n_out = n;
  d_out = dcur;
  return true;
}

// intersectAABB, src/metal/bvh.metal:1285-1290 (closed intervals).
YSIM_HD bool overlap(f3 qmin, f3 qmax, f3 nmin, f3 nmax) {
  return !(qmax.x < nmin.x || qmin.x > nmax.x || qmax.y < nmin.y || qmin.y > nmax.y ||
           qmax.z < nmin.z || qmin.z > nmax.z);
}

// ------------------------------------------------------------------ springs
// calc_spring, src/metal/physics.metal:146-153.
YSIM_HD f3 calc_spring(f3 p0, f3 v0, f3 p1, f3 v1, float rest, float ks, float kd) {
  f3 dx = p1 - p0;
  f3 dv = v1 - v0;
  float len = sqrt_len(dot(dx, dx));
  if (len < 1e-7f) return f3{0.f, 0.f, 0.f};
  f3 ndx = {dx.x / len, dx.y / len, dx.z / len};
  float s = ks * (len - rest) + kd * dot(dv, ndx);
  return s * ndx;
}

// nonFinite3 / sanitizeIntegrateOutput, src/metal/physics.metal:40-74.
__device__ __forceinline__ bool nonfinite3(f3 p) {
  return ((__float_as_uint(p.x) & 0x7F800000u) == 0x7F800000u) ||
         ((__float_as_uint(p.y) & 0x7F800000u) == 0x7F800000u) ||
         ((__float_as_uint(p.z) & 0x7F800000u) == 0x7F800000u);
}
__device__ __forceinline__ bool sanitize(f3& pos, f3& vel, f3 fb) {
  const float B = 1.0e4f;
  bool touched = false;
  if (nonfinite3(pos)) {
    pos = nonfinite3(fb) ? f3{0.f, 0.f, 0.f} : fb;
    vel = f3{0.f, 0.f, 0.f};
    touched = true;
  }
  if (nonfinite3(vel)) {
    vel = f3{0.f, 0.f, 0.f};
    touched = true;
  }
  f3 bx = {fminf(fmaxf(pos.x, -B), B), fminf(fmaxf(pos.y, -B), B), fminf(fmaxf(pos.z, -B), B)};
  if (bx.x != pos.x || bx.y != pos.y || bx.z != pos.z) {
    pos = bx;
    vel = f3{0.f, 0.f, 0.f};
    touched = true;
  }
  return touched;
}

// ------------------------------------------------------------------ coherent access
// Device-scope (cross-XCD) coherent loads/stores used by the atomic bottom-up
// and the decoupled look-back: relaxed agent-scope atomics compile to sc1
// accesses that bypass the non-coherent L2 state of other XCDs.
__device__ __forceinline__ void st_agent(float* p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_agent(const float* p) {
  return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent_u32(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent_u32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent_u64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_agent_u64(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t pack2(float a, float b) {
  return (uint64_t)__float_as_uint(a) | ((uint64_t)__float_as_uint(b) << 32);
}
__device__ __forceinline__ float lo_f(uint64_t v) { return __uint_as_float((uint32_t)v); }
__device__ __forceinline__ float hi_f(uint64_t v) { return __uint_as_float((uint32_t)(v >> 32)); }
// Wait until every outstanding memory operation of this wave has completed.
__device__ __forceinline__ void wait_all_mem() { __builtin_amdgcn_s_waitcnt(0); }

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// XCD-aware workgroup order.  MI355X deals workgroup b of a launch to XCD
// (b mod 8), and each XCD has its own L2.  Returning a logical block index so
// that XCD x processes one contiguous chunk of the work keeps spatially
// coherent items (and the tree nodes they touch) in one L2.  A bijection on
// [0, nblocks).
constexpr uint32_t NUM_XCD = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
  const uint32_t x = b % NUM_XCD, k = b / NUM_XCD;
  const uint32_t per = nblocks / NUM_XCD, rem = nblocks % NUM_XCD;
  return x * per + min(x, rem) + k;
}

// Cross-lane reductions of a wave64 without LDS: DPP moves inside rows of 16
// lanes, then the gfx950 permlane swaps across rows.  Each step pairs every
// lane with one partner, so every lane ends with the result:
//   lane ^ 1, lane ^ 2 (quad_perm), lane 7 - i within 8 (row_half_mirror):
//     the 8-lane group; lane 15 - i within 16 (row_mirror): the row;
//   rows 0/1 and 2/3 (permlane16_swap), then the two halves (permlane32_swap).
// lane ^ 8 (row_ror:8), then the two swaps, reduce the 8 lanes of equal lane & 7.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140, DPP_ROR8 = 0x128;
template <class Op>
__device__ __forceinline__ uint32_t swap_rows_u32(uint32_t v, Op op) {
  const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  return op((uint32_t)p[0], (uint32_t)p[1]);
}
template <class Op>
__device__ __forceinline__ uint32_t swap_halves_u32(uint32_t v, Op op) {
  const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return op((uint32_t)p[0], (uint32_t)p[1]);
}
template <class Op>
__device__ __forceinline__ uint32_t group8_reduce_u32(uint32_t v, Op op) {
  v = op(v, dpp_u32<DPP_XOR1>(v));
  v = op(v, dpp_u32<DPP_XOR2>(v));
  return op(v, dpp_u32<DPP_HALF_MIRROR>(v));
}
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t v, Op op) {
  v = group8_reduce_u32(v, op);
  v = op(v, dpp_u32<DPP_MIRROR>(v));
  return swap_halves_u32(swap_rows_u32(v, op), op);
}
// over the lanes of equal lane & 7 (lane ^ 8, ^ 16, ^ 32)
template <class Op>
__device__ __forceinline__ uint32_t column8_reduce_u32(uint32_t v, Op op) {
  v = op(v, dpp_u32<DPP_ROR8>(v));
  return swap_halves_u32(swap_rows_u32(v, op), op);
}
struct MinU32 {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return min(a, b); }
};
struct MaxU32 {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return max(a, b); }
};
struct AddU32 {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) { return wave_reduce_u32(v, MinU32{}); }
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) { return wave_reduce_u32(v, MaxU32{}); }
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) { return wave_reduce_u32(v, AddU32{}); }

// Box of a whole grid of NT-thread blocks into acc[0..5] (ordered keys fkey:
// min x, y, z, max x, y, z), from one (lo, hi) per thread (+inf / -inf for a
// thread with no point).  Two levels of last-arrival reduction, so no counter
// takes more than BOX_GROUP same-address atomics (they serialize at the memory
// side): each block writes its box to parts with device-coherent stores and
// arrives at its group's counter; the last block of a group reduces the
// group's boxes, writes the group box and arrives at the top counter; the last
// group reduces the group boxes into acc.  Each last arrival resets the
// counter it read.  parts layout (grid_box_parts_words(gridDim.x) words, zero
// before the first launch): top counter, BOX_MAX_GROUPS group counters (at
// fixed places, so launches of other grid sizes sharing the buffer find them
// zero), 6 words per block, 6 words per group.  At most BOX_GROUP x
// BOX_MAX_GROUPS blocks.  Every thread of the block must call it.
constexpr uint32_t BOX_GROUP = 32, BOX_MAX_GROUPS = 4096;
inline uint32_t grid_box_parts_words(uint32_t nblocks) {
  const uint32_t ng = (nblocks + BOX_GROUP - 1) / BOX_GROUP;
  return 1 + BOX_MAX_GROUPS + 6 * nblocks + 6 * ng;
}
inline bool grid_box_fits(uint32_t nblocks) { return nblocks <= BOX_GROUP * BOX_MAX_GROUPS; }
// min (k < 3) or max (k >= 3) of 6-word records src[0 .. count) into threads
// 0..5's return value (other threads: unspecified); a block-wide step.  Each
// thread folds whole records, then wave reductions and one LDS pass over the
// waves (s_acc: at least 6 NT / 64 words).  COHERENT: records written by other
// workgroups of the running grid.
template <int NT, bool COHERENT = true>
__device__ __forceinline__ uint32_t reduce_box_records(const uint32_t* src, uint32_t count, uint32_t* s_acc) {
  constexpr uint32_t NW = NT / 64;
  uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (uint32_t r = threadIdx.x; r < count; r += NT)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const uint32_t u = COHERENT ? ld_agent_u32(src + 6 * (size_t)r + k) : src[6 * (size_t)r + k];
      v[k] = k < 3 ? min(v[k], u) : max(v[k], u);
    }
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = k < 3 ? wave_min_u32(v[k]) : wave_max_u32(v[k]);
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  __syncthreads();  // s_acc may still be read by an earlier call
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; ++k) s_acc[k * NW + w] = v[k];
  __syncthreads();
  uint32_t out = 0;
  if (threadIdx.x < 6) {
    const uint32_t k = threadIdx.x;
    out = s_acc[k * NW];
    for (uint32_t ww = 1; ww < NW; ++ww) out = k < 3 ? min(out, s_acc[k * NW + ww]) : max(out, s_acc[k * NW + ww]);
  }
  return out;
}
// The block's box (ordered keys, as grid_box_reduce) written by threads 0..5
// to out[0..5] with plain stores.  Every thread of the block must call it.
template <int NT>
__device__ __forceinline__ void block_box_store(f3 lo, f3 hi, uint32_t* out) {
  constexpr uint32_t NW = NT / 64;
  __shared__ uint32_t s_red[6][NW];
  const uint32_t r[6] = {wave_min_u32(fkey(lo.x)), wave_min_u32(fkey(lo.y)), wave_min_u32(fkey(lo.z)),
                         wave_max_u32(fkey(hi.x)), wave_max_u32(fkey(hi.y)), wave_max_u32(fkey(hi.z))};
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) s_red[k][w] = r[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const uint32_t k = threadIdx.x;
    uint32_t v = s_red[k][0];
    for (uint32_t ww = 1; ww < NW; ++ww) v = k < 3 ? min(v, s_red[k][ww]) : max(v, s_red[k][ww]);
    out[k] = v;
  }
}

template <int NT>
__device__ __forceinline__ void grid_box_reduce(f3 lo, f3 hi, uint32_t* parts, uint32_t* acc) {
  constexpr uint32_t NW = NT / 64;
  __shared__ uint32_t s_red[6][NW];
  __shared__ uint32_t s_acc[NT];
  __shared__ uint32_t s_last;
  const uint32_t nb = gridDim.x, ng = (nb + BOX_GROUP - 1) / BOX_GROUP;
  const uint32_t b = blockIdx.x, g = b / BOX_GROUP;
  uint32_t* top = parts;
  uint32_t* gcnt = parts + 1;
  uint32_t* bbox = parts + 1 + BOX_MAX_GROUPS;
  uint32_t* gbox = bbox + 6 * (size_t)nb;
  const uint32_t r[6] = {wave_min_u32(fkey(lo.x)), wave_min_u32(fkey(lo.y)), wave_min_u32(fkey(lo.z)),
                         wave_max_u32(fkey(hi.x)), wave_max_u32(fkey(hi.y)), wave_max_u32(fkey(hi.z))};
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) s_red[k][w] = r[k];
  __syncthreads();
  if (threadIdx.x < 64) {  // wave 0 publishes the block's box, then arrives
    if (threadIdx.x < 6) {
      const uint32_t k = threadIdx.x;
      uint32_t v = s_red[k][0];
      for (uint32_t ww = 1; ww < NW; ++ww) v = k < 3 ? min(v, s_red[k][ww]) : max(v, s_red[k][ww]);
      st_agent_u32(bbox + 6 * (size_t)b + k, v);
    }
    wait_all_mem();
    if (threadIdx.x == 0) {
      const uint32_t gsize = min(BOX_GROUP, nb - g * BOX_GROUP);
      const uint32_t old = __hip_atomic_fetch_add(gcnt + g, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_last = (old == gsize - 1) ? 1u : 0u;
    }
  }
  __syncthreads();
  if (!s_last) return;
  // last block of group g
  {
    const uint32_t gsize = min(BOX_GROUP, nb - g * BOX_GROUP);
    const uint32_t v = reduce_box_records<NT>(bbox + 6 * (size_t)g * BOX_GROUP, gsize, s_acc);
    if (threadIdx.x < 64) {
      if (threadIdx.x < 6) st_agent_u32(gbox + 6 * (size_t)g + threadIdx.x, v);
      wait_all_mem();
      if (threadIdx.x == 0) {
        gcnt[g] = 0u;
        const uint32_t old = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (old == ng - 1) ? 1u : 0u;
      }
    }
    __syncthreads();
    if (!s_last) return;
  }
  // last group
  const uint32_t v = reduce_box_records<NT>(gbox, ng, s_acc);
  if (threadIdx.x < 6) acc[threadIdx.x] = v;
  if (threadIdx.x == 0) *top = 0u;
}

}  // namespace ysim
