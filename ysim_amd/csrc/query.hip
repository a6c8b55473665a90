// query.hip — BVH traversal fused with the point-triangle narrow phase.
//
// queryPoints (src/metal/bvh.metal:1393-1412) reports, for every query vertex
// and every target tree of its collision group, each leaf whose box overlaps
// [x - m, x + m] (closed; facets containing the vertex skipped on self queries,
// bvh.metal:1366-1371); each candidate is then tested with the swept
// point-triangle narrow phase (narrow_pt_tri, src/metal/bruteforce.metal:41-178)
// and contacts are kept in per-vertex slots.  Contacts are applied later in
// (target mesh, facet) order, so the result does not depend on the order in
// which candidates are found.  Two kernels compute the same slots:
//
//   k_query_tile  one wave per 64 query vertices (an 8 x 8 vertex tile of a
//                 grid cloth).  The wave walks the target tree breadth-first
//                 with the union of its 64 boxes, one tree level per round
//                 with a whole wave of lanes, collects the overlapping leaves,
//                 distributes them to the vertices through 16 sub-boxes of
//                 2 x 2 vertices, and runs the narrow phase on the resulting
//                 (vertex, facet) pairs with every lane busy.  A tree node is
//                 loaded once per tile instead of once per vertex.
//   k_query       one thread per query vertex: a short-stack walk with the
//                 narrow phase inline (also the pair-emission broad phase).
//
// Candidate-pair statistics are per-wave running counts (each wave owns one
// word; no atomics): same-address atomics from thousands of waves serialize
// and cost tens of microseconds on this GPU.
#include "kernels.hpp"
#include "traverse.hpp"

namespace ysim {

namespace {
constexpr int SDEPTH = 8;          // LDS stack entries per thread (deeper stacks overflow no less on cloth)
constexpr uint32_t TQ_STACK = 128; // tile walk: pending child pairs
constexpr uint32_t TQ_LEAVES = 192;
constexpr uint32_t TQ_PAIRS = 256;  // one emission round adds up to 4 pairs per lane
constexpr uint32_t TQ_HEAVY = 16;   // a leaf with more candidate vertices is tested by the whole wave
static_assert(TQ_PAIRS >= 64 * 4, "an emission round must fit the pair buffer after a flush");
static_assert(TQ_LEAVES >= 2 * 64, "a walk round adds up to 2 leaves per lane");
// LDS words of the tile kernel's shared buffer: pending child pairs (2 each),
// candidate pairs (1 each), the triangles of a batch of 64 leaves (9 floats
// each) and their facet ids; the whole buffer is also the fallback per-lane stack
constexpr uint32_t TQ_BUF_WORDS = 2 * TQ_STACK + TQ_PAIRS + 64 * 9 + 64;
static_assert(TQ_BUF_WORDS % 128 == 0, "whole per-lane stack entries");
static_assert(TQ_BUF_WORDS * 4 >= 2 * 144 * 16, "the cloth step stages a 12 x 12 halo of x and v in the buffer");
constexpr int TQ_SDEPTH = TQ_BUF_WORDS / 128;  // fallback per-lane stack depth of the tile kernel

// Set bits of the wave-uniform mask m at lanes below this one.
__device__ __forceinline__ uint32_t count_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// The crossings of rows rb and columns cb (8 bits each) of an 8 x 8 tile as
// a lane mask (lane 8 r + c): bit r of rb spread to row r's byte (times
// 0x204081, a 24-bit product, moves bit i of a nibble to bit 8 i; x 255 as
// (y << 8) - y fills the byte), every byte masked by cb.
__device__ __forceinline__ uint64_t cross_mask(uint32_t rb, uint32_t cb) {
  const uint32_t c2 = cb | (cb << 8), c4 = c2 | (c2 << 16);
  const uint32_t ylo = __umul24(rb & 0xfu, 0x00204081u) & 0x01010101u;
  const uint32_t yhi = __umul24((rb >> 4) & 0xfu, 0x00204081u) & 0x01010101u;
  const uint32_t lo = ((ylo << 8) - ylo) & c4, hi = ((yhi << 8) - yhi) & c4;
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t popc64(uint64_t v) { return (uint32_t)__popcll(v); }

// LDS handoff between the lanes of one wave (the tile kernel runs one wave per
// workgroup): all earlier LDS writes of the wave are visible to later reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Adds a wave's candidate count to its own statistics word.
__device__ __forceinline__ void wave_stats(const QueryArgs& a, uint32_t wave, uint32_t npair) {
  const uint32_t s = wave_sum_u32(npair);
  if (a.wstats && __lane_id() == 0 && s) a.wstats[wave] += s;
}

// One query vertex against one target tree, narrow phase inline; cnt is the
// vertex's running slot count.  Returns false if the walk's guard fired.
template <bool EMIT_PAIRS, int SD>
__device__ __forceinline__ bool lane_target(const QueryArgs& a, uint32_t m, uint32_t p, uint32_t j, uint32_t t,
                                            f3 qcur, f3 qprev, f3 qmin, f3 qmax, float band, int2* stk,
                                            int stride, uint32_t& cnt, uint32_t& npair) {
  const uint32_t nT = a.mesh_nfacets[t];
  if (nT == 0) return true;
  const ysim_bvh_node* T = a.tree + a.mt.noff[t];
  const int32_t* P = a.parent + a.mt.noff[t];
  const uint32_t* F = a.facets + 3 * (size_t)a.mt.foff[t];
  const uint32_t tvb = a.mt.voff[t];
  const bool self = (t == m);
  const uint32_t cap = a.cap;
  const int it = traverse<SD>(T, P, nT, qmin, qmax, stk, stride, [&](int fid) {
    const uint32_t* f = F + 3 * (size_t)fid;
    const uint32_t f0 = f[0], f1 = f[1], f2 = f[2];
    if (self && (p == f0 || p == f1 || p == f2)) return;
    ++npair;
    if (a.audit) audit_lane(a, j, t, (uint32_t)fid);
    if (EMIT_PAIRS) {
      uint32_t slot = atomicAdd(a.npairs, 1u);
      if (slot < a.pair_cap) reinterpret_cast<uint4*>(a.pairs)[slot] = make_uint4(m, p, t, (uint32_t)fid);
      return;
    }
    f3 nrm;
    float d;
    if (narrow_test(qcur, qprev, ld3(a.x, tvb + f0), ld3(a.x, tvb + f1), ld3(a.x, tvb + f2), band, nrm, d)) {
      if (cnt < cap) {
        a.cnd[(size_t)j * cap + cnt] = make_float4(nrm.x, nrm.y, nrm.z, d);
        a.cid[(size_t)j * cap + cnt] = make_uint2(t, (uint32_t)fid);
      }
      ++cnt;
    }
  });
  return it >= 0;
}

// Query vertex of item tq: mesh-local vertex p (grid cloths in tile order).
__device__ __forceinline__ uint32_t query_vertex(const QueryArgs& a, uint32_t sg, uint32_t m, uint32_t tq) {
  const uint32_t gn = a.mesh_grid_n ? a.mesh_grid_n[m] : 0u;
  return gn ? tile_order(tq - a.qsegs.off[sg], gn) : tq - a.qsegs.off[sg];
}
}  // namespace

// QB threads per block (64, 128 or 256) and an occupancy target of WPE waves
// per SIMD (5: registers as needed; 6-8: fewer registers, some spilled).
template <bool EMIT_PAIRS, int QB, int WPE>
__attribute__((amdgpu_waves_per_eu(WPE, WPE))) __global__ __launch_bounds__(QB) void k_query(QueryArgs a) {
  __shared__ int2 s_stack[SDEPTH * QB];
  const uint32_t tq = xcd_block(blockIdx.x, gridDim.x) * QB + threadIdx.x;
  uint32_t npair = 0;
  if (tq < a.nitems) {
    const uint32_t sg = seg_of(a.qsegs, tq);
    const uint32_t m = a.qsegs.mesh[sg];
    const uint32_t p = query_vertex(a, sg, m, tq);
    const uint32_t j = a.qsegs.off[sg] + p;  // contact slots stay indexed by vertex
    const uint32_t g = a.mt.voff[m] + p;
    const f3 qcur = ld3(a.x, g);
    const f3 qprev = ld3(a.xprev, g);
    const float margin = a.mesh_margin[m];
    const float band = a.mesh_band[m];
    const f3 qmin = {qcur.x - margin, qcur.y - margin, qcur.z - margin};
    const f3 qmax = {qcur.x + margin, qcur.y + margin, qcur.z + margin};
    uint32_t cnt = 0;
    for (uint32_t ti = a.tgt_off[sg]; ti < a.tgt_off[sg + 1]; ++ti)
      if (!lane_target<EMIT_PAIRS, SDEPTH>(a, m, p, j, a.tgt[ti], qcur, qprev, qmin, qmax, band,
                                           s_stack + threadIdx.x, QB, cnt, npair))
        atomicOr(a.flags, 4u);
    if (!EMIT_PAIRS) {
      a.ccount[j] = cnt;
      if (cnt > a.cap) atomicOr(a.flags, 2u);
    }
  }
  if (!EMIT_PAIRS) wave_stats(a, (xcd_block(blockIdx.x, gridDim.x) * QB + threadIdx.x) >> 6, npair);
}

// The 12 x 12 halo of x and v around the full 8 x 8 tile (r0, c0) of an N x N
// grid cloth at global vertex offset vb (entry e = lane + 64 k, vertex (r0 +
// e / 12 - 2, c0 + e % 12 - 2), loaded when inside the cloth), in registers
// until halo_store writes it to LDS; every load is issued before any is used.
struct TileHalo {
  f3 x[3], v[3];
  bool ok[3];
};
__device__ __forceinline__ TileHalo halo_load(const ClothArgs& ca, uint32_t r0, uint32_t c0, uint32_t N,
                                              uint32_t vb) {
  const uint32_t lane = __lane_id();
  TileHalo h;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t e = lane + 64u * k;
    const int r = (int)(r0 + e / 12u) - 2, c = (int)(c0 + e % 12u) - 2;
    h.ok[k] = e < 144u && r >= 0 && c >= 0 && r < (int)N && c < (int)N;
    if (h.ok[k]) {
      const uint32_t q = vb + (uint32_t)r * N + (uint32_t)c;
      h.x[k] = ld3(ca.x, q);
      h.v[k] = ld3(ca.v, q);
    }
  }
  return h;
}
// Entries as float4 at HX[e] and HV[e] (GridTile layout); visible to the
// whole wave after the wave_sync.
__device__ __forceinline__ void halo_store(const TileHalo& h, float4* HX, float4* HV) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (h.ok[k]) {
      HX[lane + 64u * k] = make_float4(h.x[k].x, h.x[k].y, h.x[k].z, 0.f);
      HV[lane + 64u * k] = make_float4(h.v[k].x, h.v[k].y, h.v[k].z, 0.f);
    }
  wave_sync();
}

// After a tile's query (CLOTH variant): the cloth step of its vertices, and
// the box of their new positions as 6 ordered keys at box_parts + 6 tile.
// grid_n > 0: the lanes are vertices of one N x N grid cloth, mesh m at
// global vertex offset vb, in tile order (j local); a full 8 x 8 tile stages
// its 12 x 12 halo of x and v in LDS (s_halo, 288 float4) and the springs
// read their neighbours there.
// Contact slots were written by any lane of the wave: their stores complete
// before the wave barrier (the wave's loads of its own slot lines, never read
// before in this kernel, then come from L2); the count comes in a register.
__device__ __forceinline__ void tile_cloth(const ClothArgs& ca, bool active, uint32_t j, uint32_t count,
                                           uint32_t tile, uint32_t grid_n, uint32_t m, uint32_t vb,
                                           float4* s_halo) {
  wait_all_mem();
  wave_sync();
  const float INF = __uint_as_float(0x7f800000u);
  f3 lo = {INF, INF, INF}, hi = {-INF, -INF, -INF};
  const uint32_t nt = grid_n >> 3;
  if (grid_n && tile < nt * nt) {
    // a full tile (every lane active, lane 8 r + c on vertex (r0 + r, c0 + c))
    const uint32_t r0 = (tile / nt) * 8, c0 = (tile % nt) * 8;
    const TileHalo halo = halo_load(ca, r0, c0, grid_n, vb);
    const f3 ext = ld3(ca.ext, vb + j);
    float4* HX = s_halo;
    float4* HV = s_halo + 144;
    halo_store(halo, HX, HV);
    const uint32_t r = __lane_id() >> 3, c = __lane_id() & 7u;
    const ClothForce cf = cloth_force_at(ca, 0u, m, j, vb, GridTile{HX, HV, r0 + r, c0 + c, (r + 2) * 12 + c + 2, ext});
    lo = hi = cloth_apply_at(ca, j, m, vb + j, (int32_t)count, cf);
  } else if (active) {
    lo = hi = cloth_vertex(ca, j, (int32_t)count);
  }
  if (!ca.box_parts) return;
  const uint32_t lane = __lane_id();
  const uint32_t r0 = wave_min_u32(fkey(lo.x)), r1 = wave_min_u32(fkey(lo.y)), r2 = wave_min_u32(fkey(lo.z));
  const uint32_t r3 = wave_max_u32(fkey(hi.x)), r4 = wave_max_u32(fkey(hi.y)), r5 = wave_max_u32(fkey(hi.z));
  if (lane < 6)
    ca.box_parts[6 * (size_t)tile + lane] =
        lane == 0 ? r0 : lane == 1 ? r1 : lane == 2 ? r2 : lane == 3 ? r3 : lane == 4 ? r4 : r5;
}

// One wave per 64 consecutive query items (lane 8 r + c: row r, column c of
// an 8 x 8 vertex tile for grid cloths).  Waves whose items span two query
// segments, or that hold a non-finite query box, take the per-lane walk.
// When the leaf list fills up, the leaves found so far go through phases 2-3
// and the walk goes on.  When the walk's pending list outgrows its buffer, the
// (wave, target) is redone per lane from the slot counts it started with (the
// slots it filled are overwritten).  a.tile == 2 shrinks both buffers (128
// leaves, 32 pending pairs) so that the leaf flushes and the per-lane redo
// both run (a test mode for those paths).
// CLOTH: the tile's cloth step follows its query (tile_cloth, one call site
// for both paths); its registers keep that variant at 4 waves per SIMD (at 5
// it spills), the query alone runs 5 (LDS 8.6 KB per wave).
template <bool CLOTH>
__attribute__((amdgpu_waves_per_eu(CLOTH ? 4 : 5, CLOTH ? 4 : 5))) __global__ __launch_bounds__(64) void k_query_tile(QueryArgs a,
                                                                                               ClothArgs ca) {
  // s_buf: TQ_STACK child-link pairs of overlapping internal nodes, TQ_PAIRS
  // candidates (lane | batch slot << 6), the corners and facet ids of the
  // current batch of leaves; the whole buffer is the per-lane stack
  // (TQ_SDEPTH entries) of the fallback walks
  __shared__ __align__(16) int2 s_buf[TQ_BUF_WORDS / 2];
  __shared__ uint2 s_leaf[TQ_LEAVES];  // (leaf node, facet) overlapping the union box
  __shared__ float s_q[64 * 6];        // qcur, qprev per lane
  __shared__ float s_rc[16 * 6];       // boxes of the tile's 8 rows, then of its 8 columns
  __shared__ uint32_t s_cnt[64], s_p[64], s_j[64];
  int2* s_stk = s_buf;
  uint32_t* s_pair = reinterpret_cast<uint32_t*>(s_buf + TQ_STACK);
  float* s_tri = reinterpret_cast<float*>(s_pair + TQ_PAIRS);
  uint32_t* s_bfid = reinterpret_cast<uint32_t*>(s_tri + 64 * 9);
  const uint32_t lane = threadIdx.x;
  const uint32_t w = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t base = w * 64;
  const uint32_t tq = base + lane;
  const bool active = tq < a.nitems;
  const uint32_t last = min(base + 63u, a.nitems - 1u);
  const uint32_t sg = seg_of(a.qsegs, last);  // wave-uniform from here on
  uint32_t npair = 0;
  int2* fstk = s_buf + lane;
  // a wave whose items span two query segments, or that holds a non-finite
  // query box (a NaN box overlaps every box under the closed test; the union
  // box would ignore it), walks per lane over each lane's own targets
  bool per_lane = seg_of(a.qsegs, base) != sg;
  // the cooperative path's lane data (idle lanes take the wave's last item);
  // with one query segment (pre_t >= 0) the host gave its tables: the
  // segment starts at item 0
  const bool one = a.pre_t >= 0;
  const uint32_t m = one ? a.seg0_mesh : a.qsegs.mesh[sg];
  const uint32_t it = active ? tq : last;
  const uint32_t p = one ? (a.seg0_grid_n ? tile_order(it, a.seg0_grid_n) : it) : query_vertex(a, sg, m, it);
  const uint32_t j = one ? p : a.qsegs.off[sg] + p;
  const uint32_t g = (one ? a.seg0_voff : a.mt.voff[m]) + p;
  const uint32_t ti0 = a.tgt_off[sg], ti1 = a.tgt_off[sg + 1];
  f3 qcur = {0.f, 0.f, 0.f}, qprev = {0.f, 0.f, 0.f};
  float margin = 0.f, band = 0.f;
  // the first target's cut when it has at most 3 x 64 entries and fits the
  // leaf list: entry and node of this lane for batches 0-2 in (pn0, pv0) ..
  // (pn2, pv2), loaded while the position loads, the finiteness ballot and
  // the box reductions below run; scanned before the target loop (npre > 0)
  const uint32_t leaf_cap = a.tile == 2 ? 128u : TQ_LEAVES;  // >= 2 x 64
  const uint32_t stack_cap = a.tile == 2 ? 32u : TQ_STACK;
  uint32_t npre = 0;
  int pn0 = 0, pn1 = 0, pn2 = 0;
  NodeV pv0, pv1, pv2;
  if (!per_lane) {
    qcur = ld3(a.x, g);
    qprev = ld3(a.xprev, g);
    margin = a.mesh_margin[m];
    band = a.mesh_band[m];
    if (one) {
      // the host lists only targets with facets, and the build filled the
      // list of a one-segment group with -1 before appending: the entries
      // load without the count, which decides later whether they are used
      const int32_t* CL = a.cut + (size_t)a.pre_t * CUT_CAP;
      pn0 = CL[lane];
      pn1 = CL[64 + lane];
      pn2 = CL[128 + lane];
      // no branch between the three node loads (an unused entry loads node 0,
      // which every listed target has), so they are in flight together
      pv0 = load_node(a.pre_tree, pn0 >= 0 ? pn0 : 0);
      pv1 = load_node(a.pre_tree, pn1 >= 0 ? pn1 : 0);
      pv2 = load_node(a.pre_tree, pn2 >= 0 ? pn2 : 0);
      const uint32_t ncut = a.cut_cnt ? a.cut_cnt[a.pre_t] : 0u;
      if (ncut > 0 && ncut <= CUT_CAP && ncut <= 3 * 64 && ncut <= leaf_cap) npre = ncut;
    } else if (ti0 < ti1) {
      const uint32_t t = a.tgt[ti0];
      const uint32_t ncut = a.cut_cnt ? a.cut_cnt[t] : 0u;
      if (a.mesh_nfacets[t] > 0 && ncut > 0 && ncut <= CUT_CAP && ncut <= 3 * 64 && ncut <= leaf_cap) {
        const int32_t* CL = a.cut + (size_t)t * CUT_CAP;
        const ysim_bvh_node* T = a.tree + a.mt.noff[t];
        npre = ncut;
        if (lane < ncut) {
          pn0 = CL[lane];
          pv0 = load_node(T, pn0);
        }
        if (64 + lane < ncut) {
          pn1 = CL[64 + lane];
          pv1 = load_node(T, pn1);
        }
        if (128 + lane < ncut) {
          pn2 = CL[128 + lane];
          pv2 = load_node(T, pn2);
        }
      }
    }
    const f3 q = active ? qcur : f3{0.f, 0.f, 0.f};
    const float mg = margin;
    const bool finite = __builtin_isfinite(q.x - mg) && __builtin_isfinite(q.y - mg) &&
                        __builtin_isfinite(q.z - mg) && __builtin_isfinite(q.x + mg) &&
                        __builtin_isfinite(q.y + mg) && __builtin_isfinite(q.z + mg);
    per_lane = __ballot(!finite) != 0;
  }
  uint32_t jc = 0, cc = 0;  // the lane's slot row and contact count, for the cloth step
  if (per_lane) {
    if (active) {
      uint32_t cnt = 0;
      const uint32_t sl = seg_of(a.qsegs, tq);
      const uint32_t ml = a.qsegs.mesh[sl];
      const uint32_t pl = query_vertex(a, sl, ml, tq);
      const uint32_t jl = a.qsegs.off[sl] + pl;
      const f3 qc = ld3(a.x, a.mt.voff[ml] + pl), qp = ld3(a.xprev, a.mt.voff[ml] + pl);
      const float mg = a.mesh_margin[ml];
      const f3 qlo = {qc.x - mg, qc.y - mg, qc.z - mg}, qhi = {qc.x + mg, qc.y + mg, qc.z + mg};
      for (uint32_t ti = a.tgt_off[sl]; ti < a.tgt_off[sl + 1]; ++ti)
        if (!lane_target<false, TQ_SDEPTH>(a, ml, pl, jl, a.tgt[ti], qc, qp, qlo, qhi, a.mesh_band[ml], fstk, 64,
                                        cnt, npair))
          atomicOr(a.flags, 4u);
      a.ccount[jl] = cnt;
      if (cnt > a.cap) atomicOr(a.flags, 2u);
      jc = jl;
      cc = cnt;
    }
  } else {
    const f3 qmin = {qcur.x - margin, qcur.y - margin, qcur.z - margin};
    const f3 qmax = {qcur.x + margin, qcur.y + margin, qcur.z + margin};
    // per-lane query data for the cooperative phases; idle lanes carry an empty box
    const float INF = __uint_as_float(0x7f800000u);
    f3 lo = active ? qmin : f3{INF, INF, INF}, hi = active ? qmax : f3{-INF, -INF, -INF};
    s_q[lane * 6 + 0] = qcur.x;
    s_q[lane * 6 + 1] = qcur.y;
    s_q[lane * 6 + 2] = qcur.z;
    s_q[lane * 6 + 3] = qprev.x;
    s_q[lane * 6 + 4] = qprev.y;
    s_q[lane * 6 + 5] = qprev.z;
    s_cnt[lane] = 0;
    s_p[lane] = p;
    s_j[lane] = j;
    // min / max of the box bounds over lane groups, with the LDS-free
    // reductions of common.hpp (exact in any order; the boxes are only
    // compared, so the sign of a zero bound does not matter)
    const auto fmin_u = [](uint32_t x, uint32_t y) { return __float_as_uint(fminf(__uint_as_float(x), __uint_as_float(y))); };
    const auto fmax_u = [](uint32_t x, uint32_t y) { return __float_as_uint(fmaxf(__uint_as_float(x), __uint_as_float(y))); };
    auto reduce = [&](auto&& red) {
      lo = {__uint_as_float(red(__float_as_uint(lo.x), fmin_u)), __uint_as_float(red(__float_as_uint(lo.y), fmin_u)),
            __uint_as_float(red(__float_as_uint(lo.z), fmin_u))};
      hi = {__uint_as_float(red(__float_as_uint(hi.x), fmax_u)), __uint_as_float(red(__float_as_uint(hi.y), fmax_u)),
            __uint_as_float(red(__float_as_uint(hi.z), fmax_u))};
    };
    const auto rows8 = [](uint32_t v, auto op) { return group8_reduce_u32(v, op); };
    const auto rows_to_wave = [](uint32_t v, auto op) {
      v = op(v, dpp_u32<DPP_MIRROR>(v));
      return swap_halves_u32(swap_rows_u32(v, op), op);
    };
    const auto cols8 = [](uint32_t v, auto op) { return column8_reduce_u32(v, op); };
    // row r holds lanes 8 r .. 8 r + 7, column c lanes c, c + 8, .., c + 56; a
    // vertex whose box overlaps a leaf's lies in a row and a column that do too
    auto put = [&](uint32_t b) {
      s_rc[b * 6 + 0] = lo.x;
      s_rc[b * 6 + 1] = lo.y;
      s_rc[b * 6 + 2] = lo.z;
      s_rc[b * 6 + 3] = hi.x;
      s_rc[b * 6 + 4] = hi.y;
      s_rc[b * 6 + 5] = hi.z;
    };
    const f3 lo1 = lo, hi1 = hi;
    reduce(rows8);
    if ((lane & 7u) == 0) put(lane >> 3);
    reduce(rows_to_wave);
    const f3 Ulo = lo, Uhi = hi;  // union of the wave's boxes
    lo = lo1;
    hi = hi1;
    reduce(cols8);
    if (lane < 8) put(8 + lane);
    // the prefetched cut of the first target: pending pairs and leaves
    // (sp_pre, nl_pre; no leaf flush: the cut fits the list)
    uint32_t sp_pre = 0, nl_pre = 0;
    bool ovf_pre = false;
    auto pre_batch = [&](uint32_t c0, int node, const NodeV& A) {
      // branch-free: the flags stay lane masks for the ballots
      const bool h = (c0 + lane < npre) & overlap(Ulo, Uhi, A.lo, A.hi);
      const bool pc = h & (A.a >= 0), lc = h & (A.a < 0);
      const uint64_t bp = __ballot(pc), bl = __ballot(lc);
      if (sp_pre + popc64(bp) > stack_cap) {
        ovf_pre = true;
        return;
      }
      if (pc) s_stk[sp_pre + count_below(bp)] = make_int2(A.a, A.b);
      if (lc) s_leaf[nl_pre + count_below(bl)] = make_uint2((uint32_t)node, (uint32_t)A.b);
      sp_pre += popc64(bp);
      nl_pre += popc64(bl);
    };
    if (npre > 0) pre_batch(0, pn0, pv0);
    if (npre > 64 && !ovf_pre) pre_batch(64, pn1, pv1);
    if (npre > 128 && !ovf_pre) pre_batch(128, pn2, pv2);
    wave_sync();

    for (uint32_t ti = ti0; ti < ti1; ++ti) {
      const uint32_t t = a.tgt[ti];
      const uint32_t nT = a.mesh_nfacets[t];
      if (nT == 0) continue;
      const ysim_bvh_node* T = a.tree + a.mt.noff[t];
      const uint32_t* F = a.facets + 3 * (size_t)a.mt.foff[t];
      const uint32_t tvb = a.mt.voff[t];
      const bool self = (t == m);
      const uint32_t cnt0 = s_cnt[lane], npair0 = npair;  // for a per-lane redo of this target
      uint32_t np = 0;
      // ---- phase 3 (after each batch of leaves, and whenever the pair buffer
      // fills): narrow phase per pair, the triangle's corners from LDS
      auto flush = [&]() {
        for (uint32_t c = 0; c < np; c += 64) {
          if (c + lane < np) {
            const uint32_t pr = s_pair[c + lane];
            const uint32_t v = pr & 63u, sl = pr >> 6;
            const uint32_t fid = s_bfid[sl];
            const float* tp = s_tri + 9 * sl;
            const f3 t0 = {tp[0], tp[1], tp[2]}, t1 = {tp[3], tp[4], tp[5]}, t2 = {tp[6], tp[7], tp[8]};
            const f3 vc = {s_q[v * 6 + 0], s_q[v * 6 + 1], s_q[v * 6 + 2]};
            const f3 vp = {s_q[v * 6 + 3], s_q[v * 6 + 4], s_q[v * 6 + 5]};
            f3 nrm;
            float d;
            if (narrow_test(vc, vp, t0, t1, t2, band, nrm, d)) {
              const uint32_t slot = atomicAdd(&s_cnt[v], 1u);
              if (slot < a.cap) {
                const uint32_t jv = s_j[v];
                a.cnd[(size_t)jv * a.cap + slot] = make_float4(nrm.x, nrm.y, nrm.z, d);
                a.cid[(size_t)jv * a.cap + slot] = make_uint2(t, fid);
              }
            }
          }
        }
        np = 0;
        wave_sync();
      };
      // ---- phase 2 (whenever the leaf list fills): leaves -> (vertex, facet)
      // candidates, in batches of 64 leaves (one per lane, batch slot = lane)
      auto leaves = [&](uint32_t nl) {
        for (uint32_t c0 = 0; c0 < nl; c0 += 64) {
          const bool have = c0 + lane < nl;
          uint32_t fid = 0, f0 = 0, f1 = 0, f2 = 0;
          uint64_t cand = 0;
          NodeV L;
          f3 p0, p1, p2;
          if (have) {
            const uint2 le = s_leaf[c0 + lane];
            L = load_node(T, (int)le.x);
            fid = le.y;
            const uint32_t* f = F + 3 * (size_t)fid;
            f0 = f[0];
            f1 = f[1];
            f2 = f[2];
            p0 = ld3(a.x, tvb + f0);  // for the narrow phase; in flight during the tests
            p1 = ld3(a.x, tvb + f1);
            p2 = ld3(a.x, tvb + f2);
            // rows and columns overlapping the leaf (bit r, bit c);
            // candidates = their crossings
            uint32_t rb = 0, cb = 0;
#pragma unroll 2
            for (int r = 0; r < 8; ++r) {
              const f3 bl = {s_rc[r * 6 + 0], s_rc[r * 6 + 1], s_rc[r * 6 + 2]};
              const f3 bh = {s_rc[r * 6 + 3], s_rc[r * 6 + 4], s_rc[r * 6 + 5]};
              if (overlap(bl, bh, L.lo, L.hi)) rb |= 1u << r;
            }
#pragma unroll 2
            for (int c = 0; c < 8; ++c) {
              const f3 bl = {s_rc[48 + c * 6 + 0], s_rc[48 + c * 6 + 1], s_rc[48 + c * 6 + 2]};
              const f3 bh = {s_rc[48 + c * 6 + 3], s_rc[48 + c * 6 + 4], s_rc[48 + c * 6 + 5]};
              if (overlap(bl, bh, L.lo, L.hi)) cb |= 1u << c;
            }
            cand = cross_mask(rb, cb);
          }
          // the candidate vertices whose own box overlaps the leaf, as a mask
          // of tile lanes.  A leaf with more than TQ_HEAVY candidates (a
          // stretched or fast triangle spanning the tile) is tested by the
          // whole wave below, one vertex per lane, instead of one candidate
          // after another by its own lane
          const bool heavy = popc64(cand) > TQ_HEAVY;
          uint64_t hv = __ballot(heavy);
          const uint64_t hcand = cand;
          if (heavy) cand = 0;
          uint64_t vm = 0;
          if (hv) {
            // this lane's vertex, as the loop below reads a candidate's
            const f3 vc = {s_q[lane * 6 + 0], s_q[lane * 6 + 1], s_q[lane * 6 + 2]};
            const f3 vl = {vc.x - margin, vc.y - margin, vc.z - margin};
            const f3 vh = {vc.x + margin, vc.y + margin, vc.z + margin};
            const uint32_t pv = s_p[lane];
            const bool act = base + lane < a.nitems;
            do {  // wave-uniform: leaf l's box, corners and mask read from its lane
              const uint32_t l = (uint32_t)__builtin_ctzll(hv);
              hv &= hv - 1;
              const auto rl = [l](float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
              const f3 llo = {rl(L.lo.x), rl(L.lo.y), rl(L.lo.z)}, lhi = {rl(L.hi.x), rl(L.hi.y), rl(L.hi.z)};
              const uint32_t g0 = (uint32_t)__builtin_amdgcn_readlane((int)f0, l);
              const uint32_t g1 = (uint32_t)__builtin_amdgcn_readlane((int)f1, l);
              const uint32_t g2 = (uint32_t)__builtin_amdgcn_readlane((int)f2, l);
              const uint32_t clo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)hcand, l);
              const uint32_t chi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(hcand >> 32), l);
              const uint64_t cl = ((uint64_t)chi << 32) | clo;
              const bool ok = ((cl >> lane) & 1u) != 0 && act && overlap(vl, vh, llo, lhi) &&
                              !(self && (pv == g0 || pv == g1 || pv == g2));
              const uint64_t m1 = __ballot(ok);
              if (lane == l) vm = m1;
            } while (hv);
          }
          while (cand) {
            const uint32_t v = __builtin_ctzll(cand);
            cand &= cand - 1;
            if (base + v >= a.nitems) continue;
            const f3 vc = {s_q[v * 6 + 0], s_q[v * 6 + 1], s_q[v * 6 + 2]};
            const f3 vl = {vc.x - margin, vc.y - margin, vc.z - margin};
            const f3 vh = {vc.x + margin, vc.y + margin, vc.z + margin};
            if (!overlap(vl, vh, L.lo, L.hi)) continue;
            const uint32_t pv = s_p[v];
            if (self && (pv == f0 || pv == f1 || pv == f2)) continue;
            vm |= 1ull << v;
          }
          if (have) {
            float* tp = s_tri + 9 * lane;
            tp[0] = p0.x;
            tp[1] = p0.y;
            tp[2] = p0.z;
            tp[3] = p1.x;
            tp[4] = p1.y;
            tp[5] = p1.z;
            tp[6] = p2.x;
            tp[7] = p2.y;
            tp[8] = p2.z;
            s_bfid[lane] = fid;
          }
          wave_sync();
          // emitted in rounds of at most 4 (vertex, facet) pairs per lane
          while (__ballot(vm != 0)) {
            uint32_t nc = 0, cand = 0;  // up to 4 candidate lanes, 6 bits each
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (vm) {
                cand |= (uint32_t)__builtin_ctzll(vm) << (6 * nc);
                vm &= vm - 1;
                ++nc;
              }
            // exclusive scan of the candidate counts (nc <= 4) from the ballots
            // of their three bits
            const uint64_t c1 = __ballot(nc & 1u), c2 = __ballot(nc & 2u), c4 = __ballot(nc & 4u);
            const uint32_t excl = count_below(c1) + 2 * count_below(c2) + 4 * count_below(c4);
            const uint32_t tot = popc64(c1) + 2 * popc64(c2) + 4 * popc64(c4);
            if (np + tot > TQ_PAIRS) flush();
            const uint32_t o = np + excl;
            for (uint32_t q = 0; q < nc; ++q) s_pair[o + q] = ((cand >> (6 * q)) & 63u) | (lane << 6);
            if (a.audit) {  // one reservation per round for the wave's tot records
              uint32_t r = 0;
              if (lane == 0) r = atomicAdd(a.audit_cnt, tot);
              r = __shfl(r, 0, 64) + excl;
              for (uint32_t q = 0; q < nc; ++q)
                if (r + q < a.audit_cap)
                  a.audit[r + q] = make_uint4(s_j[(cand >> (6 * q)) & 63u], t, fid, audit_tag(AUDIT_TILE, w));
            }
            np += tot;
            npair += nc;
            wave_sync();
          }
          flush();  // the next batch overwrites the triangles
        }
      };
      // ---- phase 1: breadth-first walk with the union box, up to 64 pending
      // nodes per round.  It starts from the tree's cut when the build listed
      // one (a frontier covering every leaf once, so the leaves found are the
      // same as from the root), in one round of independent loads instead of
      // the top levels' dependent rounds; else from the root.
      uint32_t sp = 0, nl = 0;
      bool overflow = false;
      const uint32_t ncut = a.cut_cnt ? a.cut_cnt[t] : 0u;
      if (ti == ti0 && npre > 0) {  // scanned above
        sp = ovf_pre ? 0u : sp_pre;
        nl = nl_pre;
        overflow = ovf_pre;
      } else if (ncut > 0 && ncut <= CUT_CAP) {
        const int32_t* CL = a.cut + (size_t)t * CUT_CAP;
        for (uint32_t c0 = 0; c0 < ncut; c0 += 64) {
          if (nl + 64 > leaf_cap) {
            leaves(nl);
            nl = 0;
          }
          // branch-free: lanes past the cut load its first entry of this
          // batch (a valid node) and drop the result
          const bool have = c0 + lane < ncut;
          const int node = CL[c0 + (have ? lane : 0u)];
          const NodeV A = load_node(T, node);
          const bool h = have & overlap(Ulo, Uhi, A.lo, A.hi);
          const bool pc = h & (A.a >= 0), lc = h & (A.a < 0);
          const uint64_t bp = __ballot(pc), bl = __ballot(lc);
          if (sp + popc64(bp) > stack_cap) {
            overflow = true;
            break;
          }
          if (pc) s_stk[sp + count_below(bp)] = make_int2(A.a, A.b);
          if (lc) s_leaf[nl + count_below(bl)] = make_uint2((uint32_t)node, (uint32_t)A.b);
          sp += popc64(bp);
          nl += popc64(bl);
          wave_sync();
        }
        if (overflow) sp = 0;
      } else {
        const NodeV r = load_node(T, 0);
        if (overlap(Ulo, Uhi, r.lo, r.hi)) {
          if (r.a < 0) {
            if (lane == 0) s_leaf[0] = make_uint2(0u, (uint32_t)r.b);
            nl = 1;
          } else {
            if (lane == 0) s_stk[0] = make_int2(r.a, r.b);
            sp = 1;
          }
        }
      }
      wave_sync();
      uint32_t rounds = 0;
      while (sp > 0) {
        if (++rounds > 2 * nT + 2) {  // a valid tree is walked in fewer rounds
          overflow = true;
          break;
        }
        const uint32_t k = sp < 64u ? sp : 64u, b0 = sp - k;
        if (nl + 2 * k > leaf_cap) {  // this round may add 2 leaves per node
          leaves(nl);
          nl = 0;
        }
        // branch-free, so the flags stay lane masks for the ballots: lanes
        // past the k pending pairs load pair b0 (a valid one) and drop it
        const bool have = lane < k;
        const int2 e = s_stk[b0 + (have ? lane : 0u)];
        wave_sync();
        sp = b0;
        const NodeV A = load_node(T, e.x), B = load_node(T, e.y);
        const bool hA = have & overlap(Ulo, Uhi, A.lo, A.hi), hB = have & overlap(Ulo, Uhi, B.lo, B.hi);
        const bool pA = hA & (A.a >= 0), pB = hB & (B.a >= 0);
        const bool lA = hA & (A.a < 0), lB = hB & (B.a < 0);
        const uint64_t bp1 = __ballot(pA | pB), bp2 = __ballot(pA & pB);
        const uint64_t bl1 = __ballot(lA | lB), bl2 = __ballot(lA & lB);
        const uint32_t tp = popc64(bp1) + popc64(bp2), tl = popc64(bl1) + popc64(bl2);
        if (sp + tp > stack_cap) {
          overflow = true;
          break;
        }
        uint32_t op = sp + count_below(bp1) + count_below(bp2);
        uint32_t ol = nl + count_below(bl1) + count_below(bl2);
        if (pA) s_stk[op++] = make_int2(A.a, A.b);
        if (pB) s_stk[op++] = make_int2(B.a, B.b);
        if (lA) s_leaf[ol++] = make_uint2((uint32_t)e.x, (uint32_t)A.b);
        if (lB) s_leaf[ol++] = make_uint2((uint32_t)e.y, (uint32_t)B.b);
        sp += tp;
        nl += tl;
        wave_sync();
      }
      if (overflow) {
        // too wide: this target again per lane, from the slot counts it started with
        if (a.audit && lane == 0) {  // the walk's audit records of t are void
          const uint32_t r = atomicAdd(a.audit_cnt, 1u);
          if (r < a.audit_cap) a.audit[r] = make_uint4(0u, t, 0u, audit_tag(AUDIT_RETRACT, w));
        }
        wave_sync();
        uint32_t c = cnt0;
        npair = npair0;
        const f3 qc = {s_q[lane * 6 + 0], s_q[lane * 6 + 1], s_q[lane * 6 + 2]};
        const f3 qp = {s_q[lane * 6 + 3], s_q[lane * 6 + 4], s_q[lane * 6 + 5]};
        const f3 qlo = {qc.x - margin, qc.y - margin, qc.z - margin}, qhi = {qc.x + margin, qc.y + margin, qc.z + margin};
        if (active && !lane_target<false, TQ_SDEPTH>(a, m, s_p[lane], s_j[lane], t, qc, qp, qlo, qhi, band, fstk, 64, c,
                                                  npair))
          atomicOr(a.flags, 4u);
        wave_sync();
        s_cnt[lane] = c;
        wave_sync();
        continue;
      }
      leaves(nl);
    }
    wave_sync();
    cc = s_cnt[lane];
    jc = j;
    if (active) {
      a.ccount[j] = cc;
      if (cc > a.cap) atomicOr(a.flags, 2u);
    }
  }
  if (CLOTH)  // every lane (wave reductions); s_buf is free again
    tile_cloth(ca, active, jc, cc, w, one ? a.seg0_grid_n : 0u, a.seg0_mesh, a.seg0_voff,
               reinterpret_cast<float4*>(s_buf));
  wave_stats(a, w, npair);
}

uint32_t query_tiles(uint32_t nitems) { return ceil_div(nitems, 64); }

int launch_query_cloth(hipStream_t s, const QueryArgs& a, const ClothArgs& ca) {
  if (!a.nitems) return YSIM_OK;
  if (!a.tile || ca.nitems != a.nitems) return YSIM_ERR_INVALID;
  hipLaunchKernelGGL(k_query_tile<true>, dim3(query_tiles(a.nitems)), dim3(64), 0, s, a, ca);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

bool query_block_ok(uint32_t b) { return b == 64 || b == 128 || b == 256; }
bool query_waves_ok(uint32_t w) { return w == 5 || w == 6 || w == 8; }

int launch_query(hipStream_t s, const QueryArgs& a, bool emit_pairs) {
  if (!a.nitems) return YSIM_OK;
  if (!emit_pairs && a.tile) {
    hipLaunchKernelGGL(k_query_tile<false>, dim3(ceil_div(a.nitems, 64)), dim3(64), 0, s, a, ClothArgs{});
    YSIM_CHECK_LAUNCH();
    return YSIM_OK;
  }
  const uint32_t qb = a.block;
  if (!query_block_ok(qb) || !query_waves_ok(a.waves)) return YSIM_ERR_INVALID;
#define YSIM_Q(E, B, W) hipLaunchKernelGGL((k_query<E, B, W>), dim3(ceil_div(a.nitems, B)), dim3(B), 0, s, a)
#define YSIM_QB(E, W) \
  if (qb == 64) YSIM_Q(E, 64, W); else if (qb == 128) YSIM_Q(E, 128, W); else YSIM_Q(E, 256, W)
  if (emit_pairs) {
    YSIM_QB(true, 5);
  } else if (a.waves == 5) {
    YSIM_QB(false, 5);
  } else if (a.waves == 6) {
    YSIM_QB(false, 6);
  } else {
    YSIM_QB(false, 8);
  }
#undef YSIM_QB
#undef YSIM_Q
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

}  // namespace ysim
