// query.hip — BVH traversal fused with the point-triangle narrow phase.
//
// One thread per query vertex (queryPoints, src/metal/bvh.metal:1393-1412).  For
// every target tree of its collision group the thread runs a short-stack DFS
// that tests both children at once and keeps only "both children hit" siblings
// on a per-thread LDS stack; if that stack fills, the walk continues
// stacklessly through parent links (exactly the same visit set).  Every
// candidate (leaf box overlapping [x - m, x + m], closed; facets containing the
// vertex skipped on self queries, bvh.metal:1366-1371) is either emitted as a
// pair record (broad-phase mode) or immediately tested with the swept
// point-triangle narrow phase (narrow_pt_tri, src/metal/bruteforce.metal:41-178),
// and contacts are kept in per-vertex slots.  Contacts are applied later in
// (target mesh, facet) order, so the result does not depend on traversal order.
#include "kernels.hpp"
#include "traverse.hpp"

namespace ysim {

namespace {
constexpr int SDEPTH = 8;    // LDS stack entries per thread (deeper stacks overflow no less on cloth)
}  // namespace

// QB threads per block (64, 128 or 256) and an occupancy target of WPE waves
// per SIMD (5: registers as needed; 6-8: fewer registers, some spilled), both
// chosen by the autotuner: the traversal is bound by dependent node loads, so
// more waves in flight can pay for a few spills.
template <bool EMIT_PAIRS, int QB, int WPE>
__attribute__((amdgpu_waves_per_eu(WPE, WPE))) __global__ __launch_bounds__(QB) void k_query(QueryArgs a) {
  __shared__ int2 s_stack[SDEPTH * QB];
  const uint32_t tq = xcd_block(blockIdx.x, gridDim.x) * QB + threadIdx.x;
  uint32_t npair = 0;
  if (tq < a.nitems) {
    const uint32_t sg = seg_of(a.qsegs, tq);
    const uint32_t m = a.qsegs.mesh[sg];
    // grid cloths are queried in 8 x 8 vertex tiles (one tile per wave); the
    // contact slots stay indexed by vertex
    const uint32_t gn = a.mesh_grid_n ? a.mesh_grid_n[m] : 0u;
    const uint32_t p = gn ? tile_order(tq - a.qsegs.off[sg], gn) : tq - a.qsegs.off[sg];
    const uint32_t j = a.qsegs.off[sg] + p;
    const uint32_t g = a.mt.voff[m] + p;
    const f3 qcur = ld3(a.x, g);
    const f3 qprev = ld3(a.xprev, g);
    const float margin = a.mesh_margin[m];
    const float band = a.mesh_band[m];
    const f3 qmin = {qcur.x - margin, qcur.y - margin, qcur.z - margin};
    const f3 qmax = {qcur.x + margin, qcur.y + margin, qcur.z + margin};
    uint32_t cnt = 0;
    const uint32_t cap = a.cap;
    for (uint32_t ti = a.tgt_off[sg]; ti < a.tgt_off[sg + 1]; ++ti) {
      const uint32_t t = a.tgt[ti];
      const uint32_t nT = a.mesh_nfacets[t];
      if (nT == 0) continue;
      const ysim_bvh_node* T = a.tree + a.mt.noff[t];
      const int32_t* P = a.parent + a.mt.noff[t];
      const uint32_t* F = a.facets + 3 * (size_t)a.mt.foff[t];
      const uint32_t tvb = a.mt.voff[t];
      const bool self = (t == m);
      const int it = traverse<SDEPTH>(T, P, nT, qmin, qmax, s_stack + threadIdx.x, QB,
               [&](int fid) {
                 const uint32_t* f = F + 3 * (size_t)fid;
                 const uint32_t f0 = f[0], f1 = f[1], f2 = f[2];
                 if (self && (p == f0 || p == f1 || p == f2)) return;
                 ++npair;
                 if (EMIT_PAIRS) {
                   uint32_t slot = atomicAdd(a.npairs, 1u);
                   if (slot < a.pair_cap) {
                     uint4* o = reinterpret_cast<uint4*>(a.pairs) + slot;
                     *o = make_uint4(m, p, t, (uint32_t)fid);
                   }
                   return;
                 }
                 f3 nrm;
                 float d;
                 if (narrow_test(qcur, qprev, ld3(a.x, tvb + f0), ld3(a.x, tvb + f1), ld3(a.x, tvb + f2),
                                 band, nrm, d)) {
                   if (cnt < cap) {
                     a.cnd[(size_t)j * cap + cnt] = make_float4(nrm.x, nrm.y, nrm.z, d);
                     a.cid[(size_t)j * cap + cnt] = make_uint2(t, (uint32_t)fid);
                   }
                   ++cnt;
                 }
               });
      if (it < 0) atomicOr(a.flags, 4u);
    }
    if (!EMIT_PAIRS) {
      a.ccount[j] = cnt;
      if (cnt > cap) atomicOr(a.flags, 2u);
    }
  }
  if (!EMIT_PAIRS && a.npairs) {
    // candidate-pair statistics: one atomic per wave
    uint32_t s = npair;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(a.npairs, s);
  }
}

bool query_block_ok(uint32_t b) { return b == 64 || b == 128 || b == 256; }
bool query_waves_ok(uint32_t w) { return w == 5 || w == 6 || w == 8; }

int launch_query(hipStream_t s, const QueryArgs& a, bool emit_pairs) {
  if (!a.nitems) return YSIM_OK;
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
