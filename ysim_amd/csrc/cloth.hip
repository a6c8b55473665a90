// cloth.hip — FastGrid or TriangularCloth springs + symplectic integrate +
// contact projection, fused into one pass per vertex.  integrate_cloth
// (physics.metal:380-466, TriangularCloth) applies the same update as
// integrate_cloth_grid.
//
// compute_cloth_grid_forces_fast (src/metal/physics.metal:171-224) gathers 12
// grid springs; integrate_cloth_grid (physics.metal:226-312) then updates v,
// applies the contact bucket (dedup by target facet, gate distance < thickness,
// drop the inward normal velocity, push out to thickness), updates x and runs
// the world-bounds guard.  Forces read (x, v) and the results go to separate
// buffers (x_out, v_out), so the two reference passes fuse without a race and
// the old x buffer becomes the next substep's xPrev (snapshot_xprev_range,
// physics.metal:520-529) without a copy.
#include "kernels.hpp"

namespace ysim {

// One vertex j < a.nitems: forces, integrate, contacts, guard; stores v_out
// and x_out and returns the new position.
__device__ __forceinline__ f3 cloth_vertex(const ClothArgs& a, uint32_t j) {
  const uint32_t sg = seg_of(a.dsegs, j);
  const uint32_t m = a.dsegs.mesh[sg];
  const uint32_t i = j - a.dsegs.off[sg];
  const uint32_t vb = a.mt.voff[m];
  const uint32_t g = vb + i;
  const ysim_fastgrid_params P = a.cloth[m];
  const f3 pos0 = ld3(a.x, g), vel0 = ld3(a.v, g);
  const float* KS = (a.kspring && sg == 0) ? a.kspring : nullptr;
  const f3 force = (a.behavior && a.behavior[m] == YSIM_BEHAVIOR_TRIANGULAR_CLOTH)
                       ? tri_force(P, g, a.tri_off, a.tri_mid, a.tri_ent, a.x, a.v, pos0, vel0, ld3(a.ext, g), a.kair,
                                   a.kd)
                       : grid_force(P, i, vb, a.x, a.v, pos0, vel0, ld3(a.ext, g), a.kair, a.kd, KS);
  if (a.f_out) st3(a.f_out, g, force);

  // integrate_cloth_grid
  const float mask = a.fixed[g];
  const float mass = a.mass[g];
  const float subh = a.subh;
  f3 pos = pos0, vel = vel0;
  vel.x = vel.x + ((subh * force.x) / mass) * mask;
  vel.y = vel.y + ((subh * force.y) / mass) * mask;
  vel.z = vel.z + ((subh * force.z) / mass) * mask;

  const float thickness = P.thickness;
  auto apply = [&](float4 nd) {
    f3 n = {nd.x, nd.y, nd.z};
    float nlen2 = dot(n, n);
    if (nlen2 < 1e-12f) return;
    float inv = 1.0f / sqrtf(nlen2);
    n = n * inv;
    float vn = dot(vel, n);
    float dist = nd.w;
    if (dist < thickness) {
      if (vn < 0.0f) vel = vel - vn * n;
      pos = pos + (thickness - dist) * n;
    }
  };
  if (a.csr_off) {
    // bucket already ordered; skip repeats of a (target, facet) (physics.metal:257-271)
    const uint32_t b = a.csr_off[g], e = a.csr_off[g + 1];
    for (uint32_t c = b; c < e; ++c) {
      bool dup = false;
      for (uint32_t d = b; d < c; ++d)
        if (a.csr[d].target_obj == a.csr[c].target_obj && a.csr[d].triangle == a.csr[c].triangle) {
          dup = true;
          break;
        }
      if (dup) continue;
      const float* q = a.csr[c].normal_distance;
      apply(make_float4(q[0], q[1], q[2], q[3]));
    }
  } else {
    uint32_t cnt = a.ccount ? a.ccount[j] : 0u;
    if (cnt > a.cap) cnt = a.cap;
    const float4* ND = a.cnd + (size_t)j * a.cap;
    const uint2* ID = a.cid + (size_t)j * a.cap;
    // visit contacts in ascending (target mesh, facet) order; equal keys are
    // duplicates and are applied once (physics.metal:257-271)
    uint64_t last = 0;
    bool first = true;
    for (uint32_t it = 0; it < cnt; ++it) {
      uint64_t best = ~0ull;
      uint32_t bi = 0;
      for (uint32_t c = 0; c < cnt; ++c) {
        uint2 id = ID[c];
        uint64_t key = ((uint64_t)id.x << 32) | id.y;
        if ((first || key > last) && key < best) {
          best = key;
          bi = c;
        }
      }
      if (best == ~0ull) break;
      first = false;
      last = best;
      apply(ND[bi]);
    }
  }
  pos = pos + (subh * vel) * mask;
  if (sanitize(pos, vel, pos0)) atomicOr(a.flags, 1u);
  st3(a.v_out, g, vel);
  st3(a.x_out, g, pos);
  return pos;
}

// BS threads per block (128, 256 or 512; chosen by the autotuner).  With
// box_parts set, each block also writes the box of its new positions (6
// ordered keys at box_parts + 6 blockIdx.x); the next build's Morton kernel
// reduces them, so the build needs no box kernel of its own.
template <int BS>
__global__ __launch_bounds__(BS) void k_cloth(ClothArgs a) {
  const uint32_t j = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (!a.box_parts) {
    if (j < a.nitems) cloth_vertex(a, j);
    return;
  }
  const float INF = __uint_as_float(0x7f800000u);
  f3 lo = {INF, INF, INF}, hi = {-INF, -INF, -INF};
  if (j < a.nitems) lo = hi = cloth_vertex(a, j);
  block_box_store<BS>(lo, hi, a.box_parts + 6 * (size_t)blockIdx.x);
}

bool cloth_block_ok(uint32_t b) { return b == 128 || b == 256 || b == 512; }
uint32_t cloth_box_parts_words(uint32_t nitems) { return 6 * ceil_div(nitems, 128); }
uint32_t cloth_box_parts(uint32_t nitems, uint32_t block) { return ceil_div(nitems, block); }

int launch_cloth(hipStream_t s, const ClothArgs& a) {
  if (!a.nitems) return YSIM_OK;
  if (!cloth_block_ok(a.block)) return YSIM_ERR_INVALID;
#define YSIM_C(B) hipLaunchKernelGGL((k_cloth<B>), dim3(ceil_div(a.nitems, B)), dim3(B), 0, s, a)
  if (a.block == 128) YSIM_C(128); else if (a.block == 256) YSIM_C(256); else YSIM_C(512);
#undef YSIM_C
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

}  // namespace ysim
