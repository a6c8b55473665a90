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
// records of either producer: cloth blocks of >= 128 vertices or query tiles of 64
uint32_t cloth_box_parts_words(uint32_t nitems) { return 6 * ceil_div(nitems, 64); }
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
