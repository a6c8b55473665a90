// Host build of the fused query's traversal (ysim_amd/csrc/traverse.hpp): runs
// the exact kernel logic on the CPU so tests can check, without a GPU, that
// the short-stack + stackless walk visits every overlapping leaf once at full
// scale (stack overflows included) and count its iterations.
#include <cstdint>
#include <vector>

#include "../../ysim_amd/csrc/traverse.hpp"

using namespace ysim;

extern "C" void yt_tile_order(uint32_t N, uint32_t* out) {
  for (uint32_t i = 0; i < N * N; ++i) out[i] = tile_order(i, N);
}

extern "C" int yt_query(const float* x, uint32_t nq, uint32_t qobj, const ysim_bvh_node* tree, const int32_t* parent,
                        uint32_t nT, const uint32_t* tf, uint32_t tobj, float margin, int sdepth_small,
                        uint32_t* out_pairs, uint64_t cap, uint64_t* out_count, uint64_t* out_iters,
                        uint32_t* out_max_iters) {
  uint64_t count = 0, iters = 0;
  uint32_t maxit = 0;
  std::vector<int2> stack(64);
  for (uint32_t p = 0; p < nq; ++p) {
    f3 q = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
    f3 qmin = {q.x - margin, q.y - margin, q.z - margin}, qmax = {q.x + margin, q.y + margin, q.z + margin};
    auto leaf = [&](int fid) {
      const uint32_t* f = tf + 3 * (size_t)fid;
      if (qobj == tobj && (p == f[0] || p == f[1] || p == f[2])) return;
      if (out_pairs && count < cap) {
        out_pairs[2 * count] = p;
        out_pairs[2 * count + 1] = (uint32_t)fid;
      }
      ++count;
    };
    int it = sdepth_small == 1 ? traverse<2>(tree, parent, nT, qmin, qmax, stack.data(), 1, leaf)
             : sdepth_small == 8 ? traverse<8>(tree, parent, nT, qmin, qmax, stack.data(), 1, leaf)
             : sdepth_small == 12 ? traverse<12>(tree, parent, nT, qmin, qmax, stack.data(), 1, leaf)
                                 : traverse<16>(tree, parent, nT, qmin, qmax, stack.data(), 1, leaf);
    if (it < 0) return -1;
    iters += (uint64_t)it;
    if ((uint32_t)it > maxit) maxit = (uint32_t)it;
  }
  *out_count = count;
  *out_iters = iters;
  *out_max_iters = maxit;
  return 0;
}
