// Host build of the fused query's traversal (ysim_amd/csrc/traverse.hpp): runs
// the exact kernel logic on the CPU so tests can check, without a GPU, that
// the short-stack + stackless walk visits every overlapping leaf once at full
// scale (stack overflows included) and count its iterations.
#include <cstdint>
#include <vector>

#include "../../ysim_amd/csrc/audit.hpp"
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

// The library's audit filter (ysim_amd/csrc/audit.hpp) on host records:
// returns the kept count, writes up to cap (j, t, fid) triples.
extern "C" uint64_t yt_audit_collect(const uint32_t* rec, uint64_t n, uint32_t* out3, uint64_t cap,
                                     uint32_t* n_retract, uint32_t* n_bad) {
  const std::vector<AuditTriple> kept = audit_collect(rec, n, n_retract, n_bad);
  for (uint64_t i = 0; i < kept.size() && i < cap; ++i) {
    out3[3 * i] = kept[i].j;
    out3[3 * i + 1] = kept[i].t;
    out3[3 * i + 2] = kept[i].fid;
  }
  return kept.size();
}
