// dropin.hip — the unfused stage kernels behind the drop-in entry points:
// broad pairs -> narrow contacts -> per-vertex contact CSR, exactly the data
// flow of the reference's detectCollisions / narrowAndSortByVertices
// (include/sim/bvh.hpp:2853-2949, include/sim/bruteforce.hpp:311-384).  The
// scene substep fuses these stages (query.hip); these kernels exist for
// callers that drive the stages one by one and for stage-level parity tests.
#include "kernels.hpp"

namespace ysim {

namespace {
constexpr int QB = 128;

struct NodeV {
  f3 lo, hi;
  int a, b;
};
__device__ __forceinline__ NodeV load_node(const ysim_bvh_node* T, int i) {
  const float4* q = reinterpret_cast<const float4*>(T + i);
  float4 u = q[0], w = q[1];
  return NodeV{{u.x, u.y, u.z}, {w.x, w.y, w.z}, __float_as_int(u.w), __float_as_int(w.w)};
}
}  // namespace

// queryPoints (src/metal/bvh.metal:1393-1412) with an explicit per-thread
// stack in scratch-free form: plain DFS over a private stack of depth 64
// (the reference's depth) backed by the parent-link walk when it fills.
// SEG: queryPointsSegmented + scanReserveSegmented + compactSegmented
// (bvh.metal:1732-1827) in one kernel: a workgroup stages its pairs in LDS,
// reserves its output range with one atomic and copies them out coalesced;
// pairs beyond the LDS capacity reserve per wave (one atomic for the wave's
// spilling lanes).  Output order is per workgroup instead of per pair; the
// set is the same.
constexpr uint32_t SEG_CAP = 512;
template <bool SEG>
__global__ __launch_bounds__(QB) void k_broad_pairs(const float* __restrict__ xq, uint32_t nq, uint32_t qobj,
                                                    uint32_t qbeh, uint32_t qshape,
                                                    const ysim_bvh_node* __restrict__ T,
                                                    const int32_t* __restrict__ P, uint32_t nT,
                                                    const uint32_t* __restrict__ tf, uint32_t tobj,
                                                    uint32_t tbeh, uint32_t tshape, float margin,
                                                    ysim_broad_collision* __restrict__ out, uint32_t cap,
                                                    uint32_t* __restrict__ count, uint32_t* __restrict__ flags) {
  __shared__ ysim_broad_collision s_rec[SEG ? SEG_CAP : 1];
  __shared__ uint32_t s_cnt, s_base;
  if (SEG) {
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
  }
  const uint32_t p = blockIdx.x * QB + threadIdx.x;
  if (p < nq && nT != 0) {
    const f3 x = ld3(xq, p);
    const f3 qmin = {x.x - margin, x.y - margin, x.z - margin};
    const f3 qmax = {x.x + margin, x.y + margin, x.z + margin};
    const bool self = (qobj == tobj);
    auto emit = [&](int fid) {
      const uint32_t* f = tf + 3 * (size_t)fid;
      if (self && (p == f[0] || p == f[1] || p == f[2])) return;
      ysim_broad_collision r;
      r.point = p; r.triangle = (uint32_t)fid; r.query_obj = qobj; r.target_obj = tobj;
      r.query_behavior = qbeh; r.target_behavior = tbeh; r.query_shape = qshape; r.target_shape = tshape;
      if (SEG) {
        // LDS staging; past its capacity the wave's spilling lanes share one
        // output reservation (ballot + prefix rank)
        const uint32_t k = atomicAdd(&s_cnt, 1u);
        const bool spill = k >= SEG_CAP;
        if (!spill) s_rec[k] = r;
        const uint64_t m = __ballot(spill);
        if (m) {
          const int leader = __ffsll((unsigned long long)m) - 1;
          const int lane = (int)__lane_id();
          uint32_t base = 0;
          if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
          base = __shfl(base, leader);
          if (spill) {
            const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (slot < cap) out[slot] = r;
          }
        }
        return;
      }
      const uint32_t slot = atomicAdd(count, 1u);
      if (slot < cap) out[slot] = r;
    };
    // stackless DFS through parent links (node 0 = root)
    int c = 0;
    uint32_t guard = 0;
    const uint32_t guard_max = 6 * nT + 16;
    bool down = true;
    while (true) {
      if (++guard > guard_max) {
        atomicOr(flags, 4u);
        break;
      }
      if (down) {
        NodeV N = load_node(T, c);
        if (overlap(qmin, qmax, N.lo, N.hi)) {
          if (N.a < 0) {
            emit(N.b);
          } else {
            c = N.a;
            continue;
          }
        }
      }
      // climb to the next unvisited right sibling
      if (c == 0) break;
      int pp = P[c];
      if (pp < 0 || pp >= (int)nT - 1) {
        atomicOr(flags, 4u);
        break;
      }
      NodeV PN = load_node(T, pp);
      if (PN.a == c) {
        c = PN.b;
        down = true;
      } else {
        c = pp;
        down = false;
      }
    }
  }
  if (SEG) {
    __syncthreads();
    const uint32_t nl = s_cnt < SEG_CAP ? s_cnt : SEG_CAP;
    if (threadIdx.x == 0) s_base = nl ? atomicAdd(count, nl) : 0u;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nl; k += QB) {
      const uint32_t slot = s_base + k;
      if (slot < cap) out[slot] = s_rec[k];
    }
  }
}

int launch_broad_pairs(hipStream_t s, const float* xq, uint32_t nq, uint32_t qobj, uint32_t qbeh,
                       uint32_t qshape, const ysim_bvh_node* tree, const int32_t* parent, uint32_t nT,
                       const uint32_t* tfacets, uint32_t tobj, uint32_t tbeh, uint32_t tshape, float margin,
                       ysim_broad_collision* out, uint32_t cap, uint32_t* count, uint32_t* flags, bool segmented) {
  if (!nq || !nT) return YSIM_OK;
  if (segmented)
    hipLaunchKernelGGL(k_broad_pairs<true>, dim3(ceil_div(nq, QB)), dim3(QB), 0, s, xq, nq, qobj, qbeh, qshape, tree,
                       parent, nT, tfacets, tobj, tbeh, tshape, margin, out, cap, count, flags);
  else
    hipLaunchKernelGGL(k_broad_pairs<false>, dim3(ceil_div(nq, QB)), dim3(QB), 0, s, xq, nq, qobj, qbeh, qshape,
                       tree, parent, nT, tfacets, tobj, tbeh, tshape, margin, out, cap, count, flags);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// narrow_pt_tri, src/metal/bruteforce.metal:41-178 (one thread per pair).
__global__ __launch_bounds__(256) void k_narrow_pairs(const ysim_broad_collision* __restrict__ pairs, uint32_t np,
                                                      const float* __restrict__ x, const float* __restrict__ xprev,
                                                      const uint32_t* __restrict__ soff,
                                                      const uint32_t* __restrict__ facets,
                                                      const uint32_t* __restrict__ foff,
                                                      const uint32_t* __restrict__ adj_off,
                                                      const uint32_t* __restrict__ adj, float band,
                                                      ysim_narrow_collision* __restrict__ out, uint32_t cap,
                                                      uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  const ysim_broad_collision bc = pairs[i];
  const uint32_t* tri = facets + 3 * (size_t)(bc.triangle + foff[bc.target_obj]);
  if (bc.query_obj == bc.target_obj && adj_off) {
    const uint32_t gp = bc.point + soff[bc.target_obj];
    for (uint32_t k = adj_off[gp]; k < adj_off[gp + 1]; ++k)
      if (adj[k] == bc.triangle) return;
  }
  const uint32_t qo = soff[bc.query_obj], to = soff[bc.target_obj];
  f3 n;
  float d;
  if (!narrow_test(ld3(x, bc.point + qo), ld3(xprev, bc.point + qo), ld3(x, tri[0] + to), ld3(x, tri[1] + to),
                   ld3(x, tri[2] + to), band, n, d))
    return;
  uint32_t slot = atomicAdd(count, 1u);
  if (slot >= cap) return;
  ysim_narrow_collision r;
  r.point = bc.point; r.triangle = bc.triangle; r.query_obj = bc.query_obj; r.target_obj = bc.target_obj;
  r.normal_distance[0] = n.x; r.normal_distance[1] = n.y; r.normal_distance[2] = n.z; r.normal_distance[3] = d;
  r.query_behavior = bc.query_behavior; r.target_behavior = bc.target_behavior;
  r.query_shape = bc.query_shape; r.target_shape = bc.target_shape;
  out[slot] = r;
}

int launch_narrow_pairs(hipStream_t s, const ysim_broad_collision* pairs, uint32_t np, const float* x,
                        const float* xprev, const uint32_t* states_off, const uint32_t* facets,
                        const uint32_t* facets_off, const uint32_t* adj_off, const uint32_t* adj, float band,
                        ysim_narrow_collision* out, uint32_t cap, uint32_t* count) {
  if (!np) return YSIM_OK;
  hipLaunchKernelGGL(k_narrow_pairs, dim3(ceil_div(np, 256)), dim3(256), 0, s, pairs, np, x, xprev, states_off,
                     facets, facets_off, adj_off, adj, band, out, cap, count);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// ------------------------------------------------------------------ contact CSR
// narrow_bucket_zero/count/scan/scatter (src/metal/physics.metal:546-616) with a
// parallel scan and a deterministic in-bucket order.
__global__ void k_csr_count(const ysim_narrow_collision* __restrict__ c, uint32_t nc,
                            const uint32_t* __restrict__ soff, uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  atomicAdd(&cnt[soff[c[i].query_obj] + c[i].point], 1u);
}

// exclusive scan of n values in place into out[0..n] (out[n] = total), using
// per-block sums in bsum (<= 1024 blocks of 1024 items)
__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t n,
                                                     uint32_t* __restrict__ out, uint32_t* __restrict__ bsum) {
  __shared__ uint32_t s[256];
  const uint32_t base = blockIdx.x * 1024;
  uint32_t v[4], tot = 0;
  for (int k = 0; k < 4; ++k) {
    uint32_t i = base + threadIdx.x * 4 + k;
    v[k] = i < n ? in[i] : 0;
    tot += v[k];
  }
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    uint32_t t = threadIdx.x >= (uint32_t)o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tot;
  for (int k = 0; k < 4; ++k) {
    uint32_t i = base + threadIdx.x * 4 + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ bsum, uint32_t nb) {
  __shared__ uint32_t s[1024];
  uint32_t v = threadIdx.x < nb ? bsum[threadIdx.x] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t t = threadIdx.x >= (uint32_t)o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (threadIdx.x < nb) bsum[threadIdx.x] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) bsum[1024] = s[1023];
}
__global__ void k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ bsum) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] += bsum[i / 1024];
  if (i == 0) out[n] = bsum[1024];
}
__global__ void k_csr_scatter(const ysim_narrow_collision* __restrict__ c, uint32_t nc,
                              const uint32_t* __restrict__ soff, const uint32_t* __restrict__ off,
                              uint32_t* __restrict__ fill, ysim_narrow_collision* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  const uint32_t v = soff[c[i].query_obj] + c[i].point;
  out[off[v] + atomicAdd(&fill[v], 1u)] = c[i];
}
// insertion sort of each bucket by (target_obj, triangle)
__global__ void k_csr_sort(const uint32_t* __restrict__ off, uint32_t nv, ysim_narrow_collision* __restrict__ r) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const uint32_t b = off[v], e = off[v + 1];
  for (uint32_t i = b + 1; i < e; ++i) {
    ysim_narrow_collision k = r[i];
    uint64_t key = ((uint64_t)k.target_obj << 32) | k.triangle;
    uint32_t j = i;
    while (j > b) {
      uint64_t pk = ((uint64_t)r[j - 1].target_obj << 32) | r[j - 1].triangle;
      if (pk <= key) break;
      r[j] = r[j - 1];
      --j;
    }
    r[j] = k;
  }
}

int launch_contact_csr(hipStream_t s, const ysim_narrow_collision* c, uint32_t nc, const uint32_t* states_off,
                       uint32_t nv, uint32_t* offsets, ysim_narrow_collision* sorted, uint32_t* scratch) {
  if (nv > 1024u * 1024u) return fail(YSIM_ERR_UNSUPPORTED, "contact_csr: more than 2^20 vertices");
  uint32_t* cnt = scratch;            // nv
  uint32_t* bsum = scratch + nv;      // 1025
  const uint32_t zb = std::max<uint32_t>(1, ceil_div(nv, 256));
  hipLaunchKernelGGL(k_zero_u32, dim3(zb), dim3(256), 0, s, cnt, nv);
  YSIM_CHECK_LAUNCH();
  if (nc) {
    hipLaunchKernelGGL(k_csr_count, dim3(ceil_div(nc, 256)), dim3(256), 0, s, c, nc, states_off, cnt);
    YSIM_CHECK_LAUNCH();
  }
  const uint32_t nb = std::max<uint32_t>(1, ceil_div(nv, 1024));
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, s, cnt, nv, offsets, bsum);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, bsum, nb);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_add, dim3(zb), dim3(256), 0, s, offsets, nv, bsum);
  YSIM_CHECK_LAUNCH();
  if (nc) {
    hipLaunchKernelGGL(k_zero_u32, dim3(zb), dim3(256), 0, s, cnt, nv);
    YSIM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_csr_scatter, dim3(ceil_div(nc, 256)), dim3(256), 0, s, c, nc, states_off, offsets, cnt,
                       sorted);
    YSIM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_csr_sort, dim3(zb), dim3(256), 0, s, offsets, nv, sorted);
    YSIM_CHECK_LAUNCH();
  }
  return YSIM_OK;
}

}  // namespace ysim
