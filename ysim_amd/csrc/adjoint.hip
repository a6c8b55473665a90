// adjoint.hip — reverse-mode derivative of the cloth substep (springs +
// symplectic integrate + contact projection) for one FastGrid cloth whose
// contacts are against static meshes.  The reference has no differentiation;
// the map differentiated is its substep (integrate_cloth_grid,
// src/metal/physics.metal:226-312, over the gather of physics.metal:171-224)
// with the contact set and the two gates (distance < thickness, v.n < 0) of the
// forward step held fixed.
//
// One adjoint step is two vertex-parallel passes, both HBM-bound gathers:
//   k_adj_integrate: recompute the vertex's force and replay its contact
//     sequence (same order and arithmetic as k_cloth) to recover the gates, then
//     pull (x_bar', v_bar') back through x' = p + (h u) mask, the contact pushes
//     and projections (in reverse order) and u = v + (h f / m) mask:
//       u_bar = v_bar' + (h mask) x_bar';  x_bar = x_bar' - sum n (n.x_bar');
//       u_bar <- (I - n n^T) u_bar for gated contacts;  f_bar = u_bar (h mask)/m;
//       v_bar = u_bar + kair f_bar.
//   k_adj_springs: for each of the 12 springs, g = f_bar_i - f_bar_j and
//       x_bar_i -= (k n + kd M dv)(n.g) + s M g,  v_bar_i -= kd (n.g) n,
//     with M = (I - n n^T)/len; the lower-index end adds (len - r)(n.g) to the
//     spring's stiffness accumulator and (dv.n)(n.g) to its damping sum.  Every
//     accumulator has exactly one writer, so the result is deterministic.
#include "kernels.hpp"

namespace ysim {

namespace {
constexpr int AB = 256;

// Contact slots of a vertex visited in ascending (descending) unique
// (target mesh, facet) key order — the order k_cloth applies them in.
struct SlotOrder {
  const uint2* id;
  uint32_t cnt;
  __device__ uint64_t key(uint32_t c) const { return ((uint64_t)id[c].x << 32) | id[c].y; }
  // smallest key > last (any key when first); returns false when none
  __device__ bool next_up(bool first, uint64_t last, uint32_t& bi, uint64_t& best) const {
    best = ~0ull;
    bool found = false;
    for (uint32_t c = 0; c < cnt; ++c) {
      const uint64_t k = key(c);
      if ((first || k > last) && k < best) {
        best = k;
        bi = c;
        found = true;
      }
    }
    return found;
  }
  // largest key < last (any key when first)
  __device__ bool next_down(bool first, uint64_t last, uint32_t& bi, uint64_t& best) const {
    best = 0;
    bool found = false;
    for (uint32_t c = 0; c < cnt; ++c) {
      const uint64_t k = key(c);
      if ((first || k < last) && (!found || k > best)) {
        best = k;
        bi = c;
        found = true;
      }
    }
    return found;
  }
};
}  // namespace

__global__ __launch_bounds__(AB) void k_adj_exp(const float* theta, float* k, uint32_t n) {
  const uint32_t s = blockIdx.x * AB + threadIdx.x;
  if (s < n) k[s] = (float)exp((double)theta[s]);
}

__global__ __launch_bounds__(AB) void k_adj_integrate(AdjArgs a) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * AB + threadIdx.x;
  if (i >= a.V) return;
  const uint32_t g = a.voff + i;
  const f3 pos0 = ld3(a.x, g), vel0 = ld3(a.v, g);
  const f3 force = grid_force(a.P, i, a.voff, a.x, a.v, pos0, vel0, ld3(a.ext, g), a.kair, a.kd, a.ks);
  const float mask = a.fixed[g];
  const float mass = a.mass[g];
  const float subh = a.subh;
  const float thickness = a.P.thickness;
  f3 u = vel0;
  u.x = u.x + ((subh * force.x) / mass) * mask;
  u.y = u.y + ((subh * force.y) / mass) * mask;
  u.z = u.z + ((subh * force.z) / mass) * mask;

  uint32_t cnt = a.ccount[i];
  if (cnt > a.cap) cnt = a.cap;
  const float4* ND = a.cnd + (size_t)i * a.cap;
  const SlotOrder ord{a.cid + (size_t)i * a.cap, cnt};
  auto unit = [](float4 nd, f3& n) {
    n = f3{nd.x, nd.y, nd.z};
    const float nlen2 = dot(n, n);
    if (nlen2 < 1e-12f) return false;
    n = n * (1.0f / sqrtf(nlen2));
    return true;
  };
  // forward replay: gate bits by rank in application order
  uint64_t act = 0, proj = 0;
  uint32_t nsel = 0;
  {
    uint64_t last = 0, best;
    uint32_t bi = 0;
    for (bool first = true; ord.next_up(first, last, bi, best); first = false, last = best, ++nsel) {
      const float4 nd = ND[bi];
      f3 n;
      if (!unit(nd, n)) continue;
      const float vn = dot(u, n);
      if (nd.w < thickness) {
        act |= 1ull << nsel;
        if (vn < 0.0f) {
          u = u - vn * n;
          proj |= 1ull << nsel;
        }
      }
    }
  }
  // reverse
  const f3 pb = ld3(a.xb, i);
  f3 ub = ld3(a.vb, i) + (subh * mask) * pb;
  f3 xbar = pb;
  {
    uint64_t last = 0, best;
    uint32_t bi = 0;
    uint32_t r = nsel;
    for (bool first = true; r > 0 && ord.next_down(first, last, bi, best); first = false, last = best) {
      --r;
      if (!((act >> r) & 1ull)) continue;
      f3 n;
      unit(ND[bi], n);
      xbar = xbar - n * dot(n, pb);
      if ((proj >> r) & 1ull) ub = ub - dot(ub, n) * n;
    }
  }
  const float w = (subh * mask) / mass;
  const f3 fbi = ub * w;
  st3(a.fb, i, fbi);
  st3(a.xb, i, xbar);
  st3(a.vb, i, ub + a.kair * fbi);
}

__global__ __launch_bounds__(AB) void k_adj_springs(AdjArgs a) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * AB + threadIdx.x;
  if (i >= a.V) return;
  const uint32_t g = a.voff + i;
  const f3 p = ld3(a.x, g), vel = ld3(a.v, g), fbi = ld3(a.fb, i);
  const float kd = a.kd;
  f3 xa = {0.f, 0.f, 0.f}, va = {0.f, 0.f, 0.f};
  float kdl = 0.f;
  for_grid_springs(a.P, i, [&](uint32_t j, float rest, float, uint32_t sidx) {
    const f3 dx = ld3(a.x, a.voff + j) - p;
    const f3 dv = ld3(a.v, a.voff + j) - vel;
    const float len = sqrtf(dot(dx, dx));
    if (len < 1e-7f) return;
    const f3 n = {dx.x / len, dx.y / len, dx.z / len};
    const float ks = a.ks[sidx];
    const float ndv = dot(dv, n);
    const float e = len - rest;
    const float s = ks * e + kd * ndv;
    const f3 gg = fbi - ld3(a.fb, j);
    const float ng = dot(n, gg);
    const float inv = 1.0f / len;
    const f3 Mdv = (dv - n * ndv) * inv;
    const f3 Mg = (gg - n * ng) * inv;
    const f3 JT = (ks * n + kd * Mdv) * ng + s * Mg;
    xa = xa - JT;
    va = va - (kd * ng) * n;
    if (i < j) {
      a.gk[sidx] = a.gk[sidx] + e * ng;
      kdl = kdl + ndv * ng;
    }
  });
  st3(a.xb, i, ld3(a.xb, i) + xa);
  st3(a.vb, i, ld3(a.vb, i) + va);
  a.kdv[i] = a.kdv[i] + kdl;
}

__global__ __launch_bounds__(AB) void k_adj_scale(const float* gk, const float* k, float* out, uint32_t n) {
  const uint32_t s = blockIdx.x * AB + threadIdx.x;
  if (s < n) out[s] = gk[s] * k[s];
}

// One block: thread t sums kdv[t], kdv[t + 256], ... then a halving tree.
__global__ __launch_bounds__(AB) void k_adj_sum(const float* kdv, uint32_t V, float* out) {
  __shared__ float sh[AB];
  float acc = 0.f;
  for (uint32_t i = threadIdx.x; i < V; i += AB) acc = acc + kdv[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t w = AB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

int launch_adj_exp(hipStream_t s, const float* theta, float* k, uint32_t n) {
  if (!n) return YSIM_OK;
  hipLaunchKernelGGL(k_adj_exp, dim3(ceil_div(n, AB)), dim3(AB), 0, s, theta, k, n);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

int launch_adj_step(hipStream_t s, const AdjArgs& a) {
  if (!a.V) return YSIM_OK;
  hipLaunchKernelGGL(k_adj_integrate, dim3(ceil_div(a.V, AB)), dim3(AB), 0, s, a);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_adj_springs, dim3(ceil_div(a.V, AB)), dim3(AB), 0, s, a);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

int launch_adj_finish(hipStream_t s, const float* gk, const float* k, float* g_log_k, uint32_t n, const float* kdv,
                      uint32_t V, float* g_kd) {
  if (g_log_k && n) {
    hipLaunchKernelGGL(k_adj_scale, dim3(ceil_div(n, AB)), dim3(AB), 0, s, gk, k, g_log_k, n);
    YSIM_CHECK_LAUNCH();
  }
  if (g_kd) {
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(AB), 0, s, kdv, V, g_kd);
    YSIM_CHECK_LAUNCH();
  }
  return YSIM_OK;
}

}  // namespace ysim
