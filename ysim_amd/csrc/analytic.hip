// analytic.hip — contacts of cloth vertices with analytic colliders (sphere /
// ellipsoid, box, capped cylinder, half-space).
//
// Reference: the analytic branch of BroadPhase::detectCollisions
// (include/sim/bvh.hpp:2897-2949: one (cloth, shape) marker when the cloth's
// root box meets the shape's swept box, objectBroadAABB 2486-2518; a plane
// always), the marker carry-over (PackedCollisionData::endAnalyticPairs,
// include/sim/scene.hpp:511-529: a pair runs when marked now or at the
// previous detection) and narrow_pt_analytic (src/metal/bruteforce.metal:
// 198-499) with the band of BruteForce::narrowAnalytic
// (include/sim/bruteforce.hpp:192-308).
//
// One thread per dynamic vertex runs every pair of its cloth: the pair test
// (two boxes, identical for the whole cloth; the cloth's first vertex stores
// it for the next substep), then the swept motion-inverted narrow test, and
// appends at most one contact per shape (target = the shape's mesh, facet 0)
// after the triangle contacts the query kernel left in the vertex's slots.
// The contact record is what the integrate kernel consumes for triangles, so
// the response needs no change.
#include "kernels.hpp"

namespace ysim {

namespace {

// quatRotate (bruteforce.metal:198-200): v + 2 cross(qv, cross(qv, v) + qw v).
__device__ __forceinline__ f3 quat_rotate(float qw, f3 qv, f3 v) {
  return v + 2.0f * cross(qv, cross(qv, v) + qw * v);
}
__device__ __forceinline__ float len3(f3 v) { return sqrtf(dot(v, v)); }
// normalize: v * rsqrt(dot(v, v)) with a correctly rounded 1 / sqrt
__device__ __forceinline__ f3 unit(f3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
__device__ __forceinline__ float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// boxDominantAxis (bruteforce.metal:209-217)
__device__ __forceinline__ void dominant_axis(f3 q, f3 h, int& axis, float& sg) {
  const float gx = fabsf(q.x) - h.x, gy = fabsf(q.y) - h.y, gz = fabsf(q.z) - h.z;
  axis = 0;
  float best = gx;
  if (gy > best) { best = gy; axis = 1; }
  if (gz > best) { best = gz; axis = 2; }
  sg = comp(q, axis) >= 0.0f ? 1.0f : -1.0f;
}

// narrow_pt_analytic geometry for one vertex: a / b are the previous / current
// positions in the shape's previous / current local frame.
__device__ bool analytic_contact(const ysim_analytic_shape& sh, f3 pcur, f3 pprev, float band, f3& n_out,
                                 float& d_out) {
  const float4 C = *reinterpret_cast<const float4*>(sh.center_radius);
  const float4 H = *reinterpret_cast<const float4*>(sh.half_ext_height);
  const float4 Q = *reinterpret_cast<const float4*>(sh.rot_quat);
  const float4 PC = *reinterpret_cast<const float4*>(sh.prev_center_pad);
  const float4 PQ = *reinterpret_cast<const float4*>(sh.prev_rot_quat);
  const f3 a = quat_rotate(PQ.x, f3{-PQ.y, -PQ.z, -PQ.w}, pprev - f3{PC.x, PC.y, PC.z});
  const f3 b = quat_rotate(Q.x, f3{-Q.y, -Q.z, -Q.w}, pcur - f3{C.x, C.y, C.z});
  const f3 seg = b - a;
  f3 nl{0.0f, 1.0f, 0.0f};
  float d = 1e30f;
  const uint32_t kind = sh.kind;
  if (kind == YSIM_COLLIDER_PLANE) {
    d = b.y;
  } else if (kind == YSIM_COLLIDER_SPHERE) {
    const f3 e{fmaxf(H.x, 1e-8f), fmaxf(H.y, 1e-8f), fmaxf(H.z, 1e-8f)};
    const float R = 1.0f + band / fminf(e.x, fminf(e.y, e.z));
    const f3 A{a.x / e.x, a.y / e.y, a.z / e.z}, B{b.x / e.x, b.y / e.y, b.z / e.z};
    const float lenA = len3(A);
    f3 u{0.0f, 1.0f, 0.0f};
    bool got = false;
    if (lenA < 1.0f) {
      // entry-side rule: already inside, push along a's direction
      u = (lenA > 1e-6f) ? A : ((len3(B) > 1e-6f) ? B : f3{0.0f, 1.0f, 0.0f});
      got = true;
    } else {
      const f3 D = B - A;
      const float qa = dot(D, D), qb = dot(A, D), qc = dot(A, A) - R * R;
      float t = -1.0f;
      if (qa > 1e-20f) {
        const float disc = qb * qb - qa * qc;
        if (disc >= 0.0f) {
          const float t0 = (-qb - sqrtf(disc)) / qa;
          if (t0 >= 0.0f && t0 <= 1.0f) t = t0;
        }
      }
      if (t >= 0.0f) {
        u = A + t * D;
        got = true;
      } else if (len3(B) > 1e-6f) {
        u = B;
        got = true;
      }
    }
    if (!got) return false;
    const f3 uh = unit(u);
    const f3 s{e.x * uh.x, e.y * uh.y, e.z * uh.z};
    nl = unit(f3{uh.x / e.x, uh.y / e.y, uh.z / e.z});
    d = dot(nl, b - s);
  } else if (kind == YSIM_COLLIDER_BOX) {
    const f3 h{fmaxf(H.x, 1e-8f), fmaxf(H.y, 1e-8f), fmaxf(H.z, 1e-8f)};
    const f3 HB{h.x + band, h.y + band, h.z + band};
    int k = 0;
    float sg = 1.0f;
    if (fabsf(a.x) < h.x && fabsf(a.y) < h.y && fabsf(a.z) < h.z) {
      dominant_axis(a, h, k, sg);
    } else {
      // slab test of the segment against the band-inflated box
      float tenter = 0.0f, texit = 1.0f, esg = 1.0f;
      int ek = -1;
      bool miss = false;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (miss) break;
        const float di = comp(seg, i), ai = comp(a, i), Hi = comp(HB, i);
        if (fabsf(di) < 1e-12f) {
          if (fabsf(ai) > Hi) miss = true;
        } else {
          const float inv = 1.0f / di;
          const float t1 = (-Hi - ai) * inv, t2 = (Hi - ai) * inv;
          const float tn = fminf(t1, t2), tf = fmaxf(t1, t2);
          if (tn > tenter) {
            tenter = tn;
            ek = i;
            esg = (di > 0.0f) ? -1.0f : 1.0f;
          }
          texit = fminf(texit, tf);
        }
      }
      if (!miss && ek >= 0 && tenter <= texit && tenter <= 1.0f) {
        k = ek;
        sg = esg;
      } else {
        dominant_axis(b, h, k, sg);
      }
    }
    nl = f3{k == 0 ? sg : 0.0f, k == 1 ? sg : 0.0f, k == 2 ? sg : 0.0f};
    d = sg * comp(b, k) - comp(h, k);
  } else if (kind == YSIM_COLLIDER_CYLINDER) {
    const float r = fmaxf(C.w, 1e-8f), hh = fmaxf(H.w, 1e-8f);
    const float R = r + band, HH = hh + band;
    int feat = 0;
    float capsg = 1.0f, nx = 1.0f, nz = 0.0f;
    const bool ain = (sqrtf(a.x * a.x + a.z * a.z) < r) && (fabsf(a.y) < hh);
    bool resolved = false;
    if (!ain) {
      // radial quadratic of the band-inflated wall intersected with the cap slab
      float tenter = 0.0f, texit = 1.0f, ecap = 1.0f, ex = 1.0f, ez = 0.0f;
      int ef = -1;
      bool miss = false;
      const float qa = seg.x * seg.x + seg.z * seg.z, qb = a.x * seg.x + a.z * seg.z,
                  qc = (a.x * a.x + a.z * a.z) - R * R;
      if (qa > 1e-20f) {
        const float disc = qb * qb - qa * qc;
        if (disc < 0.0f) {
          miss = true;
        } else {
          const float sq = sqrtf(disc);
          const float t1 = (-qb - sq) / qa, t2 = (-qb + sq) / qa;
          if (t1 > tenter) {
            tenter = t1;
            ef = 0;
          }
          texit = fminf(texit, t2);
        }
      } else if (qc > 0.0f) {
        miss = true;
      }
      if (!miss) {
        if (fabsf(seg.y) < 1e-12f) {
          if (fabsf(a.y) > HH) miss = true;
        } else {
          const float inv = 1.0f / seg.y;
          const float t1 = (-HH - a.y) * inv, t2 = (HH - a.y) * inv;
          const float tn = fminf(t1, t2), tf = fmaxf(t1, t2);
          if (tn > tenter) {
            tenter = tn;
            ef = 1;
            ecap = (seg.y > 0.0f) ? -1.0f : 1.0f;
          }
          texit = fminf(texit, tf);
        }
      }
      if (!miss && ef >= 0 && tenter <= texit && tenter <= 1.0f) {
        if (ef == 0) {
          const float hx = a.x + tenter * seg.x, hz = a.z + tenter * seg.z;
          const float l = sqrtf(hx * hx + hz * hz);
          if (l > 1e-9f) {
            ex = hx / l;
            ez = hz / l;
          }
        }
        feat = ef;
        capsg = ecap;
        nx = ex;
        nz = ez;
        resolved = true;
      }
    }
    if (!resolved) {
      const f3 q = ain ? a : b;
      const float lxz = sqrtf(q.x * q.x + q.z * q.z);
      feat = ((r - lxz) <= (hh - fabsf(q.y))) ? 0 : 1;
      capsg = (q.y >= 0.0f) ? 1.0f : -1.0f;
      if (lxz > 1e-9f) {
        nx = q.x / lxz;
        nz = q.z / lxz;
      } else {
        nx = 1.0f;
        nz = 0.0f;
      }
    }
    if (feat == 0) {
      nl = f3{nx, 0.0f, nz};
      d = dot(nl, b) - r;
    } else {
      nl = f3{0.0f, capsg, 0.0f};
      d = capsg * b.y - hh;
    }
  } else {
    return false;
  }
  if (!(d < band)) return false;
  n_out = unit(quat_rotate(Q.x, f3{Q.y, Q.z, Q.w}, nl));
  d_out = d;
  return true;
}

__global__ __launch_bounds__(256) void k_analytic(AnalyticArgs a) {
  const uint32_t j = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (j >= a.nitems) return;
  const uint32_t sg = seg_of(a.dsegs, j);
  const uint32_t pb = a.pair_off[sg], pe = a.pair_off[sg + 1];
  if (pb == pe) return;
  const uint32_t m = a.dsegs.mesh[sg];
  const uint32_t i = j - a.dsegs.off[sg];
  const uint32_t g = a.mt.voff[m] + i;
  const ysim_bvh_node root = a.tree[a.mt.noff[m]];
  const float band = a.mesh_band[m];
  const float thick = a.mesh_thick[m];
  const f3 pcur = ld3(a.x, g), pprev = ld3(a.xprev, g);
  uint32_t cnt = a.ccount[j];
  const uint32_t cnt0 = cnt;
  for (uint32_t p = pb; p < pe; ++p) {
    const AShape& S = a.shapes[a.pair_shape[p]];
    bool hit = true;
    if (S.s.kind != YSIM_COLLIDER_PLANE) {
      const float pad = band + S.rot_band;
      for (int c = 0; c < 3; ++c)
        if (S.hi[c] + pad < root.min[c] || S.lo[c] - pad > root.max[c]) hit = false;
    }
    if (i == 0) a.mark_cur[p] = hit ? 1u : 0u;
    if (!(hit || a.mark_prev[p]) || !(S.s.flags & 1u)) continue;
    const float rotm =
        S.theta > 0.0 ? (float)(S.theta * ((S.rmax + (double)a.radius) + (double)thick)) : 0.0f;
    f3 n;
    float d;
    if (!analytic_contact(S.s, pcur, pprev, band + rotm, n, d)) continue;
    if (cnt < a.cap) {
      const size_t slot = (size_t)j * a.cap + cnt;
      a.cnd[slot] = make_float4(n.x, n.y, n.z, d);
      a.cid[slot] = make_uint2(S.s.obj_index, 0u);
    }
    ++cnt;
  }
  if (cnt != cnt0) {
    a.ccount[j] = cnt;
    if (cnt > a.cap) atomicOr(a.flags, 2u);
  }
}

}  // namespace

int launch_analytic(hipStream_t s, const AnalyticArgs& a) {
  if (!a.nitems) return YSIM_OK;
  hipLaunchKernelGGL(k_analytic, dim3(ceil_div(a.nitems, 256)), dim3(256), 0, s, a);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

}  // namespace ysim
