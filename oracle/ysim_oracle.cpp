// ysim_oracle.cpp — CPU oracle for the cloth collision + spring substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product path (ysim_amd/ and its HIP library must not link or call it).
//
// It restates, in strict IEEE arithmetic (built with -ffp-contract=off, no
// fast-math), the semantics of the reference's Metal kernels and host loops.
// The reference itself cannot be built here (metal-cpp needs objc/runtime.h,
// Eigen3 and xcrun metal are absent), so this restatement *is* the reference
// CPU path.  Each function cites the reference file:line it follows.
// Templated on the real type: float is the parity path, double is the FP64
// twin used for error bars and for classifying threshold ties.
//
// Parity pins (tests/test_oracle_*.py): Morton vs an independent bit loop
// (exhaustive per axis), sort vs numpy stable argsort, Karras vs a top-down
// radix-split construction, refit vs recursive union, candidate set vs brute
// force O(V*T), narrow vs the FP64 twin (ties reported), forces vs the FP64
// twin and closed-form free fall, CSR order vs lexsort.

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ysim_hip.h"

namespace {

// ---------------------------------------------------------------- Morton
// expandBits, src/metal/bvh.metal:26-32 (10 bits -> 30 bits).
inline uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
// 21 bits -> 63 bits (new in this build; the reference has only 30-bit codes).
inline uint64_t expand21(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x001f00000000ffffull;
  v = (v | (v << 16)) & 0x001f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

inline float fmin_(float a, float b) { return std::fmin(a, b); }
inline float fmax_(float a, float b) { return std::fmax(a, b); }

// mortonCode + fillMortons_Tri, src/metal/bvh.metal:34-63.
inline uint64_t morton_tri(const float* x, const uint32_t* f, const float* bmin,
                           const float* bmax, uint32_t bits) {
  uint64_t e[3];
  for (int k = 0; k < 3; ++k) {
    float x0 = x[3 * f[0] + k], x1 = x[3 * f[1] + k], x2 = x[3 * f[2] + k];
    float center = ((x0 + x1) + x2) / 3.0f;
    float width = fmax_(bmax[k] - bmin[k], 1e-8f);
    float c = (center - bmin[k]) / width;
    if (bits == 30) {
      float p = fmin_(fmax_(c * 1024.0f, 0.0f), 1023.0f);
      e[k] = expand10((uint32_t)p);
    } else {
      float p = fmin_(fmax_(c * 2097152.0f, 0.0f), 2097151.0f);
      e[k] = expand21((uint64_t)(uint32_t)p);
    }
  }
  if (bits == 30) return e[0] * 4 + e[1] * 2 + e[2];
  return (e[0] << 2) | (e[1] << 1) | e[2];
}

inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

// delta / determineRange / findSplit, src/metal/bvh.metal:197-258, with the
// 64-bit variant (clz64, 64 + clz(idx ^ idx) tie-break) for 63-bit codes.
struct Karras {
  const uint64_t* code;
  const uint32_t* idx;
  int n;
  bool wide;
  int delta(int i, int j) const {
    if (j < 0 || j >= n) return -1;
    uint64_t ci = code[i], cj = code[j];
    if (ci != cj) return wide ? clz64(ci ^ cj) : clz32((uint32_t)(ci ^ cj));
    return (wide ? 64 : 32) + clz32(idx[i] ^ idx[j]);
  }
  void range(int index, int& lo, int& hi) const {
    int d = (delta(index, index + 1) - delta(index, index - 1) >= 0) ? 1 : -1;
    int dmin = delta(index, index - d);
    int lmax = 2;
    while (delta(index, index + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
      if (delta(index, index + (l + t) * d) > dmin) l += t;
    int j = index + l * d;
    if (d < 0) { lo = j; hi = index; } else { lo = index; hi = j; }
  }
  // findSplit, bvh.metal:197-224, with one deliberate change: the common
  // prefix is measured with delta() (code, then index tie-break) in every
  // case.  The reference returns the midpoint when the end codes are equal,
  // which disagrees with determineRange's index tie-break for runs of >= 3
  // equal codes and yields nodes that claim the same child.  When the end codes
  // differ the two formulations pick the same split.
  uint32_t split(int i0, int i1) const {
    int common = delta(i0, i1);
    uint32_t s = (uint32_t)i0, step = (uint32_t)(i1 - i0);
    do {
      step = (step + 1) >> 1;
      uint32_t ns = s + step;
      if (ns < (uint32_t)i1 && delta(i0, (int)ns) > common) s = ns;
    } while (step > 1);
    return s;
  }
};

template <class R>
struct V3 {
  R x, y, z;
};
template <class R>
inline V3<R> ld(const R* a, size_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
template <class R>
inline void st(R* a, size_t i, V3<R> v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }
template <class R>
inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R>
inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R>
inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R>
inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R>
inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R>
inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R>
inline R sqrt_(R v) { return std::sqrt(v); }

// calc_spring, src/metal/physics.metal:146-153.
template <class R>
inline V3<R> calc_spring(V3<R> p0, V3<R> v0, V3<R> p1, V3<R> v1, R rest, R ks, R kd) {
  V3<R> dx = p1 - p0;
  V3<R> dv = v1 - v0;
  R len = sqrt_(dot(dx, dx));
  if (len < (R)1e-7) return {0, 0, 0};
  V3<R> ndx = {dx.x / len, dx.y / len, dx.z / len};
  R s = ks * (len - rest) + kd * dot(dv, ndx);
  return s * ndx;
}

// Spring numbering of a FastGrid cloth with side N (this build's convention for
// per-spring stiffness; the reference uses one k per spring family):
//   struct-x (r,c)-(r,c+1): r*(N-1)+c              [N*(N-1)]
//   struct-y (r,c)-(r+1,c): SY + r*N+c             [(N-1)*N]
//   shear-a  (r,c)-(r+1,c+1): SA + r*(N-1)+c       [(N-1)^2]
//   shear-b  (r,c+1)-(r+1,c): SB + r*(N-1)+c       [(N-1)^2]
//   bend-x   (r,c)-(r,c+2): BX + r*(N-2)+c         [N*(N-2)]
//   bend-y   (r,c)-(r+2,c): BY + r*N+c             [(N-2)*N]
struct SpringIndex {
  uint32_t N, SY, SA, SB, BX, BY, total;
  explicit SpringIndex(uint32_t n) : N(n) {
    SY = N * (N - 1);
    SA = SY + (N - 1) * N;
    SB = SA + (N - 1) * (N - 1);
    BX = SB + (N - 1) * (N - 1);
    BY = BX + N * (N >= 2 ? N - 2 : 0);
    total = BY + (N >= 2 ? N - 2 : 0) * N;
  }
};

// compute_cloth_grid_forces_fast, src/metal/physics.metal:171-224 (gather of 12
// grid springs in this exact order).  kspring (optional) holds one stiffness per
// spring in SpringIndex order and replaces the per-family k.
template <class R>
void fastgrid_forces(const ysim_fastgrid_params& P, uint32_t nV, R kair, R kd, const R* x,
                     const R* v, const R* ext, const R* kspring, R* f) {
  const uint32_t N = P.particle_num_1d;
  SpringIndex S(N);
  const R rX = P.stretch_rest_x, rY = P.stretch_rest_y, rA = P.shear_rest_a, rB = P.shear_rest_b,
          bX = P.bend_rest_x, bY = P.bend_rest_y;
  const R ks = P.kstretch, kh = P.kshear, kb = P.kbend;
#pragma omp parallel for schedule(static)
  for (int64_t id_ = 0; id_ < (int64_t)nV; ++id_) {
    uint32_t id = (uint32_t)id_;
    uint32_t row = id / N, col = id % N;
    V3<R> pos = ld(x, id), vel = ld(v, id);
    V3<R> force = ld(ext, id) + vel * kair;
    auto nb = [&](uint32_t r, uint32_t c, R rest, R k, uint32_t sidx) {
      uint32_t j = r * N + c;
      R kk = kspring ? kspring[sidx] : k;
      force = force + calc_spring(pos, vel, ld(x, j), ld(v, j), rest, kk, kd);
    };
    if (col > 0) nb(row, col - 1, rX, ks, row * (N - 1) + col - 1);
    if (col < N - 1) nb(row, col + 1, rX, ks, row * (N - 1) + col);
    if (row > 0) nb(row - 1, col, rY, ks, S.SY + (row - 1) * N + col);
    if (row < N - 1) nb(row + 1, col, rY, ks, S.SY + row * N + col);
    if (col > 0 && row > 0) nb(row - 1, col - 1, rA, kh, S.SA + (row - 1) * (N - 1) + col - 1);
    if (col < N - 1 && row > 0) nb(row - 1, col + 1, rB, kh, S.SB + (row - 1) * (N - 1) + col);
    if (col > 0 && row < N - 1) nb(row + 1, col - 1, rB, kh, S.SB + row * (N - 1) + col - 1);
    if (col < N - 1 && row < N - 1) nb(row + 1, col + 1, rA, kh, S.SA + row * (N - 1) + col);
    if (col > 1) nb(row, col - 2, bX, kb, S.BX + row * (N - 2) + col - 2);
    if (col < N - 2) nb(row, col + 2, bX, kb, S.BX + row * (N - 2) + col);
    if (row > 1) nb(row - 2, col, bY, kb, S.BY + (row - 2) * N + col);
    if (row < N - 2) nb(row + 2, col, bY, kb, S.BY + row * N + col);
    st(f, id, force);
  }
}

// nonFinite3 / sanitizeIntegrateOutput, src/metal/physics.metal:40-74.
template <class R>
inline bool nonfinite(V3<R> p) {
  return !std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z);
}
template <class R>
inline bool sanitize(V3<R>& pos, V3<R>& vel, V3<R> fb) {
  const R B = (R)1.0e4f;
  bool touched = false;
  if (nonfinite(pos)) {
    pos = nonfinite(fb) ? V3<R>{0, 0, 0} : fb;
    vel = {0, 0, 0};
    touched = true;
  }
  if (nonfinite(vel)) {
    vel = {0, 0, 0};
    touched = true;
  }
  V3<R> bx = {std::fmin(std::fmax(pos.x, -B), B), std::fmin(std::fmax(pos.y, -B), B),
              std::fmin(std::fmax(pos.z, -B), B)};
  if (bx.x != pos.x || bx.y != pos.y || bx.z != pos.z) {
    pos = bx;
    vel = {0, 0, 0};
    touched = true;
  }
  return touched;
}

struct Contact {  // one narrow contact, bucketed per global query vertex
  uint32_t gv;    // statesOffsets[q] + point
  uint32_t tobj, tri, qobj, point;
  float n[3], d;
  uint32_t qbeh, tbeh, qshape, tshape;
};

template <class R>
struct ContactR {
  uint32_t tobj, tri;
  R n[3], d;
};

// integrate_cloth_grid, src/metal/physics.metal:226-312 (one vertex).  cs is the
// vertex's contact bucket in (target_obj, triangle) order.
template <class R>
inline bool integrate_vertex(V3<R>& pos, V3<R>& vel, V3<R> force, R m, R mask, R subh,
                             R thickness, const ContactR<R>* cs, uint32_t nc) {
  V3<R> xold = pos;
  vel.x = vel.x + ((subh * force.x) / m) * mask;
  vel.y = vel.y + ((subh * force.y) / m) * mask;
  vel.z = vel.z + ((subh * force.z) / m) * mask;
  for (uint32_t i = 0; i < nc; ++i) {
    bool dup = false;
    for (uint32_t j = 0; j < i; ++j)
      if (cs[j].tobj == cs[i].tobj && cs[j].tri == cs[i].tri) { dup = true; break; }
    if (dup) continue;
    V3<R> n = {cs[i].n[0], cs[i].n[1], cs[i].n[2]};
    R nlen2 = dot(n, n);
    if (nlen2 < (R)1e-12f) continue;
    R inv = (R)1 / sqrt_(nlen2);
    n = n * inv;
    R vn = dot(vel, n);
    R dist = cs[i].d;
    if (dist < thickness) {
      if (vn < (R)0) vel = vel - vn * n;
      pos = pos + (thickness - dist) * n;
    }
  }
  pos = pos + (subh * vel) * mask;
  return sanitize(pos, vel, xold);
}

// pointInTriangleBary, src/metal/bruteforce.metal:9-28.
template <class R>
inline bool point_in_tri(V3<R> p, V3<R> v0, V3<R> v1) {
  R d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(p, v0), d21 = dot(p, v1);
  R denom = d00 * d11 - d01 * d01;
  if (std::fabs(denom) < (R)1e-12f) return false;
  R b = (d11 * d20 - d01 * d21) / denom;
  R c = (d00 * d21 - d01 * d20) / denom;
  R a = (R)1 - b - c;
  return a >= (R)0 && b >= (R)0 && c >= (R)0;
}

// narrow_pt_tri geometry, src/metal/bruteforce.metal:114-158.  Returns true on a
// contact, with the oriented unit normal and the signed current distance.
template <class R>
inline bool narrow_test(V3<R> qcur, V3<R> qprev, V3<R> t0, V3<R> t1, V3<R> t2, R band,
                        V3<R>& n_out, R& d_out) {
  V3<R> v0 = t1 - t0, v1 = t2 - t0;
  V3<R> n = cross(v0, v1);
  R nlen2 = dot(n, n);
  if (nlen2 < (R)1e-12f) return false;
  R inv = (R)1 / sqrt_(nlen2);
  n = n * inv;
  R dprev = dot(n, qprev - t0);
  R dcur = dot(n, qcur - t0);
  R dref = (dprev != (R)0) ? dprev : dcur;
  if (dref < (R)0) {
    n = {-n.x, -n.y, -n.z};
    dprev = -dprev;
    dcur = -dcur;
  }
  bool crossed = dcur < (R)0;
  bool inmargin = dcur < band;
  if (!crossed && !inmargin) return false;
  V3<R> cp;
  if (crossed && dprev > (R)0) {
    R t = dprev / (dprev - dcur);
    cp = qprev + (qcur - qprev) * t;
  } else {
    cp = qcur;
  }
  V3<R> pc = cp - t0;
  R l = dot(n, pc);
  V3<R> inplane = pc - n * l;
  if (!point_in_tri(inplane, v0, v1)) return false;
  n_out = n;
  d_out = dcur;
  return true;
}

// Oriented current distance of narrow_test (the value compared with the band
// at src/metal/bruteforce.metal:149-151), or false for a degenerate facet.
template <class R>
inline bool narrow_dcur(V3<R> qcur, V3<R> qprev, V3<R> t0, V3<R> t1, V3<R> t2, R& d_out) {
  V3<R> n = cross(t1 - t0, t2 - t0);
  R nlen2 = dot(n, n);
  if (nlen2 < (R)1e-12f) return false;
  n = n * ((R)1 / sqrt_(nlen2));
  R dprev = dot(n, qprev - t0);
  R dcur = dot(n, qcur - t0);
  R dref = (dprev != (R)0) ? dprev : dcur;
  d_out = dref < (R)0 ? -dcur : dcur;
  return true;
}

inline V3<double> up(V3<float> a) { return {(double)a.x, (double)a.y, (double)a.z}; }

// Threshold ties of one candidate (the FP32 inputs evaluated again in FP64):
// bit 0 when the FP32 and FP64 narrow decisions differ, bit 1 when the FP64
// distance lies within 4 FP32 ulps of the band (the strict compare could go
// either way under another rounding).
inline uint32_t tie_bits(V3<float> qc, V3<float> qp, V3<float> t0, V3<float> t1, V3<float> t2, float band,
                         bool fired32) {
  V3<double> n;
  double d = 0.0;
  const bool fired64 = narrow_test<double>(up(qc), up(qp), up(t0), up(t1), up(t2), (double)band, n, d);
  uint32_t b = (fired32 != fired64) ? 1u : 0u;
  double dc = 0.0;
  if (narrow_dcur<double>(up(qc), up(qp), up(t0), up(t1), up(t2), dc)) {
    const double ulp = (double)std::nextafter(band, 2.f * band + 1.f) - (double)band;
    if (std::fabs(dc - (double)band) <= 4.0 * ulp) b |= 2u;
  }
  return b;
}

inline bool box_overlap(const float* qmin, const float* qmax, const ysim_bvh_node& nd) {
  // intersectAABB, src/metal/bvh.metal:1285-1290 (closed intervals).
  if (qmax[0] < nd.min[0] || qmin[0] > nd.max[0]) return false;
  if (qmax[1] < nd.min[1] || qmin[1] > nd.max[1]) return false;
  if (qmax[2] < nd.min[2] || qmin[2] > nd.max[2]) return false;
  return true;
}

// ---------------------------------------------------------------- TriangularCloth
// MeshAdjacencyInitializer::initialize, include/sim/mesh_state.hpp:236-454.
// Three edge records per facet ((a,b) opposite c, (b,c) opposite a, (c,a)
// opposite b, each with v0 < v1), sorted by (v0, v1) and reduced to unique
// edges.  A vertex lists its edges in edge order (stretch springs); an edge met
// twice gives each of its two opposite vertices the other one (bend springs),
// also in edge order.  Rest lengths are |x[b] - x[a]| of the given positions
// (tinym vec3 norm: sqrt((x*x + y*y) + z*z)).  Which duplicate record sorts
// first does not matter: the two opposite vertices list each other and the
// norm is symmetric.  Returns false for a facet with a repeated vertex or an
// edge shared by more than two facets (the reference leaves entries of such
// edges unset).
struct TriAdj {
  std::vector<uint32_t> edges;  // (v0, v1) per edge
  std::vector<float> rest_edge;
  std::vector<uint32_t> adj_off, adj;  // per vertex: edge ids
  std::vector<uint32_t> opp_off, opp;  // per vertex: opposite vertices
  std::vector<float> rest_opp;
};

inline float edge_length(const float* x, uint32_t a, uint32_t b) {
  const float dx = x[3 * b] - x[3 * a], dy = x[3 * b + 1] - x[3 * a + 1], dz = x[3 * b + 2] - x[3 * a + 2];
  return std::sqrt((dx * dx + dy * dy) + dz * dz);
}

bool tri_adjacency(const uint32_t* F, uint32_t nF, const float* x, uint32_t nV, TriAdj& A) {
  struct Rec { uint32_t v0, v1, o; };
  std::vector<Rec> recs;
  recs.reserve(3 * (size_t)nF);
  for (uint32_t f = 0; f < nF; ++f) {
    const uint32_t a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
    if (a == b || b == c || c == a || a >= nV || b >= nV || c >= nV) return false;
    recs.push_back({std::min(a, b), std::max(a, b), c});
    recs.push_back({std::min(b, c), std::max(b, c), a});
    recs.push_back({std::min(c, a), std::max(c, a), b});
  }
  std::stable_sort(recs.begin(), recs.end(),
                   [](const Rec& p, const Rec& q) { return p.v0 < q.v0 || (p.v0 == q.v0 && p.v1 < q.v1); });
  struct Edge { uint32_t v0, v1, o0; int64_t o1; };
  std::vector<Edge> E;
  for (size_t k = 0; k < recs.size(); ++k) {
    if (!E.empty() && E.back().v0 == recs[k].v0 && E.back().v1 == recs[k].v1) {
      if (E.back().o1 >= 0) return false;  // third facet on one edge
      E.back().o1 = recs[k].o;
    } else {
      E.push_back({recs[k].v0, recs[k].v1, recs[k].o, -1});
    }
  }
  const uint32_t nE = (uint32_t)E.size();
  A.edges.resize(2 * (size_t)nE);
  A.rest_edge.resize(nE);
  A.adj_off.assign(nV + 1, 0);
  A.opp_off.assign(nV + 1, 0);
  for (uint32_t e = 0; e < nE; ++e) {
    A.edges[2 * e] = E[e].v0;
    A.edges[2 * e + 1] = E[e].v1;
    A.rest_edge[e] = edge_length(x, E[e].v0, E[e].v1);
    A.adj_off[E[e].v0 + 1]++;
    A.adj_off[E[e].v1 + 1]++;
    if (E[e].o1 >= 0) {
      A.opp_off[E[e].o0 + 1]++;
      A.opp_off[(uint32_t)E[e].o1 + 1]++;
    }
  }
  for (uint32_t i = 0; i < nV; ++i) {
    A.adj_off[i + 1] += A.adj_off[i];
    A.opp_off[i + 1] += A.opp_off[i];
  }
  A.adj.assign(A.adj_off[nV], 0);
  A.opp.assign(A.opp_off[nV], 0);
  A.rest_opp.assign(A.opp_off[nV], 0.f);
  std::vector<uint32_t> fa(nV, 0), fo(nV, 0);
  for (uint32_t e = 0; e < nE; ++e) {
    if (E[e].o1 >= 0) {
      const uint32_t o0 = E[e].o0, o1 = (uint32_t)E[e].o1;
      const float r = edge_length(x, o0, o1);
      A.opp[A.opp_off[o0] + fo[o0]] = o1;
      A.rest_opp[A.opp_off[o0] + fo[o0]++] = r;
      A.opp[A.opp_off[o1] + fo[o1]] = o0;
      A.rest_opp[A.opp_off[o1] + fo[o1]++] = r;
    }
    A.adj[A.adj_off[E[e].v0] + fa[E[e].v0]++] = e;
    A.adj[A.adj_off[E[e].v1] + fa[E[e].v1]++] = e;
  }
  return true;
}

// compute_tri_spring_forces, src/metal/physics.metal:319-378: external force +
// air drag, then the stretch springs of the vertex's edges, then the bend
// springs of its opposite vertices, in list order (k_shear is not used).
template <class R>
void tri_forces(const TriAdj& A, uint32_t nV, R kstretch, R kbend, R kair, R kd, const R* x, const R* v,
                const R* ext, R* f) {
#pragma omp parallel for schedule(static)
  for (int64_t id_ = 0; id_ < (int64_t)nV; ++id_) {
    const uint32_t id = (uint32_t)id_;
    V3<R> pos = ld(x, id), vel = ld(v, id);
    V3<R> force = ld(ext, id) + vel * kair;
    for (uint32_t k = A.adj_off[id]; k < A.adj_off[id + 1]; ++k) {
      const uint32_t e = A.adj[k];
      const uint32_t e0 = A.edges[2 * e], e1 = A.edges[2 * e + 1];
      const uint32_t other = (e0 == id) ? e1 : e0;
      force = force + calc_spring(pos, vel, ld(x, other), ld(v, other), (R)A.rest_edge[e], kstretch, kd);
    }
    for (uint32_t k = A.opp_off[id]; k < A.opp_off[id + 1]; ++k) {
      const uint32_t other = A.opp[k];
      force = force + calc_spring(pos, vel, ld(x, other), ld(v, other), (R)A.rest_opp[k], kbend, kd);
    }
    st(f, id, force);
  }
}

// ---------------------------------------------------------------- analytic colliders
// quatRotate, src/metal/bruteforce.metal:198-200: v + 2 cross(qv, cross(qv, v) + qw v).
template <class R>
inline V3<R> quat_rotate(R qw, V3<R> qv, V3<R> v) {
  return v + (R)2 * cross(qv, cross(qv, v) + qw * v);
}
template <class R>
inline R length_(V3<R> v) { return sqrt_(dot(v, v)); }
// normalize = v * rsqrt(dot(v, v)), restated with a correctly rounded
// reciprocal square root (1 / sqrt).
template <class R>
inline V3<R> normalize_(V3<R> v) { return v * ((R)1 / sqrt_(dot(v, v))); }
template <class R>
inline R comp(V3<R> v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// boxDominantAxis, src/metal/bruteforce.metal:209-217: the axis maximising
// |q_i| - h_i (first on ties) and the side of q on it.
template <class R>
inline void box_dominant_axis(V3<R> q, V3<R> h, int& axis, R& sg) {
  const R gx = std::fabs(q.x) - h.x, gy = std::fabs(q.y) - h.y, gz = std::fabs(q.z) - h.z;
  axis = 0;
  R best = gx;
  if (gy > best) { best = gy; axis = 1; }
  if (gz > best) { best = gz; axis = 2; }
  sg = comp(q, axis) >= (R)0 ? (R)1 : (R)-1;
}

// narrow_pt_analytic, src/metal/bruteforce.metal:252-499: motion inversion of
// the collider into the vertex segment a -> b (local frame), the contact
// feature from the entry point (or a when a is already inside: entry-side
// rule; else the closest feature of b), the current distance d along the
// feature normal, gate d < band, normal rotated back with the current pose.
template <class R>
bool narrow_analytic(const ysim_analytic_shape& sh, V3<R> pcur, V3<R> pprev, R band, V3<R>& n_out, R& d_out) {
  const float *C = sh.center_radius, *H = sh.half_ext_height, *Q = sh.rot_quat, *PC = sh.prev_center_pad,
              *PQ = sh.prev_rot_quat;
  const V3<R> a = quat_rotate<R>((R)PQ[0], V3<R>{-(R)PQ[1], -(R)PQ[2], -(R)PQ[3]},
                                 pprev - V3<R>{(R)PC[0], (R)PC[1], (R)PC[2]});
  const V3<R> b = quat_rotate<R>((R)Q[0], V3<R>{-(R)Q[1], -(R)Q[2], -(R)Q[3]},
                                 pcur - V3<R>{(R)C[0], (R)C[1], (R)C[2]});
  const V3<R> seg = b - a;
  V3<R> nl{(R)0, (R)1, (R)0};
  R d = (R)1e30f;
  if (sh.kind == YSIM_COLLIDER_PLANE) {
    d = b.y;  // outside is local +Y; s = (b.x, 0, b.z)
  } else if (sh.kind == YSIM_COLLIDER_SPHERE) {
    const V3<R> e{std::fmax((R)H[0], (R)1e-8f), std::fmax((R)H[1], (R)1e-8f), std::fmax((R)H[2], (R)1e-8f)};
    const R Rr = (R)1 + band / std::fmin(e.x, std::fmin(e.y, e.z));
    const V3<R> A{a.x / e.x, a.y / e.y, a.z / e.z}, B{b.x / e.x, b.y / e.y, b.z / e.z};
    const R lenA = length_(A);
    V3<R> u{(R)0, (R)1, (R)0};
    bool got = false;
    if (lenA < (R)1) {
      u = (lenA > (R)1e-6f) ? A : ((length_(B) > (R)1e-6f) ? B : V3<R>{(R)0, (R)1, (R)0});
      got = true;
    } else {
      const V3<R> D = B - A;
      const R qa = dot(D, D), qb = dot(A, D), qc = dot(A, A) - Rr * Rr;
      R t = (R)-1;
      if (qa > (R)1e-20f) {
        const R disc = qb * qb - qa * qc;
        if (disc >= (R)0) {
          const R t0 = (-qb - sqrt_(disc)) / qa;
          if (t0 >= (R)0 && t0 <= (R)1) t = t0;
        }
      }
      if (t >= (R)0) {
        u = A + t * D;
        got = true;
      } else if (length_(B) > (R)1e-6f) {
        u = B;
        got = true;
      }
    }
    if (!got) return false;
    const V3<R> uh = normalize_(u);
    const V3<R> s{e.x * uh.x, e.y * uh.y, e.z * uh.z};
    nl = normalize_(V3<R>{uh.x / e.x, uh.y / e.y, uh.z / e.z});
    d = dot(nl, b - s);
  } else if (sh.kind == YSIM_COLLIDER_BOX) {
    const V3<R> h{std::fmax((R)H[0], (R)1e-8f), std::fmax((R)H[1], (R)1e-8f), std::fmax((R)H[2], (R)1e-8f)};
    const V3<R> HB{h.x + band, h.y + band, h.z + band};
    int k = 0;
    R sg = (R)1;
    if (std::fabs(a.x) < h.x && std::fabs(a.y) < h.y && std::fabs(a.z) < h.z) {
      box_dominant_axis(a, h, k, sg);
    } else {
      R tenter = (R)0, texit = (R)1;
      int ek = -1;
      R esg = (R)1;
      bool miss = false;
      for (int i = 0; i < 3 && !miss; ++i) {
        const R di = comp(seg, i), ai = comp(a, i), Hi = comp(HB, i);
        if (std::fabs(di) < (R)1e-12f) {
          if (std::fabs(ai) > Hi) miss = true;
        } else {
          const R inv = (R)1 / di;
          const R t1 = (-Hi - ai) * inv, t2 = (Hi - ai) * inv;
          const R tn = std::fmin(t1, t2), tf = std::fmax(t1, t2);
          if (tn > tenter) {
            tenter = tn;
            ek = i;
            esg = (di > (R)0) ? (R)-1 : (R)1;
          }
          texit = std::fmin(texit, tf);
        }
      }
      if (!miss && ek >= 0 && tenter <= texit && tenter <= (R)1) {
        k = ek;
        sg = esg;
      } else {
        box_dominant_axis(b, h, k, sg);
      }
    }
    nl = V3<R>{k == 0 ? sg : (R)0, k == 1 ? sg : (R)0, k == 2 ? sg : (R)0};
    d = sg * comp(b, k) - comp(h, k);
  } else if (sh.kind == YSIM_COLLIDER_CYLINDER) {
    const R r = std::fmax((R)C[3], (R)1e-8f), hh = std::fmax((R)H[3], (R)1e-8f);
    const R Rr = r + band, HH = hh + band;
    int feat = 0;
    R capsg = (R)1, nx = (R)1, nz = (R)0;
    const bool ain = (sqrt_(a.x * a.x + a.z * a.z) < r) && (std::fabs(a.y) < hh);
    bool resolved = false;
    if (!ain) {
      R tenter = (R)0, texit = (R)1;
      int ef = -1;
      R ecap = (R)1, ex = (R)1, ez = (R)0;
      bool miss = false;
      const R qa = seg.x * seg.x + seg.z * seg.z, qb = a.x * seg.x + a.z * seg.z,
              qc = (a.x * a.x + a.z * a.z) - Rr * Rr;
      if (qa > (R)1e-20f) {
        const R disc = qb * qb - qa * qc;
        if (disc < (R)0) {
          miss = true;
        } else {
          const R sq = sqrt_(disc);
          const R t1 = (-qb - sq) / qa, t2 = (-qb + sq) / qa;
          if (t1 > tenter) {
            tenter = t1;
            ef = 0;
          }
          texit = std::fmin(texit, t2);
        }
      } else if (qc > (R)0) {
        miss = true;
      }
      if (!miss) {
        if (std::fabs(seg.y) < (R)1e-12f) {
          if (std::fabs(a.y) > HH) miss = true;
        } else {
          const R inv = (R)1 / seg.y;
          const R t1 = (-HH - a.y) * inv, t2 = (HH - a.y) * inv;
          const R tn = std::fmin(t1, t2), tf = std::fmax(t1, t2);
          if (tn > tenter) {
            tenter = tn;
            ef = 1;
            ecap = (seg.y > (R)0) ? (R)-1 : (R)1;
          }
          texit = std::fmin(texit, tf);
        }
      }
      if (!miss && ef >= 0 && tenter <= texit && tenter <= (R)1) {
        if (ef == 0) {
          const R hx = a.x + tenter * seg.x, hz = a.z + tenter * seg.z;
          const R l = sqrt_(hx * hx + hz * hz);
          if (l > (R)1e-9f) {
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
      const V3<R> q = ain ? a : b;
      const R lxz = sqrt_(q.x * q.x + q.z * q.z);
      feat = ((r - lxz) <= (hh - std::fabs(q.y))) ? 0 : 1;
      capsg = (q.y >= (R)0) ? (R)1 : (R)-1;
      if (lxz > (R)1e-9f) {
        nx = q.x / lxz;
        nz = q.z / lxz;
      } else {
        nx = (R)1;
        nz = (R)0;
      }
    }
    if (feat == 0) {
      nl = V3<R>{nx, (R)0, nz};
      d = dot(nl, b) - r;
    } else {
      nl = V3<R>{(R)0, capsg, (R)0};
      d = capsg * b.y - hh;
    }
  } else {
    return false;
  }
  if (!(d < band)) return false;
  n_out = normalize_(quat_rotate<R>((R)Q[0], V3<R>{(R)Q[1], (R)Q[2], (R)Q[3]}, nl));
  d_out = d;
  return true;
}

// Geodesic angle between the current and previous orientations,
// 2 acos(|q . q_prev|) in double (bruteforce.hpp:204-211, bvh.hpp:2466-2472).
inline double quat_theta(const ysim_analytic_shape& s) {
  double dq = (double)s.rot_quat[0] * (double)s.prev_rot_quat[0] + (double)s.rot_quat[1] * (double)s.prev_rot_quat[1] +
              (double)s.rot_quat[2] * (double)s.prev_rot_quat[2] + (double)s.rot_quat[3] * (double)s.prev_rot_quat[3];
  dq = std::fabs(dq);
  if (dq > 1.0) dq = 1.0;
  return 2.0 * std::acos(dq);
}

// Hamilton product and rotateVector (include/sim/material_quat.hpp:28-35, 59-63):
// q * (0, v) * conj(q), products summed left to right.
struct Qf { float w, x, y, z; };
inline Qf qmul(Qf a, Qf b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
inline V3<float> rotate_vector(Qf q, V3<float> v) {
  const Qf r = qmul(qmul(q, Qf{0.f, v.x, v.y, v.z}), Qf{q.w, -q.x, -q.y, -q.z});
  return {r.x, r.y, r.z};
}

// BroadPhase::analyticWorldHalf (include/sim/bvh.hpp:2431-2460): world
// half-extents of the shape's box at orientation q (w, x, y, z).
inline V3<float> world_half(const ysim_analytic_shape& s, const float* q) {
  const float* H = s.half_ext_height;
  V3<float> h{0.f, 0.f, 0.f};
  if (s.kind == YSIM_COLLIDER_SPHERE) {
    const float r = std::max(std::max(H[0], H[1]), H[2]);
    return {r, r, r};
  } else if (s.kind == YSIM_COLLIDER_BOX) {
    h = {H[0], H[1], H[2]};
  } else if (s.kind == YSIM_COLLIDER_CYLINDER) {
    h = {s.center_radius[3], H[3], s.center_radius[3]};
  } else {
    return h;
  }
  const Qf Q{q[0], q[1], q[2], q[3]};
  const V3<float> cx = rotate_vector(Q, {1.f, 0.f, 0.f}), cy = rotate_vector(Q, {0.f, 1.f, 0.f}),
                  cz = rotate_vector(Q, {0.f, 0.f, 1.f});
  return {std::fabs(cx.x) * h.x + std::fabs(cy.x) * h.y + std::fabs(cz.x) * h.z,
          std::fabs(cx.y) * h.x + std::fabs(cy.y) * h.y + std::fabs(cz.y) * h.z,
          std::fabs(cx.z) * h.x + std::fabs(cy.z) * h.y + std::fabs(cz.z) * h.z};
}

// Per-shape constants: the swept box before padding (objectBroadAABB,
// bvh.hpp:2505-2517: min(c - hc, p - hp), max(c + hc, p + hp)), the broad
// rotation band (analyticRotBand, bvh.hpp:2465-2480) and the narrow rotation
// margin's angle and radius (BruteForce::analyticRotMargin,
// bruteforce.hpp:201-242: theta = 0 below 1e-9, plane 0).
struct ABounds {
  float lo[3], hi[3];
  float rot_band;
  double theta, rmax;
};
inline ABounds analytic_bounds(const ysim_analytic_shape& s) {
  ABounds B{};
  const float* H = s.half_ext_height;
  const float* C = s.center_radius;
  const float* P = s.prev_center_pad;
  const V3<float> hc = world_half(s, s.rot_quat), hp = world_half(s, s.prev_rot_quat);
  const float c[3] = {C[0], C[1], C[2]}, p[3] = {P[0], P[1], P[2]};
  const float hcv[3] = {hc.x, hc.y, hc.z}, hpv[3] = {hp.x, hp.y, hp.z};
  for (int k = 0; k < 3; ++k) {
    B.lo[k] = std::min(c[k] - hcv[k], p[k] - hpv[k]);
    B.hi[k] = std::max(c[k] + hcv[k], p[k] + hpv[k]);
  }
  const double theta = quat_theta(s);
  if (theta <= 0.0) {
    B.rot_band = 0.f;
  } else {
    const double rb = std::sqrt((double)H[0] * H[0] + (double)H[1] * H[1] + (double)H[2] * H[2]) + (double)C[3];
    B.rot_band = (float)(theta * rb);
  }
  B.theta = (theta > 1e-9 && s.kind != YSIM_COLLIDER_PLANE) ? theta : 0.0;
  switch (s.kind) {
    case YSIM_COLLIDER_SPHERE: B.rmax = std::max({(double)H[0], (double)H[1], (double)H[2]}); break;
    case YSIM_COLLIDER_BOX: B.rmax = std::sqrt((double)H[0] * H[0] + (double)H[1] * H[1] + (double)H[2] * H[2]); break;
    case YSIM_COLLIDER_CYLINDER: B.rmax = std::sqrt((double)C[3] * C[3] + (double)H[3] * H[3]); break;
    default: B.rmax = 0.0; break;
  }
  return B;
}
// rotMargin of one (shape, cloth) pair: theta (rMax + radius + thickness).
inline float rot_margin(const ABounds& B, float radius, float thickness) {
  return B.theta > 0.0 ? (float)(B.theta * (B.rmax + (double)radius + (double)thickness)) : 0.f;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ threads
int yo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// ------------------------------------------------------------------ S0 box
// BVH::build scene box, include/sim/bvh.hpp:1229-1231.
void yo_scene_box(const float* x, uint32_t nV, float* out6) {
  float mn[3] = {x[0], x[1], x[2]}, mx[3] = {x[0], x[1], x[2]};
  for (uint32_t i = 1; i < nV; ++i)
    for (int k = 0; k < 3; ++k) {
      mn[k] = fmin_(x[3 * i + k], mn[k]);
      mx[k] = fmax_(x[3 * i + k], mx[k]);
    }
  for (int k = 0; k < 3; ++k) { out6[k] = mn[k]; out6[3 + k] = mx[k]; }
}

// ------------------------------------------------------------------ S1 Morton
uint32_t yo_expand10(uint32_t v) { return expand10(v); }
uint64_t yo_expand21(uint64_t v) { return expand21(v); }

void yo_morton(const float* x, const uint32_t* facets, uint32_t nT, const float* box6,
               uint32_t bits, uint64_t* keys, uint32_t* idx) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)nT; ++i) {
    keys[i] = morton_tri(x, facets + 3 * i, box6, box6 + 3, bits);
    idx[i] = (uint32_t)i;
  }
}

// ------------------------------------------------------------------ S2 sort
// Stable LSD radix sort with 8-bit digits, count -> scan -> scatter per pass,
// the algorithm of RadixSorter<METAL>::sort (include/sim/spatial_hash.hpp:48-70,
// src/metal/radixsort.metal:21-112), widened to key_bits.
void yo_sort_pairs(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits) {
  std::vector<uint64_t> k2(n);
  std::vector<uint32_t> v2(n);
  uint64_t* ks = keys;
  uint32_t* vs = vals;
  uint64_t* kd = k2.data();
  uint32_t* vd = v2.data();
  uint32_t passes = (key_bits + 7) / 8;
  for (uint32_t p = 0; p < passes; ++p) {
    uint32_t shift = 8 * p;
    uint32_t hist[257] = {0};
    for (uint32_t i = 0; i < n; ++i) hist[((ks[i] >> shift) & 255u) + 1]++;
    for (int b = 1; b < 257; ++b) hist[b] += hist[b - 1];
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t b = (uint32_t)((ks[i] >> shift) & 255u);
      uint32_t dst = hist[b]++;
      kd[dst] = ks[i];
      vd[dst] = vs[i];
    }
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  if (ks != keys) {
    std::memcpy(keys, ks, sizeof(uint64_t) * n);
    std::memcpy(vals, vs, sizeof(uint32_t) * n);
  }
}

// ------------------------------------------------------------------ S3 Karras
// buildTree_Tri, src/metal/bvh.metal:352-393 (child links + parent), with
// leaves at slot n-1+i (child_a = -1, child_b = facet id) and root parent -1.
void yo_karras(const uint64_t* keys, const uint32_t* idx, uint32_t n, uint32_t bits,
               ysim_bvh_node* tree, int32_t* parent) {
  if (n == 0) return;
  Karras K{keys, idx, (int)n, bits > 32};
  for (uint32_t i = 0; i < n; ++i) {
    ysim_bvh_node& lf = tree[n - 1 + i];
    lf.child_a = -1;
    lf.child_b = (int32_t)idx[i];
  }
  parent[0] = -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n - 1; ++i) {
    int lo, hi;
    K.range((int)i, lo, hi);
    uint32_t s = K.split(lo, hi);
    int a = ((int)s == lo) ? (int)(s + n - 1) : (int)s;
    int b = ((int)s + 1 == hi) ? (int)(s + n) : (int)(s + 1);
    tree[i].child_a = a;
    tree[i].child_b = b;
    parent[a] = (int32_t)i;
    parent[b] = (int32_t)i;
  }
}

// Agglomerative (Apetrei 2014) hierarchy over the same sorted keys:
// agglomerativeBuild_Tri + agglomerativeSwapRoot (src/metal/bvh.metal:884-1011,
// 1135-1199).  Each leaf climbs: a node covering [L, R] merges toward the side
// whose neighbouring key pair differs less (code XOR, then facet-index XOR on
// equal codes), its parent is the internal node numbered by that boundary (R
// when merging right, L - 1 when merging left), and the second arrival at a
// parent continues.  The parent numbering makes every internal node the split
// of its range; the root, found at some slot k, is then swapped into slot 0
// with every link that named either slot retargeted.  Topology only (the same
// binary radix tree as yo_karras, numbered differently); boxes come from
// yo_refit.  Sequential: the result does not depend on arrival order.
void yo_agglomerative(const uint64_t* keys, const uint32_t* idx, uint32_t n, uint32_t bits, ysim_bvh_node* tree,
                      int32_t* parent) {
  if (n == 0) return;
  for (uint32_t i = 0; i < n; ++i) {
    ysim_bvh_node& lf = tree[n - 1 + i];
    lf.child_a = -1;
    lf.child_b = (int32_t)idx[i];
  }
  if (n == 1) {
    parent[0] = -1;
    return;
  }
  const uint64_t mask = (bits > 32) ? ~0ull : 0xffffffffull;
  std::vector<uint8_t> visit(n - 1, 0);
  std::vector<int> rl(n - 1, -1), rr(n - 1, -1);
  int root = -1;
  for (uint32_t i = 0; i < n; ++i) {
    int cur = (int)(n - 1 + i), L = (int)i, R = (int)i;
    while (true) {
      if (L == 0 && R == (int)n - 1) {
        root = cur;
        break;
      }
      bool right;
      if (L == 0) {
        right = true;
      } else if (R == (int)n - 1) {
        right = false;
      } else {
        const uint64_t dr = (keys[R] & mask) ^ (keys[R + 1] & mask), dl = (keys[L - 1] & mask) ^ (keys[L] & mask);
        right = (dr != dl) ? (dr < dl) : ((idx[R] ^ idx[R + 1]) < (idx[L - 1] ^ idx[L]));
      }
      int p;
      if (right) {
        p = R;
        tree[p].child_a = cur;
        rl[p] = L;
      } else {
        p = L - 1;
        tree[p].child_b = cur;
        rr[p] = R;
      }
      parent[cur] = p;
      if (visit[p]++ == 0) break;  // first arrival
      cur = p;
      L = rl[p];
      R = rr[p];
    }
  }
  // agglomerativeSwapRoot
  const int k = root;
  if (k > 0) {
    const ysim_bvh_node Rn = tree[k], On = tree[0];
    const int p0 = parent[0];
    const int ra = (Rn.child_a == 0) ? k : Rn.child_a, rb = (Rn.child_b == 0) ? k : Rn.child_b;
    tree[0] = Rn;
    tree[0].child_a = ra;
    tree[0].child_b = rb;
    tree[k] = On;
    parent[On.child_a] = k;
    parent[On.child_b] = k;
    parent[ra] = 0;
    parent[rb] = 0;
    if (p0 != k) {
      if (tree[p0].child_a == 0) tree[p0].child_a = k;
      else tree[p0].child_b = k;
      parent[k] = p0;
    }
  }
  parent[0] = -1;
}

// ------------------------------------------------------------------ S4/S5 refit
// Leaf boxes (buildLeaf_Tri / buildSweptLeaf_Tri, src/metal/bvh.metal:260-351;
// CPU enlarge bvh.hpp:1669-1680: s = x + v*dt with separate rounding) then exact
// union of the children for every internal node (bottomUpCombine,
// include/sim/bvh.hpp:1534-1553), in post-order.
void yo_refit(const float* x, const float* v, float dt, uint32_t swept, const uint32_t* facets,
              uint32_t n, ysim_bvh_node* tree) {
  if (n == 0) return;
  for (uint32_t i = 0; i < n; ++i) {
    ysim_bvh_node& lf = tree[n - 1 + i];
    const uint32_t* f = facets + 3 * (uint32_t)lf.child_b;
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) { lo[k] = x[3 * f[0] + k]; hi[k] = lo[k]; }
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) {
        float p = x[3 * f[c] + k];
        lo[k] = fmin_(lo[k], p);
        hi[k] = fmax_(hi[k], p);
        if (swept) {
          float s = p + v[3 * f[c] + k] * dt;
          lo[k] = fmin_(lo[k], s);
          hi[k] = fmax_(hi[k], s);
        }
      }
    for (int k = 0; k < 3; ++k) { lf.min[k] = lo[k]; lf.max[k] = hi[k]; }
  }
  if (n == 1) return;
  // iterative post-order from the root
  std::vector<std::pair<int, int>> st;
  st.push_back({0, 0});
  while (!st.empty()) {
    auto& top = st.back();
    int nd = top.first;
    if (tree[nd].child_a < 0) { st.pop_back(); continue; }
    if (top.second == 0) { top.second = 1; st.push_back({tree[nd].child_a, 0}); continue; }
    if (top.second == 1) { top.second = 2; st.push_back({tree[nd].child_b, 0}); continue; }
    const ysim_bvh_node& A = tree[tree[nd].child_a];
    const ysim_bvh_node& B = tree[tree[nd].child_b];
    for (int k = 0; k < 3; ++k) {
      tree[nd].min[k] = fmin_(A.min[k], B.min[k]);
      tree[nd].max[k] = fmax_(A.max[k], B.max[k]);
    }
    st.pop_back();
  }
}

// ------------------------------------------------------------------ S6 broad
// queryPoints/queryAABB, src/metal/bvh.metal:1322-1412: for each query vertex p,
// every leaf whose box overlaps [x_p - m, x_p + m] (closed), excluding facets
// that contain p when the query mesh is the target mesh.  Unbounded stack; the
// output order is DFS order, compared as a set.
uint32_t yo_broad_query(const float* xq, uint32_t nq, uint32_t qobj, const ysim_bvh_node* tree,
                        uint32_t nT, const uint32_t* tfacets, uint32_t tobj, float margin,
                        uint32_t* out_pairs /* 2 per pair, may be NULL */, uint32_t cap) {
  uint32_t count = 0;
  std::vector<int> stack;
  for (uint32_t p = 0; p < nq; ++p) {
    float qmin[3], qmax[3];
    for (int k = 0; k < 3; ++k) { qmin[k] = xq[3 * p + k] - margin; qmax[k] = xq[3 * p + k] + margin; }
    stack.clear();
    stack.push_back(0);
    while (!stack.empty()) {
      int nd = stack.back();
      stack.pop_back();
      const ysim_bvh_node& N = tree[nd];
      if (!box_overlap(qmin, qmax, N)) continue;
      if (N.child_a < 0) {
        uint32_t fid = (uint32_t)N.child_b;
        const uint32_t* f = tfacets + 3 * fid;
        if (qobj == tobj && (p == f[0] || p == f[1] || p == f[2])) continue;
        if (out_pairs && count < cap) { out_pairs[2 * count] = p; out_pairs[2 * count + 1] = fid; }
        ++count;
        continue;
      }
      stack.push_back(N.child_b);
      stack.push_back(N.child_a);
    }
  }
  (void)nT;
  return count;
}

// Brute force O(nq * nT) over explicit leaf boxes (leaf boxes given per facet
// id as 6 floats).  Independent of topology: pins the traversal.
uint32_t yo_brute_candidates(const float* xq, uint32_t nq, uint32_t qobj, const float* leafbox6,
                             uint32_t nT, const uint32_t* tfacets, uint32_t tobj, float margin,
                             uint32_t* out_pairs, uint32_t cap) {
  std::vector<uint32_t> counts(nq, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t p_ = 0; p_ < (int64_t)nq; ++p_) {
    uint32_t p = (uint32_t)p_;
    float qmin[3], qmax[3];
    for (int k = 0; k < 3; ++k) { qmin[k] = xq[3 * p + k] - margin; qmax[k] = xq[3 * p + k] + margin; }
    uint32_t c = 0;
    for (uint32_t t = 0; t < nT; ++t) {
      ysim_bvh_node nd;
      for (int k = 0; k < 3; ++k) { nd.min[k] = leafbox6[6 * t + k]; nd.max[k] = leafbox6[6 * t + 3 + k]; }
      if (!box_overlap(qmin, qmax, nd)) continue;
      const uint32_t* f = tfacets + 3 * t;
      if (qobj == tobj && (p == f[0] || p == f[1] || p == f[2])) continue;
      ++c;
    }
    counts[p] = c;
  }
  uint32_t total = 0;
  std::vector<uint32_t> off(nq);
  for (uint32_t p = 0; p < nq; ++p) { off[p] = total; total += counts[p]; }
  if (out_pairs && total <= cap) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p_ = 0; p_ < (int64_t)nq; ++p_) {
      uint32_t p = (uint32_t)p_;
      float qmin[3], qmax[3];
      for (int k = 0; k < 3; ++k) { qmin[k] = xq[3 * p + k] - margin; qmax[k] = xq[3 * p + k] + margin; }
      uint32_t o = off[p];
      for (uint32_t t = 0; t < nT; ++t) {
        ysim_bvh_node nd;
        for (int k = 0; k < 3; ++k) { nd.min[k] = leafbox6[6 * t + k]; nd.max[k] = leafbox6[6 * t + 3 + k]; }
        if (!box_overlap(qmin, qmax, nd)) continue;
        const uint32_t* f = tfacets + 3 * t;
        if (qobj == tobj && (p == f[0] || p == f[1] || p == f[2])) continue;
        out_pairs[2 * o] = p;
        out_pairs[2 * o + 1] = t;
        ++o;
      }
    }
  }
  return total;
}

// ------------------------------------------------------------------ S7 narrow
// narrow_pt_tri, src/metal/bruteforce.metal:41-178, for a list of pairs
// (point, tri) of one (query mesh, target mesh).  x/xprev of the query mesh and
// x of the target mesh are mesh-local packed arrays.  flags[i] = 1 if the pair
// fires.  Writes normal (3) + distance (1) per pair.  Self pairs with tri
// containing the point are skipped (the adjacency CSR check).
void yo_narrow_f32(const uint32_t* pairs, uint32_t np, const float* xq, const float* xqprev,
                   const float* xt, const uint32_t* tfacets, uint32_t same_obj, float radius,
                   float thickness, uint8_t* flags, float* nd4) {
  const float band = radius + thickness;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)np; ++i) {
    uint32_t p = pairs[2 * i], t = pairs[2 * i + 1];
    const uint32_t* f = tfacets + 3 * t;
    flags[i] = 0;
    if (same_obj && (p == f[0] || p == f[1] || p == f[2])) continue;
    V3<float> n;
    float d;
    if (narrow_test(ld(xq, p), ld(xqprev, p), ld(xt, f[0]), ld(xt, f[1]), ld(xt, f[2]), band, n, d)) {
      flags[i] = 1;
      nd4[4 * i] = n.x; nd4[4 * i + 1] = n.y; nd4[4 * i + 2] = n.z; nd4[4 * i + 3] = d;
    }
  }
}
void yo_narrow_f64(const uint32_t* pairs, uint32_t np, const double* xq, const double* xqprev,
                   const double* xt, const uint32_t* tfacets, uint32_t same_obj, double radius,
                   double thickness, uint8_t* flags, double* nd4) {
  const double band = radius + thickness;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)np; ++i) {
    uint32_t p = pairs[2 * i], t = pairs[2 * i + 1];
    const uint32_t* f = tfacets + 3 * t;
    flags[i] = 0;
    if (same_obj && (p == f[0] || p == f[1] || p == f[2])) continue;
    V3<double> n;
    double d;
    if (narrow_test(ld(xq, p), ld(xqprev, p), ld(xt, f[0]), ld(xt, f[1]), ld(xt, f[2]), band, n, d)) {
      flags[i] = 1;
      nd4[4 * i] = n.x; nd4[4 * i + 1] = n.y; nd4[4 * i + 2] = n.z; nd4[4 * i + 3] = d;
    }
  }
}

// ------------------------------------------------------------------ S9 springs
void yo_fastgrid_forces_f32(const ysim_fastgrid_params* P, uint32_t nV, float kair, float kd,
                            const float* x, const float* v, const float* ext, const float* kspring,
                            float* f) {
  fastgrid_forces<float>(*P, nV, kair, kd, x, v, ext, kspring, f);
}
void yo_fastgrid_forces_f64(const ysim_fastgrid_params* P, uint32_t nV, double kair, double kd,
                            const double* x, const double* v, const double* ext,
                            const double* kspring, double* f) {
  fastgrid_forces<double>(*P, nV, kair, kd, x, v, ext, kspring, f);
}
uint32_t yo_spring_count(uint32_t N) { return SpringIndex(N).total; }

// TriangularCloth adjacency (counts[0..2] = edges, stretch entries, bend
// entries); arrays are written when `edges` is not null.  Returns 0, or -1 for
// a rejected mesh.
int yo_tri_adjacency(const uint32_t* F, uint32_t nF, const float* x, uint32_t nV, uint32_t* counts,
                     uint32_t* edges, float* rest_edge, uint32_t* adj_off, uint32_t* adj, uint32_t* opp_off,
                     uint32_t* opp, float* rest_opp) {
  TriAdj A;
  if (!tri_adjacency(F, nF, x, nV, A)) return -1;
  counts[0] = (uint32_t)A.rest_edge.size();
  counts[1] = (uint32_t)A.adj.size();
  counts[2] = (uint32_t)A.opp.size();
  if (!edges) return 0;
  std::memcpy(edges, A.edges.data(), 4 * A.edges.size());
  std::memcpy(rest_edge, A.rest_edge.data(), 4 * A.rest_edge.size());
  std::memcpy(adj_off, A.adj_off.data(), 4 * A.adj_off.size());
  std::memcpy(adj, A.adj.data(), 4 * A.adj.size());
  std::memcpy(opp_off, A.opp_off.data(), 4 * A.opp_off.size());
  std::memcpy(opp, A.opp.data(), 4 * A.opp.size());
  std::memcpy(rest_opp, A.rest_opp.data(), 4 * A.rest_opp.size());
  return 0;
}
// TriangularCloth forces at (x, v), rest lengths from x0.
int yo_tri_forces_f32(const uint32_t* F, uint32_t nF, const float* x0, uint32_t nV, float kstretch, float kbend,
                      float kair, float kd, const float* x, const float* v, const float* ext, float* f) {
  TriAdj A;
  if (!tri_adjacency(F, nF, x0, nV, A)) return -1;
  tri_forces<float>(A, nV, kstretch, kbend, kair, kd, x, v, ext, f);
  return 0;
}
int yo_tri_forces_f64(const uint32_t* F, uint32_t nF, const float* x0, uint32_t nV, double kstretch,
                      double kbend, double kair, double kd, const double* x, const double* v, const double* ext,
                      double* f) {
  TriAdj A;
  if (!tri_adjacency(F, nF, x0, nV, A)) return -1;
  tri_forces<double>(A, nV, kstretch, kbend, kair, kd, x, v, ext, f);
  return 0;
}

// Integrate one FastGrid mesh in place given forces and a per-vertex contact
// CSR (offsets nV+1, records ordered by (target_obj, triangle) per vertex).
// Returns the number of vertices the world-bounds guard touched.
uint32_t yo_integrate_f32(uint32_t nV, float subh, float thickness, float* x, float* v,
                          const float* f, const float* mass, const float* fixed,
                          const uint32_t* csr_off, const ysim_narrow_collision* csr) {
  uint32_t touched = 0;
  std::vector<ContactR<float>> cs;
  for (uint32_t i = 0; i < nV; ++i) {
    V3<float> pos = ld(x, i), vel = ld(v, i);
    uint32_t b = csr_off ? csr_off[i] : 0, e = csr_off ? csr_off[i + 1] : 0;
    cs.resize(e - b);
    for (uint32_t k = b; k < e; ++k)
      cs[k - b] = {csr[k].target_obj, csr[k].triangle,
                   {csr[k].normal_distance[0], csr[k].normal_distance[1], csr[k].normal_distance[2]},
                   csr[k].normal_distance[3]};
    touched += integrate_vertex<float>(pos, vel, ld(f, i), mass[i], fixed[i], subh, thickness,
                                       cs.data(), e - b);
    st(x, i, pos);
    st(v, i, vel);
  }
  return touched;
}
uint32_t yo_integrate_f64(uint32_t nV, double subh, double thickness, double* x, double* v,
                          const double* f, const double* mass, const double* fixed,
                          const uint32_t* csr_off, const uint32_t* csr_tobj,
                          const uint32_t* csr_tri, const double* csr_nd4) {
  uint32_t touched = 0;
  std::vector<ContactR<double>> cs;
  for (uint32_t i = 0; i < nV; ++i) {
    V3<double> pos = ld(x, i), vel = ld(v, i);
    uint32_t b = csr_off ? csr_off[i] : 0, e = csr_off ? csr_off[i + 1] : 0;
    cs.resize(e - b);
    for (uint32_t k = b; k < e; ++k)
      cs[k - b] = {csr_tobj[k], csr_tri[k],
                   {csr_nd4[4 * k], csr_nd4[4 * k + 1], csr_nd4[4 * k + 2]}, csr_nd4[4 * k + 3]};
    touched += integrate_vertex<double>(pos, vel, ld(f, i), mass[i], fixed[i], subh, thickness,
                                        cs.data(), e - b);
    st(x, i, pos);
    st(v, i, vel);
  }
  return touched;
}

}  // extern "C"

// ====================================================================== scene
// The substep of Simulator::update (include/sim/simulator.hpp:3512-3815) over a
// packed scene: [build] -> refit(+enlarge) -> broad -> narrow -> CSR ->
// xPrev := x -> forces -> integrate, restricted to FastGrid cloth (dynamic) and
// Float (static) meshes.  Per-mesh LBVHs as in BVH<…,SCENE,OBJECT>::objTrees
// (include/sim/bvh.hpp:2537-2592).  Static trees are built once.
namespace {

struct OMesh {
  ysim_mesh_desc d;
  uint32_t voff, foff;
  TriAdj adj;  // TriangularCloth springs
  std::vector<uint64_t> keys;
  std::vector<uint32_t> idx;
  std::vector<ysim_bvh_node> tree;
  std::vector<int32_t> parent;
  bool built = false;
};

struct OScene {
  ysim_sim_params P;
  std::vector<OMesh> meshes;
  std::vector<float> x, xprev, v, f, mass, ext, fixed;
  std::vector<uint32_t> facets;
  uint32_t Vtot = 0, Ttot = 0;
  uint64_t step = 0;
  uint32_t anomaly = 0;
  // last substep contacts as CSR
  std::vector<uint32_t> csr_off;
  std::vector<ysim_narrow_collision> csr;
  // threshold-tie statistics of every candidate of every substep (FP32
  // decision vs FP64 on the same inputs): candidates classified, decisions
  // that differ, FP64 distances within 4 FP32 ulps of the band
  bool tie_stats = false;
  uint64_t ties[3] = {0, 0, 0};
  // double twin of the dynamics (same contacts)
  bool twin = false;
  std::vector<double> xd, vd, xprevd, fd;
  std::vector<double> twin_nd4;  // FP64 contact normals/distances, same records as csr
  std::vector<double> kspring_d;
  std::vector<float> kspring;    // optional per-spring k for mesh 0 (adjoint configs)
  double timing[YSIM_STAGE_COUNT] = {0};  // seconds per stage, accumulated
  // per-substep stage seconds of yo_scene_step (when record_times is set)
  bool record_times = false;
  std::vector<std::array<double, YSIM_STAGE_COUNT>> step_times;
  // analytic colliders: shapes, their per-shape constants, the shape of each
  // mesh (-1: none), the (cloth, shape) pairs and the previous substep's
  // pairing of each (PackedCollisionData::endAnalyticPairs: held = last fresh)
  std::vector<ysim_analytic_shape> shapes;
  std::vector<ABounds> abounds;
  std::vector<int> shape_of;
  std::vector<uint32_t> apair_q, apair_k;
  std::vector<uint8_t> held;
  bool analytic_active() const { return !shapes.empty() && P.broad_mode == 0; }
  uint32_t build_mode = 0;  // dynamic rebuilds: 0 Karras, 1 agglomerative
  float thick(const OMesh& Q) const { return P.margin > 0.f ? P.narrow_thickness : Q.d.cloth.thickness; }
};

bool is_dynamic(const ysim_mesh_desc& d) {
  return d.behavior == YSIM_BEHAVIOR_FASTGRID_CLOTH || d.behavior == YSIM_BEHAVIOR_TRIANGULAR_CLOTH;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void build_tree(OScene& S, OMesh& M) {
  uint32_t n = M.d.num_facets;
  M.keys.resize(n);
  M.idx.resize(n);
  M.tree.assign(n ? 2 * n - 1 : 0, ysim_bvh_node{});
  M.parent.assign(n ? 2 * n - 1 : 0, -1);
  if (!n) return;
  const float* x = S.x.data() + 3 * M.voff;
  const uint32_t* f = S.facets.data() + 3 * M.foff;
  float box[6];
  double t = now_s(), u;
  yo_scene_box(x, M.d.num_vertices, box);
  S.timing[YSIM_STAGE_BOX] += (u = now_s()) - t;
  yo_morton(x, f, n, box, S.P.morton_bits, M.keys.data(), M.idx.data());
  S.timing[YSIM_STAGE_MORTON] += (t = now_s()) - u;
  yo_sort_pairs(M.keys.data(), M.idx.data(), n, S.P.morton_bits == 30 ? 30 : 63);
  S.timing[YSIM_STAGE_SORT] += (u = now_s()) - t;
  if (S.build_mode == 1 && is_dynamic(M.d) && M.built)
    yo_agglomerative(M.keys.data(), M.idx.data(), n, S.P.morton_bits, M.tree.data(), M.parent.data());
  else
    yo_karras(M.keys.data(), M.idx.data(), n, S.P.morton_bits, M.tree.data(), M.parent.data());
  S.timing[YSIM_STAGE_HIERARCHY] += now_s() - u;
  M.built = true;
}

}  // namespace

extern "C" {

void* yo_scene_create(uint32_t nm, const ysim_mesh_desc* meshes, const float* x, const float* v,
                      const float* mass, const float* ext, const float* fixed,
                      const uint32_t* facets, const ysim_sim_params* P) {
  OScene* S = new OScene();
  S->P = *P;
  uint32_t vo = 0, fo = 0;
  for (uint32_t m = 0; m < nm; ++m) {
    OMesh M;
    M.d = meshes[m];
    M.voff = vo;
    M.foff = fo;
    vo += M.d.num_vertices;
    fo += M.d.num_facets;
    S->meshes.push_back(M);
  }
  S->Vtot = vo;
  S->Ttot = fo;
  S->x.assign(x, x + 3 * vo);
  S->xprev = S->x;  // scene.hpp:1048
  S->v.assign(v, v + 3 * vo);
  S->f.assign(3 * vo, 0.f);
  S->mass.assign(mass, mass + vo);
  S->ext.assign(ext, ext + 3 * vo);
  S->fixed.assign(fixed, fixed + vo);
  S->facets.assign(facets, facets + 3 * fo);
  for (auto& M : S->meshes)
    if (M.d.behavior == YSIM_BEHAVIOR_TRIANGULAR_CLOTH &&
        !tri_adjacency(S->facets.data() + 3 * M.foff, M.d.num_facets, S->x.data() + 3 * M.voff, M.d.num_vertices,
                       M.adj)) {
      delete S;
      return nullptr;
    }
  for (auto& M : S->meshes) {
    build_tree(*S, M);
    yo_refit(S->x.data() + 3 * M.voff, S->v.data() + 3 * M.voff, 0.f, 0,
             S->facets.data() + 3 * M.foff, M.d.num_facets, M.tree.data());
  }
  S->csr_off.assign(vo + 1, 0);
  S->shape_of.assign(nm, -1);
  return S;
}

void yo_scene_destroy(void* s) { delete (OScene*)s; }

// The per-call collision arguments of later substeps (the library's
// scene_set_collision): margin, radius, narrow thickness, self collision.
void yo_scene_set_collision(void* s, float margin, float radius, float narrow_thickness, uint32_t self_collision) {
  OScene* S = (OScene*)s;
  S->P.margin = margin;
  S->P.radius = radius;
  S->P.narrow_thickness = narrow_thickness;
  S->P.self_collision = self_collision;
}

// tie_bits of one candidate given as float triples (checks of the classifier).
uint32_t yo_tie_bits(const float* qc, const float* qp, const float* t0, const float* t1, const float* t2,
                     float band, int fired32) {
  auto v = [](const float* a) { return V3<float>{a[0], a[1], a[2]}; };
  return tie_bits(v(qc), v(qp), v(t0), v(t1), v(t2), band, fired32 != 0);
}

// Threshold-tie statistics from now on (each substep's candidates): out3 =
// {candidates classified, FP32/FP64 decision differences, band ties}.
void yo_scene_enable_ties(void* s) { ((OScene*)s)->tie_stats = true; }
void yo_scene_ties(void* s, uint64_t* out3) {
  OScene* S = (OScene*)s;
  for (int k = 0; k < 3; ++k) out3[k] = S->ties[k];
}

void yo_scene_enable_twin(void* s) {
  OScene* S = (OScene*)s;
  S->twin = true;
  S->xd.assign(S->x.begin(), S->x.end());
  S->vd.assign(S->v.begin(), S->v.end());
  S->xprevd.assign(S->xprev.begin(), S->xprev.end());
  S->fd.assign(S->f.size(), 0.0);
  S->kspring_d.assign(S->kspring.begin(), S->kspring.end());
}

// Per-spring stiffness for mesh `mesh` (SpringIndex order); NULL clears it.
void yo_scene_set_kspring(void* s, const float* k, uint32_t n) {
  OScene* S = (OScene*)s;
  if (!k) { S->kspring.clear(); S->kspring_d.clear(); return; }
  S->kspring.assign(k, k + n);
  S->kspring_d.assign(k, k + n);
}

// Analytic branch of one broad detection + narrowAnalytic
// (bvh.hpp:2897-2949, scene.hpp:511-529, bruteforce.hpp:256-308): a pair
// (cloth q, shape k) is marked when q's tree root box meets the shape's swept
// box padded by q's band plus the rotation band (closed intervals,
// AABB4::intersect, bvh.hpp:66-71; a plane always), and it runs when marked
// now or at the previous detection.  Each vertex of a running pair tests the
// shape with band + rotMargin and keeps one contact (shape mesh, facet 0).
static float analytic_band(const OScene& S, const OMesh& Q, uint32_t k) {
  const float th = S.thick(Q);
  return (S.P.radius + th) + rot_margin(S.abounds[k], S.P.radius, th);
}
static void analytic_contacts(OScene& S, std::vector<Contact>& contacts) {
  std::vector<uint8_t> fresh(S.apair_q.size(), 0);
  for (size_t i = 0; i < S.apair_q.size(); ++i) {
    const uint32_t q = S.apair_q[i], k = S.apair_k[i];
    const OMesh& Q = S.meshes[q];
    const ysim_analytic_shape& sh = S.shapes[k];
    const ABounds& B = S.abounds[k];
    bool hit = true;
    if (sh.kind != YSIM_COLLIDER_PLANE) {
      const float pad = (S.P.radius + S.thick(Q)) + B.rot_band;
      const ysim_bvh_node& R = Q.tree[0];
      hit = true;
      for (int c = 0; c < 3; ++c)
        if (B.hi[c] + pad < R.min[c] || B.lo[c] - pad > R.max[c]) hit = false;
    }
    fresh[i] = hit ? 1 : 0;
    if (!(hit || S.held[i]) || !(sh.flags & 1u)) continue;
    const float band = analytic_band(S, Q, k);
    const float* xq = S.x.data() + 3 * Q.voff;
    const float* xqp = S.xprev.data() + 3 * Q.voff;
    for (uint32_t p = 0; p < Q.d.num_vertices; ++p) {
      V3<float> n;
      float d;
      if (!narrow_analytic<float>(sh, ld(xq, p), ld(xqp, p), band, n, d)) continue;
      Contact c;
      c.gv = Q.voff + p; c.tobj = sh.obj_index; c.tri = 0; c.qobj = q; c.point = p;
      c.n[0] = n.x; c.n[1] = n.y; c.n[2] = n.z; c.d = d;
      c.qbeh = Q.d.behavior; c.tbeh = sh.behavior_type; c.qshape = Q.d.shape; c.tshape = sh.kind;
      contacts.push_back(c);
    }
  }
  S.held.swap(fresh);
}

// Collision pipeline at the current state: returns the per-vertex contact CSR
// (into S->csr_off / S->csr) and, optionally, the broad pair count.
static void collide(OScene& S, uint64_t* npairs_out, std::vector<uint32_t>* pairs_out) {
  const ysim_sim_params& P = S.P;
  bool rebuild = (P.rebuild_every > 0) && (S.step % P.rebuild_every == 0);
  for (auto& M : S.meshes) {
    if (!is_dynamic(M.d)) continue;
    if (rebuild || !M.built) build_tree(S, M);
    const double t = now_s();
    yo_refit(S.x.data() + 3 * M.voff, S.v.data() + 3 * M.voff, P.subh, P.swept,
             S.facets.data() + 3 * M.foff, M.d.num_facets, M.tree.data());
    S.timing[YSIM_STAGE_REFIT] += now_s() - t;
  }
  const double t_query = now_s();
  std::vector<Contact> contacts;
  uint64_t npairs = 0;
  for (uint32_t q = 0; q < S.meshes.size(); ++q) {
    OMesh& Q = S.meshes[q];
    if (!is_dynamic(Q.d) || !Q.d.collidable) continue;
    // margin <= 0 selects a per-query-mesh margin and band equal to the cloth's
    // thickness (batched instances of different resolutions).
    const float margin = P.margin > 0.f ? P.margin : Q.d.cloth.thickness;
    const float band = P.radius + (P.margin > 0.f ? P.narrow_thickness : Q.d.cloth.thickness);
    for (uint32_t t = 0; t < S.meshes.size(); ++t) {
      OMesh& T = S.meshes[t];
      if (!T.d.collidable || T.d.group != Q.d.group || T.d.num_facets == 0) continue;
      if (q == t && !P.self_collision) continue;
      if (S.analytic_active() && S.shape_of[t] >= 0) continue;  // analyticSkip, bvh.hpp:2520-2523
      const float* xq = S.x.data() + 3 * Q.voff;
      const float* xqp = S.xprev.data() + 3 * Q.voff;
      const float* xt = S.x.data() + 3 * T.voff;
      const uint32_t* tf = S.facets.data() + 3 * T.foff;
      uint32_t nq = Q.d.num_vertices;
      std::vector<std::vector<Contact>> per(nq);
      std::vector<uint64_t> pc(nq, 0);
      std::vector<std::vector<uint32_t>> pp(pairs_out ? nq : 0);
#pragma omp parallel for schedule(dynamic, 256)
      for (int64_t p_ = 0; p_ < (int64_t)nq; ++p_) {
        uint32_t p = (uint32_t)p_;
        float qmin[3], qmax[3];
        for (int k = 0; k < 3; ++k) { qmin[k] = xq[3 * p + k] - margin; qmax[k] = xq[3 * p + k] + margin; }
        std::vector<int> stack;
        stack.reserve(64);
        stack.push_back(0);
        while (!stack.empty()) {
          int nd = stack.back();
          stack.pop_back();
          const ysim_bvh_node& N = T.tree[nd];
          if (!box_overlap(qmin, qmax, N)) continue;
          if (N.child_a < 0) {
            uint32_t fid = (uint32_t)N.child_b;
            const uint32_t* f = tf + 3 * fid;
            if (q == t && (p == f[0] || p == f[1] || p == f[2])) continue;
            pc[p]++;
            if (pairs_out) { pp[p].push_back(p); pp[p].push_back(fid); }
            V3<float> n;
            float d;
            const bool fired = narrow_test(ld(xq, p), ld(xqp, p), ld(xt, f[0]), ld(xt, f[1]), ld(xt, f[2]), band, n, d);
            if (S.tie_stats) {
              const uint32_t tb = tie_bits(ld(xq, p), ld(xqp, p), ld(xt, f[0]), ld(xt, f[1]), ld(xt, f[2]), band, fired);
#pragma omp atomic
              S.ties[0]++;
              if (tb & 1u) {
#pragma omp atomic
                S.ties[1]++;
              }
              if (tb & 2u) {
#pragma omp atomic
                S.ties[2]++;
              }
            }
            if (fired) {
              Contact c;
              c.gv = Q.voff + p; c.tobj = t; c.tri = fid; c.qobj = q; c.point = p;
              c.n[0] = n.x; c.n[1] = n.y; c.n[2] = n.z; c.d = d;
              c.qbeh = Q.d.behavior; c.tbeh = T.d.behavior; c.qshape = Q.d.shape; c.tshape = T.d.shape;
              per[p].push_back(c);
            }
            continue;
          }
          stack.push_back(N.child_b);
          stack.push_back(N.child_a);
        }
      }
      for (uint32_t p = 0; p < nq; ++p) {
        npairs += pc[p];
        for (auto& c : per[p]) contacts.push_back(c);
        if (pairs_out) {
          for (size_t k = 0; k < pp[p].size(); k += 2) {
            pairs_out->push_back(q); pairs_out->push_back(pp[p][k]);
            pairs_out->push_back(t); pairs_out->push_back(pp[p][k + 1]);
          }
        }
      }
    }
  }
  if (S.analytic_active()) analytic_contacts(S, contacts);
  // CSR ordered by (vertex, target_obj, triangle): bruteforce.hpp:358-383 with
  // the deterministic in-bucket order of this build.
  std::sort(contacts.begin(), contacts.end(), [](const Contact& a, const Contact& b) {
    if (a.gv != b.gv) return a.gv < b.gv;
    if (a.tobj != b.tobj) return a.tobj < b.tobj;
    return a.tri < b.tri;
  });
  S.csr_off.assign(S.Vtot + 1, 0);
  S.csr.resize(contacts.size());
  for (size_t i = 0; i < contacts.size(); ++i) {
    const Contact& c = contacts[i];
    S.csr_off[c.gv + 1]++;
    ysim_narrow_collision& r = S.csr[i];
    r.point = c.point; r.triangle = c.tri; r.query_obj = c.qobj; r.target_obj = c.tobj;
    r.normal_distance[0] = c.n[0]; r.normal_distance[1] = c.n[1]; r.normal_distance[2] = c.n[2];
    r.normal_distance[3] = c.d;
    r.query_behavior = c.qbeh; r.target_behavior = c.tbeh; r.query_shape = c.qshape; r.target_shape = c.tshape;
  }
  for (uint32_t i = 0; i < S.Vtot; ++i) S.csr_off[i + 1] += S.csr_off[i];
  if (npairs_out) *npairs_out = npairs;
  S.timing[YSIM_STAGE_QUERY] += now_s() - t_query;
}

// FP64 twin of the contact geometry: recompute each contact record of the FP32
// CSR with double positions (same pair set), for error bars only.
static void twin_contacts(OScene& S) {
  S.twin_nd4.assign(4 * S.csr.size(), 0.0);
  for (size_t i = 0; i < S.csr.size(); ++i) {
    const ysim_narrow_collision& r = S.csr[i];
    const double band = 1.0;  // the FP32 decision stands; FP64 only re-evaluates the geometry
    const OMesh& Q = S.meshes[r.query_obj];
    if (S.analytic_active() && S.shape_of[r.target_obj] >= 0) {
      const uint32_t k = (uint32_t)S.shape_of[r.target_obj];
      V3<double> n;
      double d;
      const size_t gv = Q.voff + r.point;
      if (narrow_analytic<double>(S.shapes[k], ld(S.xd.data(), gv), ld(S.xprevd.data(), gv),
                                  (double)analytic_band(S, Q, k), n, d)) {
        S.twin_nd4[4 * i] = n.x; S.twin_nd4[4 * i + 1] = n.y; S.twin_nd4[4 * i + 2] = n.z;
        S.twin_nd4[4 * i + 3] = d;
      } else {
        for (int c = 0; c < 4; ++c) S.twin_nd4[4 * i + c] = r.normal_distance[c];
      }
      continue;
    }
    const OMesh& T = S.meshes[r.target_obj];
    const uint32_t* f = S.facets.data() + 3 * (T.foff + r.triangle);
    V3<double> n;
    double d;
    const double* xq = S.xd.data() + 3 * Q.voff;
    const double* xqp = S.xprevd.data() + 3 * Q.voff;
    const double* xt = S.xd.data() + 3 * T.voff;
    if (narrow_test(ld(xq, r.point), ld(xqp, r.point), ld(xt, f[0]), ld(xt, f[1]), ld(xt, f[2]),
                    band, n, d)) {
      S.twin_nd4[4 * i] = n.x; S.twin_nd4[4 * i + 1] = n.y; S.twin_nd4[4 * i + 2] = n.z;
      S.twin_nd4[4 * i + 3] = d;
    } else {
      // geometry rejected in FP64 (a tie): keep the FP32 record's values
      for (int k = 0; k < 4; ++k) S.twin_nd4[4 * i + k] = r.normal_distance[k];
    }
  }
}

static void dynamics(OScene& S) {
  const ysim_sim_params& P = S.P;
  // snapshot_xprev_range (physics.metal:520-529): xPrev := x for dynamic meshes
  for (auto& M : S.meshes) {
    if (!is_dynamic(M.d)) continue;
    std::memcpy(S.xprev.data() + 3 * M.voff, S.x.data() + 3 * M.voff, 12 * (size_t)M.d.num_vertices);
    if (S.twin)
      std::memcpy(S.xprevd.data() + 3 * M.voff, S.xd.data() + 3 * M.voff, 24 * (size_t)M.d.num_vertices);
  }
  bool first_cloth = true;
  for (auto& M : S.meshes) {
    if (!is_dynamic(M.d)) continue;
    const float* ks = (first_cloth && !S.kspring.empty()) ? S.kspring.data() : nullptr;
    const double* ksd = (first_cloth && !S.kspring_d.empty()) ? S.kspring_d.data() : nullptr;
    first_cloth = false;
    size_t o = M.voff;
    if (M.d.behavior == YSIM_BEHAVIOR_TRIANGULAR_CLOTH) {
      tri_forces<float>(M.adj, M.d.num_vertices, M.d.cloth.kstretch, M.d.cloth.kbend, P.kair, P.kd,
                        S.x.data() + 3 * o, S.v.data() + 3 * o, S.ext.data() + 3 * o, S.f.data() + 3 * o);
      if (S.twin) {
        std::vector<double> extd(S.ext.begin() + 3 * o, S.ext.begin() + 3 * (o + M.d.num_vertices));
        tri_forces<double>(M.adj, M.d.num_vertices, (double)M.d.cloth.kstretch, (double)M.d.cloth.kbend,
                           (double)P.kair, (double)P.kd, S.xd.data() + 3 * o, S.vd.data() + 3 * o, extd.data(),
                           S.fd.data() + 3 * o);
      }
      continue;
    }
    fastgrid_forces<float>(M.d.cloth, M.d.num_vertices, P.kair, P.kd, S.x.data() + 3 * o,
                           S.v.data() + 3 * o, S.ext.data() + 3 * o, ks, S.f.data() + 3 * o);
    if (S.twin) {
      std::vector<double> extd(S.ext.begin() + 3 * o, S.ext.begin() + 3 * (o + M.d.num_vertices));
      fastgrid_forces<double>(M.d.cloth, M.d.num_vertices, (double)P.kair, (double)P.kd,
                              S.xd.data() + 3 * o, S.vd.data() + 3 * o, extd.data(), ksd,
                              S.fd.data() + 3 * o);
    }
  }
  std::vector<ContactR<float>> cs;
  std::vector<ContactR<double>> csd;
  for (auto& M : S.meshes) {
    if (!is_dynamic(M.d)) continue;
    for (uint32_t i = 0; i < M.d.num_vertices; ++i) {
      uint32_t g = M.voff + i;
      uint32_t b = S.csr_off[g], e = S.csr_off[g + 1];
      cs.resize(e - b);
      for (uint32_t k = b; k < e; ++k)
        cs[k - b] = {S.csr[k].target_obj, S.csr[k].triangle,
                     {S.csr[k].normal_distance[0], S.csr[k].normal_distance[1], S.csr[k].normal_distance[2]},
                     S.csr[k].normal_distance[3]};
      V3<float> pos = ld(S.x.data(), g), vel = ld(S.v.data(), g);
      if (integrate_vertex<float>(pos, vel, ld(S.f.data(), g), S.mass[g], S.fixed[g], P.subh,
                                  M.d.cloth.thickness, cs.data(), e - b))
        S.anomaly |= 1u;
      st(S.x.data(), g, pos);
      st(S.v.data(), g, vel);
      if (S.twin) {
        csd.resize(e - b);
        for (uint32_t k = b; k < e; ++k)
          csd[k - b] = {S.csr[k].target_obj, S.csr[k].triangle,
                        {S.twin_nd4[4 * k], S.twin_nd4[4 * k + 1], S.twin_nd4[4 * k + 2]},
                        S.twin_nd4[4 * k + 3]};
        V3<double> pd = ld(S.xd.data(), g), vd = ld(S.vd.data(), g);
        integrate_vertex<double>(pd, vd, ld(S.fd.data(), g), (double)S.mass[g], (double)S.fixed[g],
                                 (double)P.subh, (double)M.d.cloth.thickness, csd.data(), e - b);
        st(S.xd.data(), g, pd);
        st(S.vd.data(), g, vd);
      }
    }
  }
}

void yo_scene_step(void* s, uint32_t n) {
  OScene* S = (OScene*)s;
  for (uint32_t i = 0; i < n; ++i) {
    std::array<double, YSIM_STAGE_COUNT> t0;
    std::copy(S->timing, S->timing + YSIM_STAGE_COUNT, t0.begin());
    collide(*S, nullptr, nullptr);
    if (S->twin) twin_contacts(*S);
    const double t = now_s();
    dynamics(*S);
    S->timing[YSIM_STAGE_INTEGRATE] += now_s() - t;
    S->step++;
    if (S->record_times) {
      std::array<double, YSIM_STAGE_COUNT> d;
      for (int k = 0; k < YSIM_STAGE_COUNT; ++k) d[k] = S->timing[k] - t0[k];
      S->step_times.push_back(d);
    }
  }
}

// Per-substep stage times (seconds, YSIM_STAGE_COUNT per substep) of the
// substeps run since record_times was switched on; returns the substep count
// and copies up to cap substeps.  on = 0 stops recording and clears.
void yo_scene_record_times(void* s, int on) {
  OScene* S = (OScene*)s;
  S->record_times = on != 0;
  S->step_times.clear();
}
uint32_t yo_scene_step_times(void* s, double* out, uint32_t cap) {
  OScene* S = (OScene*)s;
  const uint32_t n = (uint32_t)S->step_times.size();
  for (uint32_t i = 0; i < n && i < cap; ++i)
    std::copy(S->step_times[i].begin(), S->step_times[i].end(), out + (size_t)YSIM_STAGE_COUNT * i);
  return n;
}

// Collision pipeline only (no step): candidate pairs as (qobj, point, tobj, tri)
// quadruples.  Returns the pair count; writes up to cap quadruples.
uint64_t yo_scene_pairs(void* s, uint32_t* out4, uint64_t cap) {
  OScene* S = (OScene*)s;
  std::vector<uint32_t> pairs;
  uint64_t np = 0;
  collide(*S, &np, &pairs);
  if (out4) {
    uint64_t n = std::min<uint64_t>(cap, pairs.size() / 4);
    std::memcpy(out4, pairs.data(), 16 * n);
  }
  return np;
}

void yo_scene_get_state(void* s, float* x, float* v, float* xprev) {
  OScene* S = (OScene*)s;
  if (x) std::memcpy(x, S->x.data(), 4 * S->x.size());
  if (v) std::memcpy(v, S->v.data(), 4 * S->v.size());
  if (xprev) std::memcpy(xprev, S->xprev.data(), 4 * S->xprev.size());
}
void yo_scene_set_state(void* s, const float* x, const float* v, const float* xprev) {
  OScene* S = (OScene*)s;
  if (x) std::memcpy(S->x.data(), x, 4 * S->x.size());
  if (v) std::memcpy(S->v.data(), v, 4 * S->v.size());
  if (xprev) std::memcpy(S->xprev.data(), xprev, 4 * S->xprev.size());
}
void yo_scene_get_twin(void* s, double* x, double* v) {
  OScene* S = (OScene*)s;
  if (x) std::memcpy(x, S->xd.data(), 8 * S->xd.size());
  if (v) std::memcpy(v, S->vd.data(), 8 * S->vd.size());
}
uint32_t yo_scene_num_contacts(void* s) { return (uint32_t)((OScene*)s)->csr.size(); }
void yo_scene_get_contacts(void* s, uint32_t* off, ysim_narrow_collision* out) {
  OScene* S = (OScene*)s;
  if (off) std::memcpy(off, S->csr_off.data(), 4 * S->csr_off.size());
  if (out) std::memcpy(out, S->csr.data(), sizeof(ysim_narrow_collision) * S->csr.size());
}
void yo_scene_get_tree(void* s, uint32_t mesh, ysim_bvh_node* nodes, int32_t* parent,
                       uint64_t* keys, uint32_t* idx) {
  OScene* S = (OScene*)s;
  OMesh& M = S->meshes[mesh];
  if (nodes) std::memcpy(nodes, M.tree.data(), sizeof(ysim_bvh_node) * M.tree.size());
  if (parent) std::memcpy(parent, M.parent.data(), 4 * M.parent.size());
  if (keys) std::memcpy(keys, M.keys.data(), 8 * M.keys.size());
  if (idx) std::memcpy(idx, M.idx.data(), 4 * M.idx.size());
}
uint32_t yo_scene_anomaly(void* s) { return ((OScene*)s)->anomaly; }

// Builder of later dynamic rebuilds (the library's tuning build_mode): 0
// Karras, 1 agglomerative.  The trees built at creation are Karras.
void yo_scene_set_build_mode(void* s, uint32_t mode) { ((OScene*)s)->build_mode = mode; }

// Install the analytic colliders (same contract as the library's
// scene_set_analytic).  Returns 0, or -1 with the previous set unchanged.
int yo_scene_set_analytic(void* s, const ysim_analytic_shape* shapes, uint32_t n) {
  OScene* S = (OScene*)s;
  const uint32_t nm = (uint32_t)S->meshes.size();
  std::vector<int> shape_of(nm, -1);
  for (uint32_t k = 0; k < n; ++k) {
    const ysim_analytic_shape& sh = shapes[k];
    if (sh.kind < YSIM_COLLIDER_SPHERE || sh.kind > YSIM_COLLIDER_PLANE || sh.obj_index >= nm) return -1;
    const ysim_mesh_desc& d = S->meshes[sh.obj_index].d;
    if (d.behavior != YSIM_BEHAVIOR_FLOAT || d.shape != sh.kind || sh.behavior_type != d.behavior) return -1;
    if (shape_of[sh.obj_index] >= 0) return -1;
    shape_of[sh.obj_index] = (int)k;
  }
  bool same = n == S->shapes.size();
  for (uint32_t k = 0; same && k < n; ++k) same = shapes[k].obj_index == S->shapes[k].obj_index;
  S->shapes.assign(shapes, shapes + n);
  S->shape_of = shape_of;
  S->abounds.clear();
  for (uint32_t k = 0; k < n; ++k) S->abounds.push_back(analytic_bounds(shapes[k]));
  if (!same) {
    S->apair_q.clear();
    S->apair_k.clear();
    for (uint32_t q = 0; q < nm; ++q) {
      const ysim_mesh_desc& Q = S->meshes[q].d;
      if (!is_dynamic(Q) || !Q.collidable) continue;
      for (uint32_t k = 0; k < n; ++k) {
        const ysim_mesh_desc& T = S->meshes[shapes[k].obj_index].d;
        if (!T.collidable || T.group != Q.group) continue;
        S->apair_q.push_back(q);
        S->apair_k.push_back(k);
      }
    }
    S->held.assign(S->apair_q.size(), 0);
  }
  return 0;
}
// The (cloth, shape mesh) pairs and their pairing at the last detection.
uint32_t yo_scene_analytic_markers(void* s, uint32_t* pairs2, uint8_t* marked) {
  OScene* S = (OScene*)s;
  const uint32_t n = (uint32_t)S->apair_q.size();
  for (uint32_t i = 0; i < n; ++i) {
    if (pairs2) { pairs2[2 * i] = S->apair_q[i]; pairs2[2 * i + 1] = S->shapes[S->apair_k[i]].obj_index; }
    if (marked) marked[i] = S->held[i];
  }
  return n;
}
int yo_narrow_analytic_f32(const ysim_analytic_shape* sh, const float* pcur, const float* pprev, float band,
                           float* out4) {
  V3<float> n;
  float d;
  if (!narrow_analytic<float>(*sh, ld(pcur, 0), ld(pprev, 0), band, n, d)) return 0;
  out4[0] = n.x; out4[1] = n.y; out4[2] = n.z; out4[3] = d;
  return 1;
}
int yo_narrow_analytic_f64(const ysim_analytic_shape* sh, const double* pcur, const double* pprev, double band,
                           double* out4) {
  V3<double> n;
  double d;
  if (!narrow_analytic<double>(*sh, ld(pcur, 0), ld(pprev, 0), band, n, d)) return 0;
  out4[0] = n.x; out4[1] = n.y; out4[2] = n.z; out4[3] = d;
  return 1;
}
// Per-shape constants: lo3, hi3 (unpadded swept box), out3 = (rot_band,
// rot_margin(radius, thickness)) as floats, theta2 = (narrow theta, rmax).
void yo_analytic_bounds(const ysim_analytic_shape* sh, float radius, float thickness, float* lo3, float* hi3,
                        float* out2, double* theta2) {
  const ABounds B = analytic_bounds(*sh);
  for (int k = 0; k < 3; ++k) { lo3[k] = B.lo[k]; hi3[k] = B.hi[k]; }
  out2[0] = B.rot_band;
  out2[1] = rot_margin(B, radius, thickness);
  theta2[0] = B.theta;
  theta2[1] = B.rmax;
}

}  // extern "C"
