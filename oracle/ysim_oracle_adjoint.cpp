// ysim_oracle_adjoint.cpp — CPU rollout of one FastGrid cloth against static
// triangle obstacles and the hand-derived adjoint of its substep, in FP32 (the
// parity path, same operation order as the HIP adjoint) and FP64 (checked
// against central finite differences).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of the HIP adjoint;
// never linked by the product path.
//
// The reference has no differentiation.  The differentiated function is its
// substep for this case (Simulator::update, include/sim/simulator.hpp:3512-3815):
// contacts from (x, xPrev) by the swept point-triangle test of
// src/metal/bruteforce.metal:41-178 against the obstacle triangles whose box
// meets [x - margin, x + margin]; xPrev := x; the 12-spring FastGrid gather of
// src/metal/physics.metal:171-224 with one stiffness per spring; then
// integrate_cloth_grid (physics.metal:226-312):
//   f   = ext + kair v + sum_s F_s,   F_s = (k_s (len - r_s) + kd dv.n) n
//   u   = v + ((h f) / m) mask
//   per contact c with d_c < thickness, in (target mesh, facet) order:
//         if u.n_c < 0: u <- u - (u.n_c) n_c          (gate recorded)
//         p <- p + (thickness - d_c) n_c,  d_c = n_c.(x - t0_c)
//   x'  = p + (h u) mask,   v' = u
// The contact set and both gates are held fixed (the derivative of the
// piecewise-smooth map on the piece the trajectory lies in).  Obstacles are
// static, so n_c and t0_c are constants.
//
// Adjoint of one step (x_bar', v_bar' -> x_bar, v_bar):
//   u_bar = v_bar' + (h mask) x_bar'
//   x_bar = x_bar' ; for active c in reverse order:
//           x_bar -= n_c (n_c . x_bar');  if gated: u_bar -= (u_bar . n_c) n_c
//   f_bar = u_bar ((h mask) / m);   v_bar = u_bar + kair f_bar
//   per spring (i, j), g = f_bar_i - f_bar_j, M = (I - n n^T) / len:
//     x_bar_i -= (k n + kd M dv)(n.g) + s M g ;   v_bar_i -= kd (n.g) n
//     (j receives the opposite terms in its own gather)
//     k_bar_s += (len - r)(n.g),   kd_bar += (dv.n)(n.g)   (once per spring)
// and dL/dlog k_s = k_bar_s k_s.  Checkpoints (x, v, xPrev) every K steps; the
// backward pass recomputes each segment from its checkpoint.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/ysim_hip.h"

namespace {

template <class R>
struct Vec {
  R x, y, z;
};
template <class R> inline Vec<R> operator+(Vec<R> a, Vec<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline Vec<R> operator-(Vec<R> a, Vec<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline Vec<R> operator*(Vec<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline Vec<R> operator*(R s, Vec<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline R dotv(Vec<R> a, Vec<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline Vec<R> crossv(Vec<R> a, Vec<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline Vec<R> ldv(const std::vector<R>& a, size_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
template <class R> inline void stv(std::vector<R>& a, size_t i, Vec<R> v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }

// pointInTriangleBary, src/metal/bruteforce.metal:9-28.
template <class R>
bool in_tri(Vec<R> p, Vec<R> v0, Vec<R> v1) {
  R d00 = dotv(v0, v0), d01 = dotv(v0, v1), d11 = dotv(v1, v1), d20 = dotv(p, v0), d21 = dotv(p, v1);
  R den = d00 * d11 - d01 * d01;
  if (std::fabs(den) < (R)1e-12f) return false;
  R b = (d11 * d20 - d01 * d21) / den, c = (d00 * d21 - d01 * d20) / den, a = (R)1 - b - c;
  return a >= 0 && b >= 0 && c >= 0;
}

template <class R>
struct Contact {
  uint32_t tobj, tri;
  Vec<R> n;   // unit normal as applied by the integrator
  R d;        // signed distance of x at detection
  Vec<R> t0;  // first vertex of the facet (d = n.(x - t0))
  char act = 0, proj = 0;  // recorded gates (frozen mode)
};

// Spring of vertex i towards neighbour j in the gather order of
// compute_cloth_grid_forces_fast (physics.metal:195-218); s is the spring index
// (families struct-x, struct-y, shear-a, shear-b, bend-x, bend-y, row-major).
struct SpringRef {
  uint32_t j, s, fam;
};

struct Grid {
  uint32_t N = 0, SY = 0, SA = 0, SB = 0, BX = 0, BY = 0, total = 0;
  void init(uint32_t n) {
    N = n;
    SY = N * (N - 1);
    SA = SY + (N - 1) * N;
    SB = SA + (N - 1) * (N - 1);
    BX = SB + (N - 1) * (N - 1);
    BY = BX + N * (N - 2);
    total = BY + (N - 2) * N;
  }
  int springs(uint32_t i, SpringRef out[12]) const {
    const uint32_t r = i / N, c = i % N;
    int k = 0;
    auto add = [&](uint32_t rr, uint32_t cc, uint32_t s, uint32_t fam) { out[k++] = {rr * N + cc, s, fam}; };
    if (c > 0) add(r, c - 1, r * (N - 1) + c - 1, 0);
    if (c < N - 1) add(r, c + 1, r * (N - 1) + c, 0);
    if (r > 0) add(r - 1, c, SY + (r - 1) * N + c, 1);
    if (r < N - 1) add(r + 1, c, SY + r * N + c, 1);
    if (c > 0 && r > 0) add(r - 1, c - 1, SA + (r - 1) * (N - 1) + c - 1, 2);
    if (c < N - 1 && r > 0) add(r - 1, c + 1, SB + (r - 1) * (N - 1) + c, 3);
    if (c > 0 && r < N - 1) add(r + 1, c - 1, SB + r * (N - 1) + c - 1, 3);
    if (c < N - 1 && r < N - 1) add(r + 1, c + 1, SA + r * (N - 1) + c, 2);
    if (c > 1) add(r, c - 2, BX + r * (N - 2) + c - 2, 4);
    if (c < N - 2) add(r, c + 2, BX + r * (N - 2) + c, 4);
    if (r > 1) add(r - 2, c, BY + (r - 2) * N + c, 5);
    if (r < N - 2) add(r + 2, c, BY + r * N + c, 5);
    return k;
  }
};

template <class R>
struct Rollout {
  Grid G;
  uint32_t V = 0;
  R rest[6];
  R h, kair, kd, thick, margin, band;
  std::vector<R> mass, fixed, ext;
  std::vector<R> obs;  // 9 coordinates per obstacle triangle
  std::vector<uint32_t> obs_obj, obs_fid;
  std::vector<R> obs_box;  // 6 per obstacle triangle
  std::vector<R> k;        // per-spring stiffness
  std::vector<R> x0, v0;
  std::vector<R> x, v, xprev;
  uint32_t K = 10, steps = 0;
  std::vector<std::vector<R>> ck_x, ck_v, ck_xp;
  uint64_t contact_sig = 0;   // order-independent signature of (step, vertex, facet, gates)
  uint64_t contact_count = 0;
  // mode 0: detect; 1: detect and record every step's contacts with their gates;
  // 2: frozen — replay the recorded contacts and gates, d = n.(x - t0) from the
  // current x (the smooth map the adjoint differentiates)
  int mode = 0;
  std::vector<std::vector<std::vector<Contact<R>>>> rec;
  // planted mistakes, to show the finite-difference check has the power to see
  // them: bit0 drops the contact push's pull-back on x_bar, bit1 drops the
  // velocity projection's, bit2 drops the damping term of the spring Jacobian,
  // bit3 flips the sign of the stiffness gradient
  uint32_t plant = 0;

  void detect(std::vector<std::vector<Contact<R>>>& C, const std::vector<R>& xc, const std::vector<R>& xp) const {
    C.assign(V, {});
    const uint32_t T = (uint32_t)obs_fid.size();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i_ = 0; i_ < (int64_t)V; ++i_) {
      const uint32_t i = (uint32_t)i_;
      const Vec<R> q = ldv(xc, i), qp = ldv(xp, i);
      const R lo[3] = {q.x - margin, q.y - margin, q.z - margin}, hi[3] = {q.x + margin, q.y + margin, q.z + margin};
      std::vector<Contact<R>>& out = C[i];
      for (uint32_t t = 0; t < T; ++t) {
        const R* B = &obs_box[6 * (size_t)t];
        if (hi[0] < B[0] || lo[0] > B[3] || hi[1] < B[1] || lo[1] > B[4] || hi[2] < B[2] || lo[2] > B[5]) continue;
        const R* o = &obs[9 * (size_t)t];
        const Vec<R> t0 = {o[0], o[1], o[2]}, t1 = {o[3], o[4], o[5]}, t2 = {o[6], o[7], o[8]};
        const Vec<R> e0 = t1 - t0, e1 = t2 - t0;
        Vec<R> n = crossv(e0, e1);
        const R nl2 = dotv(n, n);
        if (nl2 < (R)1e-12f) continue;
        n = n * ((R)1 / std::sqrt(nl2));
        R dp = dotv(n, qp - t0), dc = dotv(n, q - t0);
        const R dref = dp != 0 ? dp : dc;
        if (dref < 0) { n = {-n.x, -n.y, -n.z}; dp = -dp; dc = -dc; }
        const bool crossed = dc < 0, inm = dc < band;
        if (!crossed && !inm) continue;
        const Vec<R> cp = (crossed && dp > 0) ? qp + (q - qp) * (dp / (dp - dc)) : q;
        const Vec<R> pc = cp - t0;
        if (!in_tri(pc - n * dotv(n, pc), e0, e1)) continue;
        // the integrator re-normalises the stored normal (physics.metal:273-276)
        const R l2 = dotv(n, n);
        if (l2 < (R)1e-12f) continue;
        n = n * ((R)1 / std::sqrt(l2));
        out.push_back({obs_obj[t], obs_fid[t], n, dc, t0, 0, 0});
      }
      std::sort(out.begin(), out.end(), [](const Contact<R>& a, const Contact<R>& b) {
        return a.tobj != b.tobj ? a.tobj < b.tobj : a.tri < b.tri;
      });
    }
  }

  void forces(const std::vector<R>& xc, const std::vector<R>& vc, std::vector<R>& f) const {
    f.assign(3 * (size_t)V, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i_ = 0; i_ < (int64_t)V; ++i_) {
      const uint32_t i = (uint32_t)i_;
      const Vec<R> p = ldv(xc, i), vel = ldv(vc, i);
      Vec<R> F = ldv(ext, i) + vel * kair;
      SpringRef sr[12];
      const int ns = G.springs(i, sr);
      for (int q = 0; q < ns; ++q) {
        const Vec<R> dx = ldv(xc, sr[q].j) - p, dv = ldv(vc, sr[q].j) - vel;
        const R len = std::sqrt(dotv(dx, dx));
        if (len < (R)1e-7f) continue;
        const Vec<R> nd = {dx.x / len, dx.y / len, dx.z / len};
        const R s = k[sr[q].s] * (len - rest[sr[q].fam]) + kd * dotv(dv, nd);
        F = F + s * nd;
      }
      stv(f, i, F);
    }
  }

  // u after the force update and the contact sequence; act/proj receive the
  // gates (taken from the records when `frozen`)
  Vec<R> replay(uint32_t i, const std::vector<R>& vc, const std::vector<R>& f,
                const std::vector<Contact<R>>& cs, std::vector<char>& act, std::vector<char>& proj,
                bool frozen = false) const {
    const Vec<R> F = ldv(f, i), v_ = ldv(vc, i);
    const R m = mass[i], mk = fixed[i];
    Vec<R> u = {v_.x + ((h * F.x) / m) * mk, v_.y + ((h * F.y) / m) * mk, v_.z + ((h * F.z) / m) * mk};
    act.assign(cs.size(), 0);
    proj.assign(cs.size(), 0);
    for (size_t c = 0; c < cs.size(); ++c) {
      const R vn = dotv(u, cs[c].n);
      if (frozen ? cs[c].act != 0 : cs[c].d < thick) {
        act[c] = 1;
        if (frozen ? cs[c].proj != 0 : vn < 0) { u = u - vn * cs[c].n; proj[c] = 1; }
      }
    }
    return u;
  }

  void step(std::vector<std::vector<Contact<R>>>& C, uint32_t n) {
    const bool frozen = mode == 2;
    if (frozen) {
      C = rec[n];
      for (uint32_t i = 0; i < V; ++i)
        for (Contact<R>& c : C[i]) c.d = dotv(c.n, ldv(x, i) - c.t0);
    } else {
      detect(C, x, xprev);
    }
    std::vector<R> f;
    forces(x, v, f);
    xprev = x;
    std::vector<char> act, proj;
    for (uint32_t i = 0; i < V; ++i) {
      const Vec<R> u = replay(i, v, f, C[i], act, proj, frozen);
      if (mode == 1)
        for (size_t c = 0; c < C[i].size(); ++c) {
          C[i][c].act = act[c];
          C[i][c].proj = proj[c];
        }
      Vec<R> p = ldv(x, i);
      for (size_t c = 0; c < C[i].size(); ++c) {
        if (!act[c]) continue;
        p = p + (thick - C[i][c].d) * C[i][c].n;
        contact_count++;
        contact_sig += (((uint64_t)n * 2654435761u) ^ ((uint64_t)i << 20) ^ ((uint64_t)C[i][c].tri << 2) ^
                        (uint64_t)proj[c]) * 0x9E3779B97F4A7C15ull;
      }
      p = p + (h * u) * fixed[i];
      stv(x, i, p);
      stv(v, i, u);
    }
    if (mode == 1) rec[n] = C;
  }

  void forward(uint32_t nsteps) {
    x = x0;
    v = v0;
    xprev = x0;
    steps = nsteps;
    contact_sig = 0;
    contact_count = 0;
    ck_x.clear(); ck_v.clear(); ck_xp.clear();
    if (mode == 1) rec.assign(nsteps, {});
    std::vector<std::vector<Contact<R>>> C;
    for (uint32_t n = 0; n < nsteps; ++n) {
      if (n % K == 0) { ck_x.push_back(x); ck_v.push_back(v); ck_xp.push_back(xprev); }
      step(C, n);
    }
  }

  // One adjoint step at state (xs, vs, xps): (xb, vb) hold dL/d(x_{n+1}, v_{n+1})
  // on entry and dL/d(x_n, v_n) on exit.
  void backstep(uint32_t n, const std::vector<R>& xs, const std::vector<R>& vs, const std::vector<R>& xps,
                std::vector<R>& xb, std::vector<R>& vb, std::vector<R>& fb, std::vector<R>& kbar,
                std::vector<R>& kdv) {
    std::vector<std::vector<Contact<R>>> C;
    const bool frozen = mode == 2;
    if (frozen) {
      C = rec[n];
      for (uint32_t i = 0; i < V; ++i)
        for (Contact<R>& c : C[i]) c.d = dotv(c.n, ldv(xs, i) - c.t0);
    } else {
      detect(C, xs, xps);
    }
    std::vector<R> f;
    forces(xs, vs, f);
    // per vertex: integrate + contacts
#pragma omp parallel for schedule(static)
    for (int64_t i_ = 0; i_ < (int64_t)V; ++i_) {
      const uint32_t i = (uint32_t)i_;
      std::vector<char> act, proj;
      replay(i, vs, f, C[i], act, proj, frozen);
      const R m = mass[i], mk = fixed[i];
      const Vec<R> pb = ldv(xb, i);
      Vec<R> ub = ldv(vb, i) + (h * mk) * pb;
      Vec<R> xbar = pb;
      for (size_t c = C[i].size(); c-- > 0;) {
        if (!act[c]) continue;
        const Vec<R> n = C[i][c].n;
        if (!(plant & 1u)) xbar = xbar - n * dotv(n, pb);
        if (proj[c] && !(plant & 2u)) ub = ub - dotv(ub, n) * n;
      }
      const R w = (h * mk) / m;
      const Vec<R> fbi = ub * w;
      stv(fb, i, fbi);
      stv(xb, i, xbar);
      stv(vb, i, ub + kair * fbi);
    }
    // per vertex: spring gather (each spring's parameter terms taken once, by
    // its lower-index end)
#pragma omp parallel for schedule(static)
    for (int64_t i_ = 0; i_ < (int64_t)V; ++i_) {
      const uint32_t i = (uint32_t)i_;
      const Vec<R> p = ldv(xs, i), vel = ldv(vs, i), fbi = ldv(fb, i);
      Vec<R> xa = {0, 0, 0}, va = {0, 0, 0};
      R kdl = 0;
      SpringRef sr[12];
      const int ns = G.springs(i, sr);
      for (int q = 0; q < ns; ++q) {
        const uint32_t j = sr[q].j;
        const Vec<R> dx = ldv(xs, j) - p, dv = ldv(vs, j) - vel;
        const R len = std::sqrt(dotv(dx, dx));
        if (len < (R)1e-7f) continue;
        const Vec<R> n = {dx.x / len, dx.y / len, dx.z / len};
        const R ks = k[sr[q].s];
        const R ndv = dotv(dv, n);
        const R ext_ = len - rest[sr[q].fam];
        const R s = ks * ext_ + kd * ndv;
        const Vec<R> g = fbi - ldv(fb, j);
        const R ng = dotv(n, g);
        const R inv = (R)1 / len;
        const Vec<R> Mdv = (dv - n * ndv) * inv;
        const Vec<R> Mg = (g - n * ng) * inv;
        const Vec<R> JT = (ks * n + ((plant & 4u) ? (R)0 : kd) * Mdv) * ng + s * Mg;
        xa = xa - JT;
        va = va - (kd * ng) * n;
        if (i < j) {
          kbar[sr[q].s] = kbar[sr[q].s] + ((plant & 8u) ? -ext_ : ext_) * ng;
          kdl = kdl + ndv * ng;
        }
      }
      stv(xb, i, ldv(xb, i) + xa);
      stv(vb, i, ldv(vb, i) + va);
      kdv[i] = kdv[i] + kdl;
    }
  }

  // In frozen mode the recorded contacts and gates are used throughout.
  void backward(const R* xbN, const R* vbN, R* gtheta, R* gkd, R* gx0, R* gv0) {
    if (mode == 1) mode = 0;
    std::vector<R> xb(xbN, xbN + 3 * (size_t)V), vb(3 * (size_t)V, 0), fb(3 * (size_t)V, 0);
    if (vbN) vb.assign(vbN, vbN + 3 * (size_t)V);
    std::vector<R> kbar(G.total, 0), kdv(V, 0);
    const uint32_t nseg = (uint32_t)ck_x.size();
    std::vector<std::vector<Contact<R>>> C;
    for (uint32_t sgi = nseg; sgi-- > 0;) {
      const uint32_t n0 = sgi * K, n1 = std::min(steps, n0 + K);
      std::vector<std::vector<R>> sx, sv, sxp;
      x = ck_x[sgi]; v = ck_v[sgi]; xprev = ck_xp[sgi];
      for (uint32_t n = n0; n < n1; ++n) {
        sx.push_back(x); sv.push_back(v); sxp.push_back(xprev);
        step(C, n);
      }
      for (uint32_t n = n1; n-- > n0;) backstep(n, sx[n - n0], sv[n - n0], sxp[n - n0], xb, vb, fb, kbar, kdv);
    }
    for (uint32_t s = 0; s < G.total; ++s) gtheta[s] = kbar[s] * k[s];
    // kd: 256 strided partial sums, then a halving tree (the device reduction order)
    R part[256];
    for (uint32_t t = 0; t < 256; ++t) {
      R a = 0;
      for (uint32_t i = t; i < V; i += 256) a = a + kdv[i];
      part[t] = a;
    }
    for (uint32_t w = 128; w > 0; w >>= 1)
      for (uint32_t t = 0; t < w; ++t) part[t] = part[t] + part[t + w];
    if (gkd) *gkd = part[0];
    if (gx0) std::copy(xb.begin(), xb.end(), gx0);
    if (gv0) std::copy(vb.begin(), vb.end(), gv0);
  }
};

template <class R>
Rollout<R>* make(const ysim_mesh_desc* cloth, const ysim_sim_params* P, const float* x0, const float* v0,
                 const float* mass, const float* fixed, const float* ext, uint32_t nobs, const float* obs9,
                 const uint32_t* obs_obj, const uint32_t* obs_fid) {
  Rollout<R>* r = new Rollout<R>();
  const ysim_fastgrid_params& c = cloth->cloth;
  r->G.init(c.particle_num_1d);
  r->V = cloth->num_vertices;
  r->rest[0] = c.stretch_rest_x; r->rest[1] = c.stretch_rest_y; r->rest[2] = c.shear_rest_a;
  r->rest[3] = c.shear_rest_b; r->rest[4] = c.bend_rest_x; r->rest[5] = c.bend_rest_y;
  r->h = P->subh; r->kair = P->kair; r->kd = P->kd; r->thick = c.thickness;
  // margin <= 0 selects margin = band = the cloth's thickness (the scene
  // convention); the band is formed in float as the device forms it
  const float margin = P->margin > 0.f ? P->margin : c.thickness;
  const float band = P->radius + (P->margin > 0.f ? P->narrow_thickness : c.thickness);
  r->margin = margin;
  r->band = band;
  const size_t V3 = 3 * (size_t)r->V;
  r->x0.assign(x0, x0 + V3);
  r->v0.assign(v0, v0 + V3);
  r->mass.assign(mass, mass + r->V);
  r->fixed.assign(fixed, fixed + r->V);
  r->ext.assign(ext, ext + V3);
  r->obs.assign(obs9, obs9 + 9 * (size_t)nobs);
  r->obs_obj.assign(obs_obj, obs_obj + nobs);
  r->obs_fid.assign(obs_fid, obs_fid + nobs);
  r->obs_box.resize(6 * (size_t)nobs);
  for (uint32_t t = 0; t < nobs; ++t)
    for (int a = 0; a < 3; ++a) {
      R lo = r->obs[9 * t + a], hi = lo;
      for (int q = 1; q < 3; ++q) {
        lo = std::min(lo, r->obs[9 * t + 3 * q + a]);
        hi = std::max(hi, r->obs[9 * t + 3 * q + a]);
      }
      r->obs_box[6 * t + a] = lo;
      r->obs_box[6 * t + 3 + a] = hi;
    }
  r->k.assign(r->G.total, (R)1);
  return r;
}

// theta -> k as the device computes it: exp in double, rounded once to R.
template <class R>
void forward_(void* h, const double* theta, double kd, uint32_t steps, uint32_t K, uint32_t mode, double* x_out,
              double* v_out) {
  Rollout<R>* r = (Rollout<R>*)h;
  r->mode = (mode == 2 && r->rec.size() != steps) ? 0 : (int)mode;
  for (uint32_t s = 0; s < r->G.total; ++s) r->k[s] = (R)std::exp(theta[s]);
  r->kd = (R)kd;
  r->K = K ? K : 1;
  r->forward(steps);
  for (size_t q = 0; q < r->x.size(); ++q) {
    if (x_out) x_out[q] = (double)r->x[q];
    if (v_out) v_out[q] = (double)r->v[q];
  }
}

template <class R>
void backward_(void* h, const double* xbar, const double* vbar, double* gtheta, double* gkd, double* gx0,
               double* gv0) {
  Rollout<R>* r = (Rollout<R>*)h;
  const size_t V3 = 3 * (size_t)r->V;
  std::vector<R> xb(V3), vb(V3, 0), gt(r->G.total), x0(V3), v0(V3);
  for (size_t q = 0; q < V3; ++q) {
    xb[q] = (R)xbar[q];
    if (vbar) vb[q] = (R)vbar[q];
  }
  R kd = 0;
  r->backward(xb.data(), vb.data(), gt.data(), &kd, x0.data(), v0.data());
  for (uint32_t s = 0; s < r->G.total; ++s) gtheta[s] = (double)gt[s];
  if (gkd) *gkd = (double)kd;
  for (size_t q = 0; q < V3; ++q) {
    if (gx0) gx0[q] = (double)x0[q];
    if (gv0) gv0[q] = (double)v0[q];
  }
}

}  // namespace

extern "C" {

// precision 32 is the parity path (the device's operation order), 64 the
// finite-difference-checked restatement.  Values cross this interface as
// doubles; in precision 32 they are rounded to float on entry.
void* yo_adj_create(uint32_t precision, const ysim_mesh_desc* cloth, const ysim_sim_params* P, const float* x0,
                    const float* v0, const float* mass, const float* fixed, const float* ext, uint32_t nobs,
                    const float* obs9, const uint32_t* obs_obj, const uint32_t* obs_fid) {
  if (precision == 64) return make<double>(cloth, P, x0, v0, mass, fixed, ext, nobs, obs9, obs_obj, obs_fid);
  return make<float>(cloth, P, x0, v0, mass, fixed, ext, nobs, obs9, obs_obj, obs_fid);
}
void yo_adj_destroy(uint32_t precision, void* h) {
  if (precision == 64) delete (Rollout<double>*)h;
  else delete (Rollout<float>*)h;
}
uint32_t yo_adj_num_springs(uint32_t n) {
  Grid g;
  g.init(n);
  return g.total;
}
// mode 0: detect contacts; 1: detect and record them with their gates; 2: replay
// the recorded contacts and gates (frozen; falls back to 0 if no record of
// `steps` steps exists).
void yo_adj_forward(uint32_t precision, void* h, const double* theta, double kd, uint32_t steps, uint32_t K,
                    uint32_t mode, double* x_out, double* v_out) {
  if (precision == 64) forward_<double>(h, theta, kd, steps, K, mode, x_out, v_out);
  else forward_<float>(h, theta, kd, steps, K, mode, x_out, v_out);
}
// Copy the contact record (contacts and gates of every step) of one rollout to
// another of the same cloth, e.g. FP32's record into the FP64 rollout so that
// FP64 evaluates the map on the FP32 trajectory's piece; `dst` switches to
// frozen mode.
void yo_adj_copy_record(uint32_t src_prec, void* src, uint32_t dst_prec, void* dst) {
  auto conv = [](auto* a, auto* b) {
    using RB = decltype(b->h);
    b->rec.assign(a->rec.size(), {});
    for (size_t n = 0; n < a->rec.size(); ++n) {
      b->rec[n].assign(a->rec[n].size(), {});
      for (size_t i = 0; i < a->rec[n].size(); ++i)
        for (auto& c : a->rec[n][i]) {
          Contact<RB> d;
          d.tobj = c.tobj;
          d.tri = c.tri;
          d.n = {(RB)c.n.x, (RB)c.n.y, (RB)c.n.z};
          d.d = (RB)c.d;
          d.t0 = {(RB)c.t0.x, (RB)c.t0.y, (RB)c.t0.z};
          d.act = c.act;
          d.proj = c.proj;
          b->rec[n][i].push_back(d);
        }
    }
    b->mode = 2;
  };
  if (src_prec == 64 && dst_prec == 64) conv((Rollout<double>*)src, (Rollout<double>*)dst);
  else if (src_prec == 64) conv((Rollout<double>*)src, (Rollout<float>*)dst);
  else if (dst_prec == 64) conv((Rollout<float>*)src, (Rollout<double>*)dst);
  else conv((Rollout<float>*)src, (Rollout<float>*)dst);
}

// Replace the initial state (x_0, v_0; xPrev_0 = x_0) of later forwards.
void yo_adj_set_initial(uint32_t precision, void* h, const double* x0, const double* v0) {
  auto set = [&](auto* r) {
    using RR = decltype(r->h);
    for (size_t q = 0; q < r->x0.size(); ++q) {
      if (x0) r->x0[q] = (RR)x0[q];
      if (v0) r->v0[q] = (RR)v0[q];
    }
  };
  if (precision == 64) set((Rollout<double>*)h);
  else set((Rollout<float>*)h);
}
void yo_adj_plant(uint32_t precision, void* h, uint32_t bits) {
  if (precision == 64) ((Rollout<double>*)h)->plant = bits;
  else ((Rollout<float>*)h)->plant = bits;
}

// per-vertex number of contacts applied (gate distance < thickness) over the
// recorded forward
void yo_adj_contact_hist(uint32_t precision, void* h, uint32_t* per_vertex) {
  auto fill = [&](auto* r) {
    std::fill(per_vertex, per_vertex + r->V, 0u);
    for (auto& st : r->rec)
      for (uint32_t i = 0; i < r->V; ++i)
        for (auto& c : st[i]) per_vertex[i] += c.act ? 1u : 0u;
  };
  if (precision == 64) fill((Rollout<double>*)h);
  else fill((Rollout<float>*)h);
}
void yo_adj_backward(uint32_t precision, void* h, const double* xbar, const double* vbar, double* gtheta,
                     double* gkd, double* gx0, double* gv0) {
  if (precision == 64) backward_<double>(h, xbar, vbar, gtheta, gkd, gx0, gv0);
  else backward_<float>(h, xbar, vbar, gtheta, gkd, gx0, gv0);
}
// contacts applied over the last forward: count and an order-independent signature
void yo_adj_contacts(uint32_t precision, void* h, uint64_t* count, uint64_t* sig) {
  if (precision == 64) {
    *count = ((Rollout<double>*)h)->contact_count;
    *sig = ((Rollout<double>*)h)->contact_sig;
  } else {
    *count = ((Rollout<float>*)h)->contact_count;
    *sig = ((Rollout<float>*)h)->contact_sig;
  }
}

}  // extern "C"
