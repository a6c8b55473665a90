// C++ self-test of the HIP backend tag (ysim_amd/host/ysim_hip_backend.hpp),
// in the style of the reference's headless self-test (test/self_test_inline.hpp):
// PASS/FAIL lines on stderr, exit code 0 when everything passes.
//   Block 6  (drape, :1413-1517): mean y falls, contacts fire, no tunnelling
//   Block 11 (determinism, :1971-2053): two runs are bit-identical
//   Block 13 (contacts, :2168-2250): contacts for overlap, none when separated
//   staged driving (refit/query/integrate) == one-call scene_step
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../ysim_amd/host/ysim_hip_backend.hpp"

// The CPU oracle (oracle/liboracle.so, test infrastructure): the same packed
// scene stepped on the host, the reference for bit-exact comparisons.
extern "C" {
void* yo_scene_create(uint32_t nm, const ysim_mesh_desc* meshes, const float* x, const float* v, const float* mass,
                      const float* ext, const float* fixed, const uint32_t* facets, const ysim_sim_params* P);
void yo_scene_destroy(void* s);
void yo_scene_step(void* s, uint32_t n);
void yo_scene_get_state(void* s, float* x, float* v, float* xprev);
int yo_scene_set_analytic(void* s, const ysim_analytic_shape* shapes, uint32_t n);
void yo_scene_set_collision(void* s, float margin, float radius, float narrow_thickness, uint32_t self_collision);
}

static int failures = 0;
static void report(bool ok, const char* name) {
  std::fprintf(stderr, "%s %s\n", ok ? "PASS" : "FAIL", name);
  if (!ok) ++failures;
}

static void uv_sphere(int slices, int stacks, float r, const float c[3], std::vector<float>& x,
                      std::vector<uint32_t>& f) {
  const double PI = 3.14159265358979323846;
  x = {c[0], c[1] + r, c[2]};
  for (int i = 1; i < stacks; ++i)
    for (int j = 0; j < slices; ++j) {
      double ph = PI * i / stacks, th = 2 * PI * j / slices;
      x.push_back((float)(c[0] + r * std::sin(ph) * std::cos(th)));
      x.push_back((float)(c[1] + r * std::cos(ph)));
      x.push_back((float)(c[2] + r * std::sin(ph) * std::sin(th)));
    }
  x.insert(x.end(), {c[0], c[1] - r, c[2]});
  auto ring = [&](int i, int j) { return (uint32_t)(1 + (i - 1) * slices + (j % slices)); };
  uint32_t bot = (uint32_t)(x.size() / 3 - 1);
  for (int j = 0; j < slices; ++j) f.insert(f.end(), {0u, ring(1, j + 1), ring(1, j)});
  for (int i = 1; i < stacks - 1; ++i)
    for (int j = 0; j < slices; ++j)
      f.insert(f.end(), {ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i, j), ring(i + 1, j + 1), ring(i + 1, j)});
  for (int j = 0; j < slices; ++j) f.insert(f.end(), {bot, ring(stacks - 1, j), ring(stacks - 1, j + 1)});
}

static void build(Simulator<HIP, float>& sim, float height, uint32_t n) {
  const float center[3] = {0.5f, height, 0.5f};
  const float l = 1.0f / (n - 1);
  std::vector<float> jitter(n * n);
  for (uint32_t i = 0; i < n * n; ++i) jitter[i] = 1e-3f * l * std::sin(0.7f * i);  // deterministic
  sim.scene.addClothGridFastAt(n, 1.0f, center, 500.f, 50.f, 5.f, 0.5f * l, 1.0f, &jitter);
  std::vector<float> sx;
  std::vector<uint32_t> sf;
  const float sc[3] = {0.5f, 0.5f, 0.5f};
  uv_sphere(24, 12, 0.3f, sc, sx, sf);
  sim.scene.addFloatMesh(sx, sf);
  sim.margin = 0.5f * l;
  sim.radius = 0.f;
  sim.system.kd = 0.01f;
  sim.initialize();
}

int main() {
  int ndev = 0;
  if (ysim_hip_device_count(&ndev) != YSIM_OK || ndev == 0) {
    std::fprintf(stderr, "SKIP no HIP device\n");
    return 0;
  }
  const uint32_t n = 24;
  SymplecticSystem<HIP, float> sysA(1.0f / 60.f, 10), sysB(1.0f / 60.f, 10);
  Simulator<HIP, float> A(sysA), B(sysB);
  build(A, 0.8f, n);
  build(B, 0.8f, n);
  std::vector<float> x0 = A.scene.positions();
  double y0 = 0;
  for (uint32_t i = 0; i < n * n; ++i) y0 += x0[3 * i + 1];
  size_t fired = 0;
  for (int frame = 0; frame < 8; ++frame) {
    A.update();
    std::vector<uint32_t> off;
    std::vector<ysim_narrow_collision> c;
    A.narrowPhase.contacts(off, c);
    fired += c.size();
  }
  std::vector<float> x1 = A.scene.positions();
  double y1 = 0, rmin = 1e9;
  for (uint32_t i = 0; i < n * n; ++i) {
    y1 += x1[3 * i + 1];
    double dx = x1[3 * i] - 0.5, dy = x1[3 * i + 1] - 0.5, dz = x1[3 * i + 2] - 0.5;
    rmin = std::fmin(rmin, std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  report(y1 < y0, "Block 6 / mean y falls");
  report(fired > 0, "Block 6 / contacts fire");
  const double sagitta = 0.3 * (1 - std::cos(3.14159265358979 / 12));
  report(rmin > 0.3 - sagitta - 2 * (0.5 / (n - 1)), "Block 6 / no tunnelling into the sphere");
  report(A.system.anomalyFlag == 0, "NAN-GUARD / no anomaly on a sane drape");

  // Block 11: B driven with the one-call substep must equal A's staged run
  ysim_hip_detail::check(ysim_hip_scene_step(B.scene.handle, 8 * 10, 1), "scene_step");
  std::vector<float> xb = B.scene.positions();
  report(std::memcmp(x1.data(), xb.data(), 4 * x1.size()) == 0, "Block 11 / staged == graph substeps, bit-identical");

  // Block 13: a cloth far above the sphere has no contacts
  SymplecticSystem<HIP, float> sysC(1.0f / 60.f, 10);
  Simulator<HIP, float> Cfar(sysC);
  build(Cfar, 1.6f, n);
  Cfar.update();
  std::vector<uint32_t> off;
  std::vector<ysim_narrow_collision> c;
  Cfar.narrowPhase.contacts(off, c);
  report(c.empty(), "Block 13 / separated objects have no contacts");

  // TriangularCloth: the same grid geometry as a general mesh (edge and
  // opposite-vertex springs) drapes onto the sphere without anomalies
  {
    SymplecticSystem<HIP, float> sysT(1.0f / 60.f, 10);
    Simulator<HIP, float> Tr(sysT);
    const float l = 1.0f / (n - 1);
    std::vector<float> tx;
    std::vector<uint32_t> tf;
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t c = 0; c < n; ++c) tx.insert(tx.end(), {c * l, 0.8f, r * l});
    for (uint32_t r = 0; r + 1 < n; ++r)
      for (uint32_t c = 0; c + 1 < n; ++c) {
        const uint32_t p00 = r * n + c, p10 = p00 + 1, p01 = p00 + n, p11 = p01 + 1;
        tf.insert(tf.end(), {p00, p01, p11, p00, p11, p10});
      }
    Tr.scene.addTriangularCloth(tx, tf, 500.f, 50.f, 5.f, 0.5f * l, 1.0f);
    std::vector<float> sx;
    std::vector<uint32_t> sf;
    const float sc[3] = {0.5f, 0.5f, 0.5f};
    uv_sphere(24, 12, 0.3f, sc, sx, sf);
    Tr.scene.addFloatMesh(sx, sf);
    Tr.margin = 0.5f * l;
    Tr.radius = 0.f;
    Tr.system.kd = 0.01f;
    Tr.initialize();
    size_t tfired = 0;
    for (int frame = 0; frame < 8; ++frame) {
      Tr.update();
      std::vector<uint32_t> toff;
      std::vector<ysim_narrow_collision> tc;
      Tr.narrowPhase.contacts(toff, tc);
      tfired += tc.size();
    }
    std::vector<float> xt = Tr.scene.positions();
    double ty = 0;
    bool finite = true;
    for (uint32_t i = 0; i < n * n; ++i) {
      ty += xt[3 * i + 1];
      finite = finite && std::isfinite(xt[3 * i]) && std::isfinite(xt[3 * i + 1]) && std::isfinite(xt[3 * i + 2]);
    }
    report(finite && ty < 0.8 * n * n && tfired > 0 && Tr.system.anomalyFlag == 0,
           "TriangularCloth / drape falls, contacts fire, no anomaly");
  }

  // RadixSorter<HIP> through VectorBase<HIP>
  std::vector<uint64_t> keys = {5, 3, 5, 1, 3, 0};
  std::vector<uint32_t> vals = {0, 1, 2, 3, 4, 5};
  VectorBase<HIP, uint64_t> dk(keys.size());
  VectorBase<HIP, uint32_t> dv(vals.size());
  dk.upload(keys);
  dv.upload(vals);
  RadixSorter<HIP, uint64_t>().sort(dk, dv, 63);
  std::vector<uint32_t> sv = dv.download();
  report(sv == std::vector<uint32_t>({5, 3, 1, 4, 0, 2}), "RadixSorter<HIP> / stable order");

  // ByteMemoryPool<HIP>: 256-byte bump allocation, fill, overflow reported
  {
    ByteMemoryPool<HIP> pool(4096);
    auto a = pool.zeros<uint32_t>(10);
    auto b = pool.allocFill<float>(5, 2.5f);
    std::vector<float> hb(5);
    ysim_hip_detail::check(ysim_hip_download(ysim_hip_detail::global_ctx(), hb.data(), b.ptr, 20), "download");
    std::vector<uint32_t> ha(10, 7u);
    ysim_hip_detail::check(ysim_hip_download(ysim_hip_detail::global_ctx(), ha.data(), a.ptr, 40), "download");
    bool threw = false;
    try {
      pool.alloc<char>(4096);
    } catch (const std::length_error&) {
      threw = true;
    }
    report(a.offset == 0 && b.offset == 256 && hb == std::vector<float>(5, 2.5f) &&
               ha == std::vector<uint32_t>(10, 0u) && threw,
           "ByteMemoryPool<HIP> / alignment, zeros, fill, overflow");
  }

  // FastGridClothBehavior<HIP>: clothStiffnessScale scales k*, not rest lengths,
  // and a scaled sheet moves differently from an unscaled one
  {
    SymplecticSystem<HIP, float> sysS(1.0f / 60.f, 10);
    Simulator<HIP, float> S1(sysS), S2(sysS);
    build(S1, 0.8f, n);
    S2.scene.meshes.clear();
    build(S2, 0.8f, n);
    S2.scene.meshes[0].clothStiffnessScale = 4.0f;
    ysim_fastgrid_params p1 = FastGridClothBehavior<HIP, float>::params(S1.scene.meshes[0]);
    ysim_fastgrid_params p2 = FastGridClothBehavior<HIP, float>::params(S2.scene.meshes[0]);
    S2.initialize();
    S1.update();
    S2.update();
    std::vector<float> y1 = S1.scene.positions(), y2 = S2.scene.positions();
    report(p2.kstretch == 4.0f * p1.kstretch && p2.kbend == 4.0f * p1.kbend &&
               p2.stretch_rest_x == p1.stretch_rest_x && std::memcmp(y1.data(), y2.data(), 4 * y1.size()) != 0,
           "FastGridClothBehavior<HIP> / stiffness scale");
  }

  // Simulator<HIP>::autotune keeps the state and the results
  {
    SymplecticSystem<HIP, float> sysT(1.0f / 60.f, 10);
    Simulator<HIP, float> T1(sysT), T2(sysT);
    build(T1, 0.8f, n);
    build(T2, 0.8f, n);
    std::vector<float> before = T1.scene.positions();
    ysim_tuning t = T1.autotune(4);
    std::vector<float> after = T1.scene.positions();
    T1.update();
    T2.update();
    std::vector<float> z1 = T1.scene.positions(), z2 = T2.scene.positions();
    report(before == after && z1 == z2 && t.ms_per_substep > 0.f, "Simulator<HIP>::autotune / results unchanged");
  }

  // The reference's default scene: a cloth over a sphere whose collider is the
  // analytic sphere (scene.hpp:326-331), at the reference's rebuild cadence
  // (frame % 10 == 0, substep 0: simulator.hpp:3253), with the per-call
  // margin and radius changed mid-run; bit-exact against the oracle.
  {
    SymplecticSystem<HIP, float> sysA2(1.0f / 60.f, 10);
    Simulator<HIP, float> S(sysA2);
    const float l = 1.0f / (n - 1);
    const float center[3] = {0.5f, 0.8f, 0.5f};
    std::vector<float> jitter(n * n);
    for (uint32_t i = 0; i < n * n; ++i) jitter[i] = 1e-3f * l * std::cos(0.3f * i);
    S.scene.addClothGridFastAt(n, 1.0f, center, 500.f, 50.f, 5.f, 0.5f * l, 1.0f, &jitter);
    std::vector<float> sx;
    std::vector<uint32_t> sf;
    const float sc[3] = {0.5f, 0.5f, 0.5f};
    uv_sphere(24, 12, 0.3f, sc, sx, sf);
    S.scene.addFloatMesh(sx, sf, YSIM_COLLIDER_SPHERE);
    S.scene.addAnalyticSphere(1, sc, 0.3f);
    S.margin = 0.5f * l;
    S.radius = 0.f;
    S.system.kd = 0.01f;
    S.enableRefit = true;  // the reference cadence
    S.initialize();
    const Scene<HIP, float>::Packed pk = S.scene.packed();
    const ysim_sim_params P = S.params();
    void* O = yo_scene_create((uint32_t)pk.descs.size(), pk.descs.data(), pk.x.data(), pk.v.data(), pk.mass.data(),
                              pk.ext.data(), pk.fixed.data(), pk.facets.data(), &P);
    report(O && yo_scene_set_analytic(O, S.scene.analyticShapes.data(), (uint32_t)S.scene.analyticShapes.size()) == 0,
           "analytic sphere / oracle scene created");
    std::vector<float> xo(pk.x.size());
    size_t fired = 0;
    bool same = true;
    for (int frame = 0; frame < 12; ++frame) {
      if (frame == 6) {  // a caller changes Simulator::margin and radius between frames
        S.margin = 0.75f * l;
        S.radius = 0.002f;
        yo_scene_set_collision(O, S.margin, S.radius, S.margin, 1u);
      }
      S.update();
      yo_scene_step(O, (uint32_t)S.system.subSteps);
      std::vector<uint32_t> off;
      std::vector<ysim_narrow_collision> c;
      S.narrowPhase.contacts(off, c);
      fired += c.size();
      yo_scene_get_state(O, xo.data(), nullptr, nullptr);
      std::vector<float> xg = S.scene.positions();
      same = same && std::memcmp(xg.data(), xo.data(), 4 * xo.size()) == 0;
    }
    yo_scene_destroy(O);
    report(P.rebuild_every == 100u, "Simulator<HIP> / enableRefit maps to the frame % 10 rebuild cadence");
    report(fired > 0, "analytic sphere / contacts fire");
    report(same, "analytic sphere / 12 frames across the frame-10 rebuild and a margin change, bit-exact vs oracle");
    bool threw = false;
    try {
      S.broadPhase.refit();
      S.broadPhase.detectCollisions(S.margin, true, true, /*resetCounter*/ false);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    report(threw, "BVH<HIP>::detectCollisions / appending to another pass's pairs is refused");
  }

  std::fprintf(stderr, "%s: %d failure(s)\n", failures ? "FAILED" : "ALL PASSED", failures);
  return failures ? 1 : 0;
}
