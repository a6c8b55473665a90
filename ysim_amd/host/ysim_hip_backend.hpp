// ysim_hip_backend.hpp — the `HIP` backend tag of ysim's templated core.
//
// Adds `struct HIP : Backend {}` beside the reference's CPU/CUDA/METAL tags
// (include/sim/core_types.hpp:7-10) and HIP specializations of the stage
// objects the Simulator holds (include/sim/simulator.hpp:7-38):
//   VectorBase<HIP, T>            device array view + storage (core_types.hpp:555-611)
//   RadixSorter<HIP, E>           stable LSD sort (spatial_hash.hpp:7-71)
//   Scene<HIP, PR>                packed scene (scene.hpp:783-1182)
//   BVH<HIP, PR, SCENE, OBJECT>   broad phase stage methods (bvh.hpp:2537-2949)
//   BruteForce<HIP, PR>           narrow phase + contact CSR (bruteforce.hpp:311-384)
//   SymplecticSystem<HIP, PR>     springs + integrate (symplectic_system.hpp:160-409)
//   Simulator<HIP, PR>            initialize / update substep loop (simulator.hpp:2942-3815)
// Every specialization is a thin shim over the C ABI in include/ysim_hip.h:
// no HIP headers, no CUDA compatibility layer, host C++17 only.
//
// Inside the reference, define YSIM_HIP_IN_REFERENCE before including this
// header so the reference's own Backend tag and primary templates are used;
// standalone (this repository's tests) the header declares them itself.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ysim_hip.h"

#ifndef YSIM_HIP_IN_REFERENCE
struct Backend {};
struct CPU : Backend {};
struct CUDA : Backend {};
struct METAL : Backend {};
template <typename BE, typename T> struct VectorBase;
template <typename BE> struct ByteMemoryPool;
template <typename BE, typename PR> struct FastGridClothBehavior;
template <typename BE, typename E> struct RadixSorter;
template <typename BE, typename PR> struct Scene;
struct SCENE {};
struct OBJECT {};
template <typename BE, typename PR, typename LEVEL, typename PRIM> struct BVH;
template <typename BE, typename PR> struct BruteForce;
template <typename BE, typename PR> struct SymplecticSystem;
template <typename BE, typename PR> struct Simulator;
enum class BehaviorType { Float = 0, FastGridCloth = 1, TriangularCloth = 2 };
#endif

struct HIP : Backend {};

namespace ysim_hip_detail {
inline void check(int rc, const char* what) {
  if (rc != YSIM_OK) throw std::runtime_error(std::string(what) + ": " + ysim_hip_last_error());
}
// One device context per process (the MetalGlobalContext analogue).
inline ysim_ctx* global_ctx(int device = 0) {
  static ysim_ctx* ctx = nullptr;
  if (!ctx) check(ysim_hip_ctx_create(device, 0, &ctx), "ysim_hip_ctx_create");
  return ctx;
}
}  // namespace ysim_hip_detail

// ---------------------------------------------------------------- VectorBase
// Owning device array.  Unlike the Metal backend, MI355X memory is not
// unified: host reads go through download().
template <typename T>
struct VectorBase<HIP, T> {
  T* ptr = nullptr;  // device pointer
  size_t size = 0;
  VectorBase() = default;
  explicit VectorBase(size_t n) : size(n) {
    void* p = nullptr;
    ysim_hip_detail::check(ysim_hip_alloc(ysim_hip_detail::global_ctx(), n * sizeof(T), &p), "alloc");
    ptr = static_cast<T*>(p);
  }
  VectorBase(const VectorBase&) = delete;
  VectorBase& operator=(const VectorBase&) = delete;
  VectorBase(VectorBase&& o) noexcept : ptr(o.ptr), size(o.size) { o.ptr = nullptr; o.size = 0; }
  ~VectorBase() {
    if (ptr) ysim_hip_free(ysim_hip_detail::global_ctx(), ptr);
  }
  void upload(const std::vector<T>& h) {
    if (h.size() != size) throw std::invalid_argument("VectorBase<HIP>::upload: size mismatch");
    ysim_hip_detail::check(ysim_hip_upload(ysim_hip_detail::global_ctx(), ptr, h.data(), size * sizeof(T)), "upload");
  }
  std::vector<T> download() const {
    std::vector<T> h(size);
    ysim_hip_detail::check(ysim_hip_download(ysim_hip_detail::global_ctx(), h.data(), ptr, size * sizeof(T)), "download");
    return h;
  }
};

// ---------------------------------------------------------------- ByteMemoryPool
// ByteMemoryPool<METAL> (include/sim/core_types.hpp:268-330): one device arena,
// bump allocation with 256-byte alignment.  alloc() returns a non-owning block;
// overflow throws with the required size instead of printing and returning an
// empty block.
template <typename PR>
struct MemoryBlockHIP {
  PR* ptr = nullptr;   // device pointer
  size_t size = 0;
  size_t offset = 0;   // bytes from the arena start
};

template <>
struct ByteMemoryPool<HIP> {
  char* pool = nullptr;  // device
  size_t marker = 0;
  size_t capacity = 0;
  ByteMemoryPool() = default;
  explicit ByteMemoryPool(size_t n) : capacity(n) {
    void* p = nullptr;
    ysim_hip_detail::check(ysim_hip_alloc(ysim_hip_detail::global_ctx(), n, &p), "ByteMemoryPool<HIP> alloc");
    pool = static_cast<char*>(p);
  }
  ByteMemoryPool(const ByteMemoryPool&) = delete;
  ByteMemoryPool& operator=(const ByteMemoryPool&) = delete;
  ByteMemoryPool(ByteMemoryPool&& o) noexcept : pool(o.pool), marker(o.marker), capacity(o.capacity) {
    o.pool = nullptr;
  }
  ByteMemoryPool& operator=(ByteMemoryPool&& o) noexcept {
    if (this != &o) {
      release();
      pool = o.pool;
      marker = o.marker;
      capacity = o.capacity;
      o.pool = nullptr;
    }
    return *this;
  }
  ~ByteMemoryPool() { release(); }
  void release() {
    if (pool) ysim_hip_free(ysim_hip_detail::global_ctx(), pool);
    pool = nullptr;
  }
  template <typename PR>
  MemoryBlockHIP<PR> alloc(size_t count) {
    MemoryBlockHIP<PR> b;
    if (count == 0) return b;
    const size_t bytes = (sizeof(PR) * count + 255) & ~size_t(255);
    if (marker + bytes > capacity)
      throw std::length_error("ByteMemoryPool<HIP>: needs " + std::to_string(marker + bytes) + " bytes, capacity " +
                              std::to_string(capacity));
    b.ptr = reinterpret_cast<PR*>(pool + marker);
    b.size = count;
    b.offset = marker;
    marker += bytes;
    return b;
  }
  template <typename PR>
  MemoryBlockHIP<PR> zeros(size_t count) {
    auto b = alloc<PR>(count);
    if (count)
      ysim_hip_detail::check(ysim_hip_memset(ysim_hip_detail::global_ctx(), b.ptr, 0, count * sizeof(PR)), "memset");
    return b;
  }
  template <typename PR>
  MemoryBlockHIP<PR> allocFill(size_t count, PR fill) {
    auto b = alloc<PR>(count);
    if (count) {
      std::vector<PR> h(count, fill);
      ysim_hip_detail::check(ysim_hip_upload(ysim_hip_detail::global_ctx(), b.ptr, h.data(), count * sizeof(PR)),
                             "upload");
    }
    return b;
  }
  void reset() { marker = 0; }   // the pool rewind of Simulator::initialize (core_types.hpp:462-473)
  char* bytePtr() { return pool; }
};

// ---------------------------------------------------------------- RadixSorter
// Keys and values as separate arrays (the reference sorts MortonNode pairs).
template <typename E>
struct RadixSorter<HIP, E> {
  void sort(VectorBase<HIP, E>& keys, VectorBase<HIP, uint32_t>& vals, uint32_t key_bits = 8 * sizeof(E)) {
    static_assert(sizeof(E) == 4 || sizeof(E) == 8, "uint32 or uint64 keys");
    if (keys.size != vals.size) throw std::invalid_argument("RadixSorter<HIP>::sort: size mismatch");
    if ((sizeof(E) == 4) != (key_bits <= 32)) throw std::invalid_argument("RadixSorter<HIP>::sort: key_bits vs key type");
    ysim_hip_detail::check(ysim_hip_sort_pairs(ysim_hip_detail::global_ctx(), keys.ptr, vals.ptr,
                                               (uint32_t)keys.size, key_bits), "sort_pairs");
  }
};

// ---------------------------------------------------------------- Scene
// Host-side mesh list packed into the device-resident scene by pack().
template <typename PR>
struct Scene<HIP, PR> {
  static_assert(sizeof(PR) == 4, "the HIP backend computes in FP32");
  struct Mesh {
    std::vector<float> x;          // packed float3
    std::vector<uint32_t> facets;  // packed uint3, mesh-local ids
    BehaviorType behaviorType = BehaviorType::Float;
    ysim_fastgrid_params cloth{};
    float clothStiffnessScale = 1.0f;  // GeneralMesh::clothStiffnessScale: scales k*, not rest lengths
    float mass = 1.0f;             // per vertex
    bool collidable = true;
    uint32_t group = 0;
    uint32_t shape = YSIM_COLLIDER_MESH;  // ColliderKind lane (collision_pod.hpp:32)
  };
  std::vector<Mesh> meshes;
  // AnalyticShape records of the analytic colliders (Scene::refreshAnalyticShapes,
  // scene.hpp:659-745): set them (poses included) each frame; the broad phase
  // installs them when detectCollisions is asked for analytic colliders.
  std::vector<ysim_analytic_shape> analyticShapes;
  bool analyticActive() const { return !analyticShapes.empty(); }
  float gravity[3] = {0.f, -9.81f, 0.f};   // SceneEnvironment.gravity (scene.hpp:187)
  ysim_scene* handle = nullptr;
  uint32_t numMeshes() const { return (uint32_t)meshes.size(); }

  // addClothGridFastAt (simulator.hpp:853-922): XZ grid at `center`, rest
  // lengths measured from the (optionally jittered) geometry
  // (recomputeFastGridRest, collision_pod.hpp:418-442).
  void addClothGridFastAt(uint32_t n, float size1D, const float center[3], float kstretch, float kshear,
                          float kbend, float thickness, float mass, const std::vector<float>* jitter = nullptr) {
    Mesh m;
    m.behaviorType = BehaviorType::FastGridCloth;
    m.mass = mass;
    const double l = (double)size1D / (n - 1);
    m.x.resize(3 * (size_t)n * n);
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t c = 0; c < n; ++c) {
        size_t b = 3 * ((size_t)r * n + c);
        m.x[b] = (float)(c * l - size1D / 2.0 + center[0]);
        m.x[b + 1] = (float)center[1] + (jitter ? (*jitter)[(size_t)r * n + c] : 0.f);
        m.x[b + 2] = (float)(r * l - size1D / 2.0 + center[2]);
      }
    for (uint32_t r = 0; r + 1 < n; ++r)  // MeshGridInitializer facet order (initializers.hpp:95-122)
      for (uint32_t c = 0; c + 1 < n; ++c) {
        uint32_t p00 = r * n + c, p10 = p00 + 1, p01 = p00 + n, p11 = p01 + 1;
        if (((r + c) & 1) == 0) m.facets.insert(m.facets.end(), {p00, p01, p11, p00, p11, p10});
        else m.facets.insert(m.facets.end(), {p00, p01, p10, p10, p01, p11});
      }
    auto dist = [&](uint32_t a, uint32_t b) {
      float dx = m.x[3 * a] - m.x[3 * b], dy = m.x[3 * a + 1] - m.x[3 * b + 1], dz = m.x[3 * a + 2] - m.x[3 * b + 2];
      return std::sqrt((dx * dx + dy * dy) + dz * dz);
    };
    m.cloth.particle_num_1d = n;
    m.cloth.stretch_rest_x = dist(0, 1);
    m.cloth.stretch_rest_y = dist(0, n);
    m.cloth.shear_rest_a = dist(0, n + 1);
    m.cloth.shear_rest_b = dist(1, n);
    m.cloth.bend_rest_x = dist(0, 2);
    m.cloth.bend_rest_y = dist(0, 2 * n);
    m.cloth.kstretch = kstretch;
    m.cloth.kshear = kshear;
    m.cloth.kbend = kbend;
    m.cloth.thickness = thickness;
    meshes.push_back(std::move(m));
  }
  // A TriangularCloth from any manifold triangle mesh (MeshFile / general mesh
  // initializers): ClothBehaviorParams {stretch, shear, bend, thickness}
  // (collision_pod.hpp:350-352); the library builds the edge and
  // opposite-vertex springs (MeshAdjacencyInitializer, mesh_state.hpp:236-454)
  // with rest lengths from x.
  void addTriangularCloth(std::vector<float> x, std::vector<uint32_t> facets, float stretch, float shear, float bend,
                          float thickness, float mass) {
    Mesh m;
    m.behaviorType = BehaviorType::TriangularCloth;
    m.x = std::move(x);
    m.facets = std::move(facets);
    m.mass = mass;
    m.cloth.kstretch = stretch;
    m.cloth.kshear = shear;
    m.cloth.kbend = bend;
    m.cloth.thickness = thickness;
    meshes.push_back(std::move(m));
  }
  // A static triangle obstacle (BehaviorType::Float).  shape: its ColliderKind
  // (YSIM_COLLIDER_*; not Mesh for a mesh answered by an analytic collider).
  void addFloatMesh(std::vector<float> x, std::vector<uint32_t> facets, uint32_t shape = YSIM_COLLIDER_MESH) {
    Mesh m;
    m.x = std::move(x);
    m.facets = std::move(facets);
    m.shape = shape;
    meshes.push_back(std::move(m));
  }
  // The analytic sphere of a static mesh (the reference's default collider of
  // a sphere object, scene.hpp:326-331): centre c and radius r, at rest.
  void addAnalyticSphere(uint32_t meshIndex, const float c[3], float r) {
    if (meshIndex >= meshes.size() || meshes[meshIndex].shape != YSIM_COLLIDER_SPHERE)
      throw std::invalid_argument("Scene<HIP>::addAnalyticSphere: the mesh must be a static sphere-collider mesh");
    ysim_analytic_shape a{};
    a.kind = YSIM_COLLIDER_SPHERE;
    a.obj_index = meshIndex;
    a.behavior_type = YSIM_BEHAVIOR_FLOAT;
    for (int k = 0; k < 3; ++k) {
      a.center_radius[k] = c[k];
      a.prev_center_pad[k] = c[k];
      a.half_ext_height[k] = r;
    }
    a.rot_quat[0] = a.prev_rot_quat[0] = 1.f;
    a.flags = 1u;  // collidable
    analyticShapes.push_back(a);
  }

  // Scene::pack (scene.hpp:783-1182): the packed buffers the resident scene
  // is created from.
  struct Packed {
    std::vector<ysim_mesh_desc> descs;
    std::vector<float> x, v, mass, ext, fixed;
    std::vector<uint32_t> facets;
  };
  Packed packed() const {
    Packed pk;
    pk.descs.resize(meshes.size());
    for (size_t i = 0; i < meshes.size(); ++i) {
      const Mesh& m = meshes[i];
      const uint32_t nv = (uint32_t)(m.x.size() / 3);
      ysim_mesh_desc& d = pk.descs[i];
      d = ysim_mesh_desc{};
      d.num_vertices = nv;
      d.num_facets = (uint32_t)(m.facets.size() / 3);
      d.behavior = m.behaviorType == BehaviorType::FastGridCloth     ? YSIM_BEHAVIOR_FASTGRID_CLOTH
                   : m.behaviorType == BehaviorType::TriangularCloth ? YSIM_BEHAVIOR_TRIANGULAR_CLOTH
                                                                     : YSIM_BEHAVIOR_FLOAT;
      d.collidable = m.collidable ? 1u : 0u;
      d.group = m.group;
      d.shape = m.shape;
      d.cloth = FastGridClothBehavior<HIP, PR>::params(m);
      pk.x.insert(pk.x.end(), m.x.begin(), m.x.end());
      pk.v.insert(pk.v.end(), m.x.size(), 0.f);
      pk.mass.insert(pk.mass.end(), nv, m.mass);
      pk.fixed.insert(pk.fixed.end(), nv, 1.f);
      for (uint32_t k = 0; k < nv; ++k)  // applyEnvironmentForces: ext = g * m (simulator.hpp:3148-3164)
        for (int a = 0; a < 3; ++a)
          pk.ext.push_back(d.behavior != YSIM_BEHAVIOR_FLOAT ? gravity[a] * m.mass : 0.f);
      pk.facets.insert(pk.facets.end(), m.facets.begin(), m.facets.end());
    }
    return pk;
  }
  // ... and the creation of the resident scene from them.
  void pack(const ysim_sim_params& params) {
    release();
    const Packed pk = packed();
    ysim_hip_detail::check(ysim_hip_scene_create(ysim_hip_detail::global_ctx(), (uint32_t)pk.descs.size(),
                                                 pk.descs.data(), pk.x.data(), pk.v.data(), pk.mass.data(),
                                                 pk.ext.data(), pk.fixed.data(), pk.facets.data(), &params, &handle),
                           "scene_create");
  }
  void release() {
    if (handle) ysim_hip_scene_destroy(handle);
    handle = nullptr;
  }
  ~Scene() { release(); }
  std::vector<float> positions() const {
    size_t n = 0;
    for (auto& m : meshes) n += m.x.size();
    std::vector<float> x(n);
    ysim_hip_detail::check(ysim_hip_scene_get_state(handle, x.data(), nullptr, nullptr), "scene_get_state");
    return x;
  }
};

// ---------------------------------------------------------------- FastGridClothBehavior
// FastGridClothBehavior<METAL> (include/sim/collision_pod.hpp:516-555) binds a
// mesh's buffers and the stiffness-scaled FastGridClothBehaviorParams, then
// dispatches compute_cloth_grid_forces_fast; integrate_cloth_grid follows in a
// second pass (symplectic_system.hpp:355-409).  On this backend forces and
// integration are one fused kernel over every cloth of the resident scene, so
// the behaviour packs the parameters once (at Scene::pack) and update() is the
// fused pass.
template <typename PR>
struct FastGridClothBehavior<HIP, PR> {
  // kstretch/kshear/kbend scaled by the mesh's clothStiffnessScale, rest lengths untouched
  static ysim_fastgrid_params params(const typename Scene<HIP, PR>::Mesh& m) {
    ysim_fastgrid_params fp = m.cloth;
    fp.kstretch *= m.clothStiffnessScale;
    fp.kshear *= m.clothStiffnessScale;
    fp.kbend *= m.clothStiffnessScale;
    return fp;
  }
  static void update(Scene<HIP, PR>& s) {
    ysim_hip_detail::check(ysim_hip_scene_integrate(s.handle), "scene_integrate");
  }
};

// ---------------------------------------------------------------- BVH (broad phase)
// The broad phase of the HIP scene: per-mesh LBVHs in HBM.  refit() and
// refitSwept() run the swept leaf + bottom-up pass (a full Morton/sort/Karras
// rebuild when the scene's rebuild cadence says so); enlargeTrajectory() is
// folded into that pass.  detectCollisions() records its arguments and
// installs or removes the analytic colliders; the traversal runs fused with
// the narrow phase in BruteForce::narrowAndSortByVertices, which hands the
// substep's margin, radius, thickness and self flag to the scene.
template <typename PR>
struct BVH<HIP, PR, SCENE, OBJECT> {
  Scene<HIP, PR>* scene = nullptr;
  bool syncEachPhase = false;
  bool refitted = false;
  bool detectRequested = false;
  float margin = 0.f;
  bool self = true;
  bool analytic = false;
  bool analyticInstalled = false;
  void build(Scene<HIP, PR>& s) { scene = &s; }  // trees are built by Scene::pack
  void refit() {
    ysim_hip_detail::check(ysim_hip_scene_refit(scene->handle), "scene_refit");
    refitted = true;
  }
  void enlargeTrajectory(PR) {
    if (!refitted) refit();  // the swept hull is part of the refit pass
  }
  void refitSwept(PR) { refit(); }
  // detectCollisions(margin, self, analytic, resetCounter) (bvh.hpp:2853).
  // resetCounter = false asks to append this pass to pairs another source
  // already wrote in the substep (the reference's CPU self-collision rows);
  // the fused query produces the substep's whole candidate set itself, so
  // that is refused.
  void detectCollisions(PR m, bool selfCollide = true, bool analyticColliders = false, bool resetCounter = true) {
    if (!resetCounter)
      throw std::invalid_argument("BVH<HIP>::detectCollisions: resetCounter = false (appending to another "
                                  "detection's pairs) is not supported; the fused query covers every pair");
    if (analyticColliders && !scene->analyticActive())
      throw std::invalid_argument("BVH<HIP>::detectCollisions: analytic colliders requested but the scene has none");
    if (analyticColliders) {
      // poses of this frame (same mesh sequence: a captured graph stays valid)
      ysim_hip_detail::check(ysim_hip_scene_set_analytic(scene->handle, scene->analyticShapes.data(),
                                                         (uint32_t)scene->analyticShapes.size()),
                             "scene_set_analytic");
      analyticInstalled = true;
    } else if (analyticInstalled) {
      ysim_hip_detail::check(ysim_hip_scene_set_analytic(scene->handle, nullptr, 0), "scene_set_analytic");
      analyticInstalled = false;
    }
    margin = (float)m;
    self = selfCollide;
    analytic = analyticColliders;
    detectRequested = true;
  }
  // Candidate pairs at the current state (for inspection / parity).
  std::vector<ysim_broad_collision> pairs() const {
    uint32_t n = 0;
    ysim_hip_detail::check(ysim_hip_scene_get_pairs(scene->handle, nullptr, 0, &n), "scene_get_pairs");
    std::vector<ysim_broad_collision> out(n);
    ysim_hip_detail::check(ysim_hip_scene_get_pairs(scene->handle, out.data(), n, &n), "scene_get_pairs");
    return out;
  }
};

// ---------------------------------------------------------------- BruteForce (narrow phase)
template <typename PR>
struct BruteForce<HIP, PR> {
  BVH<HIP, PR, SCENE, OBJECT>* broad = nullptr;
  bool syncEachPhase = false;
  // narrowAndSortByVertices(radius, thickness, analytic) (bruteforce.hpp:311):
  // the band is radius + thickness (strict <), the query box the margin of
  // the preceding detectCollisions.
  void narrowAndSortByVertices(PR radius, PR thickness, bool analyticColliders = false) {
    if (!broad || !broad->refitted) throw std::logic_error("BruteForce<HIP>: refit the broad phase first");
    if (!broad->detectRequested) throw std::logic_error("BruteForce<HIP>: call detectCollisions first");
    if (analyticColliders != broad->analytic)
      throw std::invalid_argument("BruteForce<HIP>: the analytic flag differs from the substep's detectCollisions");
    ysim_hip_detail::check(ysim_hip_scene_set_collision(broad->scene->handle, broad->margin, (float)radius,
                                                        (float)thickness, broad->self ? 1u : 0u),
                           "scene_set_collision");
    ysim_hip_detail::check(ysim_hip_scene_query(broad->scene->handle), "scene_query");
    broad->refitted = false;
    broad->detectRequested = false;
  }
  // The per-vertex contact CSR of the last substep, ordered (vertex, target, facet).
  void contacts(std::vector<uint32_t>& offsets, std::vector<ysim_narrow_collision>& out) const {
    Scene<HIP, PR>& s = *broad->scene;
    size_t nv = 0;
    for (auto& m : s.meshes) nv += m.x.size() / 3;
    offsets.assign(nv + 1, 0);
    uint32_t n = 0;
    ysim_hip_detail::check(ysim_hip_scene_get_contacts(s.handle, offsets.data(), nullptr, 0, &n), "scene_get_contacts");
    out.resize(n);
    ysim_hip_detail::check(ysim_hip_scene_get_contacts(s.handle, offsets.data(), out.data(), n, &n), "scene_get_contacts");
  }
};

// ---------------------------------------------------------------- SymplecticSystem
template <typename PR>
struct SymplecticSystem<HIP, PR> {
  PR h = PR(1) / PR(60);
  size_t subSteps = 50;
  PR subh = h / subSteps;
  PR kair = PR(-0.001);
  PR kd = PR(0.5);
  PR acctime = 0;
  uint32_t anomalyFlag = 0;
  SymplecticSystem(PR h_ = PR(1) / PR(60), size_t subSteps_ = 50) : h(h_), subSteps(subSteps_), subh(h_ / subSteps_) {}
  void snapshotXPrev(Scene<HIP, PR>&) {}  // implicit: x is double-buffered on the device
  void update(Scene<HIP, PR>& s) {
    FastGridClothBehavior<HIP, PR>::update(s);
    acctime += subh;
  }
  void pollAnomaly(Scene<HIP, PR>& s) {
    uint32_t f = 0;
    ysim_hip_detail::check(ysim_hip_scene_flags(s.handle, &f), "scene_flags");
    anomalyFlag = f & 1u;
  }
};

// ---------------------------------------------------------------- Simulator
template <typename PR>
struct Simulator<HIP, PR> {
  SymplecticSystem<HIP, PR>& system;
  Scene<HIP, PR> scene;
  BVH<HIP, PR, SCENE, OBJECT> broadPhase;
  BruteForce<HIP, PR> narrowPhase;
  PR margin = PR(0.015);
  PR radius = PR(0.012);
  bool enableSelfCollisions = true;
  // true: the reference's cadence, a full rebuild at substep 0 of every 10th
  // frame and refits otherwise (simulator.hpp:3253); false: rebuild the
  // dynamic trees every substep
  bool enableRefit = true;
  uint32_t mortonBits = 30;
  explicit Simulator(SymplecticSystem<HIP, PR>& s) : system(s) {}

  // The scene parameters initialize() packs with.
  ysim_sim_params params() const {
    ysim_sim_params p{};
    p.subh = (float)system.subh;
    p.kair = (float)system.kair;
    p.kd = (float)system.kd;
    p.margin = (float)margin;
    p.radius = (float)radius;
    p.narrow_thickness = (float)margin;  // narrowAndSortByVertices(radius, margin) (simulator.hpp:3755)
    p.morton_bits = mortonBits;
    p.rebuild_every = enableRefit ? (uint32_t)(10 * system.subSteps) : 1u;
    p.self_collision = enableSelfCollisions ? 1u : 0u;
    p.swept = 1;
    return p;
  }

  // Simulator::initialize (simulator.hpp:2942-3115): pack + initial build.
  void initialize() {
    scene.pack(params());
    broadPhase.build(scene);
    narrowPhase.broad = &broadPhase;
  }

  // Launch tuning for this scene (no counterpart in the reference): times the
  // substep under each setting, keeps the fastest; state and results unchanged.
  ysim_tuning autotune(uint32_t trialSubsteps = 10) {
    ysim_tuning t{};
    ysim_hip_detail::check(ysim_hip_scene_autotune(scene.handle, trialSubsteps, &t), "scene_autotune");
    return t;
  }

  // One frame: the substep loop of Simulator::update (simulator.hpp:3512-3815).
  void update() {
    for (size_t i = 0; i < system.subSteps; ++i) {
      broadPhase.refit();
      broadPhase.enlargeTrajectory(system.subh);
      broadPhase.detectCollisions(margin, enableSelfCollisions, scene.analyticActive(), true);
      narrowPhase.narrowAndSortByVertices(radius, margin, scene.analyticActive());
      system.snapshotXPrev(scene);
      system.update(scene);
    }
    system.pollAnomaly(scene);  // the per-frame anomaly poll (simulator.hpp:3195-3203)
  }
};
