/*
 * ysim_hip.h — C ABI of the MI355X (gfx950) cloth-collision and spring-dynamics
 * library.  This is the whole drop-in boundary of the HIP backend: the C++
 * `HIP` backend tag (ysim_amd/host/ysim_hip_backend.hpp), the Python test
 * harness (ctypes) and the benchmark call only the functions declared here.
 *
 * Conventions
 *   - every entry point is `extern "C"`, takes plain pointers and sizes and
 *     returns an int status (YSIM_OK == 0); on failure the message is in
 *     ysim_hip_last_error() (thread-local);
 *   - pointers named d_* are device pointers obtained from ysim_hip_alloc;
 *     pointers named h_* are host pointers;
 *   - vectors are packed float3 (3 floats per vertex) and facets packed uint3
 *     with mesh-local vertex ids, as in the reference's packed scene buffers
 *     (include/sim/scene.hpp:390-422, 800-840);
 *   - a context owns one HIP stream; calls on one context are ordered, calls on
 *     distinct contexts are independent.  A context is not thread-safe.
 *
 * Reference interfaces each group replaces are cited per function
 * (paths relative to the yeduk3/ysim repository).
 */
#ifndef YSIM_HIP_H
#define YSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
  YSIM_OK = 0,
  YSIM_ERR_INVALID = 1,   /* bad argument / shape */
  YSIM_ERR_HIP = 2,       /* HIP runtime error */
  YSIM_ERR_CAPACITY = 3,  /* an output buffer was too small; required size reported */
  YSIM_ERR_UNSUPPORTED = 4,
  YSIM_ERR_STATE = 5      /* call out of order (e.g. refit before build) */
};

/* ------------------------------------------------------------------ PODs   */
/* Byte-identical to the reference kernel structs (src/metal/bvh.metal:7-24,
 * src/metal/common_types.metalh:4-17, include/sim/collision_pod.hpp:256-280,
 * 389-396).  static_asserts in the library pin the sizes. */
typedef struct ysim_aabb4 {        /* AABB4, include/sim/bvh.hpp:41-102 */
  float min[3]; int32_t i0;        /* i0 = numPrimitives when used as scene box */
  float max[3]; int32_t i1;
} ysim_aabb4;

typedef struct ysim_bvh_node {     /* BVHNode, src/metal/bvh.metal:14-19 */
  float min[3]; int32_t child_a;   /* child_a < 0 marks a leaf */
  float max[3]; int32_t child_b;   /* leaf: facet id (mesh-local) */
} ysim_bvh_node;

typedef struct ysim_broad_collision {  /* BroadCollision, 32 B */
  uint32_t point, triangle;            /* indexPair (mesh-local ids) */
  uint32_t query_obj, target_obj;      /* objPair (mesh array indices) */
  uint32_t query_behavior, target_behavior;
  uint32_t query_shape, target_shape;
} ysim_broad_collision;

typedef struct ysim_narrow_collision { /* NarrowCollision, 48 B */
  uint32_t point, triangle;
  uint32_t query_obj, target_obj;
  float normal_distance[4];            /* xyz: unit normal, w: signed distance */
  uint32_t query_behavior, target_behavior;
  uint32_t query_shape, target_shape;
} ysim_narrow_collision;

typedef struct ysim_fastgrid_params {  /* FastGridClothBehaviorParams<float>, 44 B */
  uint32_t particle_num_1d;
  float stretch_rest_x, stretch_rest_y;
  float shear_rest_a, shear_rest_b;
  float bend_rest_x, bend_rest_y;
  float kstretch, kshear, kbend;
  float thickness;                     /* contact-response gate (physics.metal:284-291) */
} ysim_fastgrid_params;

/* BehaviorType subset on this path (include/sim/mesh_state.hpp:108-121). */
enum {
  YSIM_BEHAVIOR_FLOAT = 0,            /* static obstacle: target only, never moves */
  YSIM_BEHAVIOR_FASTGRID_CLOTH = 1,   /* grid cloth: query + target, integrated */
  YSIM_BEHAVIOR_TRIANGULAR_CLOTH = 2  /* any manifold triangle mesh: springs along its edges (k_stretch)
                                       * and between the opposite vertices of every interior edge
                                       * (k_bend), rest lengths from the creation positions
                                       * (compute_tri_spring_forces, physics.metal:319-378) */
};

/* One mesh of a packed scene (the Scene::pack contract, scene.hpp:783-1182). */
typedef struct ysim_mesh_desc {
  uint32_t num_vertices;
  uint32_t num_facets;
  uint32_t behavior;                   /* YSIM_BEHAVIOR_* */
  uint32_t collidable;                 /* 0: leaves the broad phase entirely */
  uint32_t group;                      /* meshes interact only within one group (batching) */
  uint32_t shape;                      /* ColliderKind lane carried into pair records; 0 = Mesh */
  ysim_fastgrid_params cloth;          /* FASTGRID_CLOTH; TRIANGULAR_CLOTH uses kstretch, kbend and
                                        * thickness only */
} ysim_mesh_desc;

/* Per-scene stepping parameters: SymplecticSystem fields
 * (include/sim/symplectic_system.hpp:190-222) plus the Simulator collision knobs
 * (include/sim/simulator.hpp:121-181, bvh.hpp:2322-2409). */
typedef struct ysim_sim_params {
  float subh;                /* substep dt */
  float kair;                /* air drag coefficient (force += kair * v) */
  float kd;                  /* spring damping */
  float margin;              /* broad query half-width (Simulator::margin) */
  float radius;              /* narrow radius (Simulator::radius) */
  float narrow_thickness;    /* narrow band = radius + narrow_thickness, strict < */
  uint32_t morton_bits;      /* 30 or 63 */
  uint32_t rebuild_every;    /* 1: rebuild the dynamic LBVHs every substep; 0: build once then refit */
  uint32_t self_collision;   /* query a cloth against its own tree */
  uint32_t swept;            /* leaf = hull{x, x + v*subh} (enlargeTrajectory) */
  uint32_t broad_mode;       /* 0: LBVH traversal, 1: spatial hash (same candidate set) */
  uint32_t hash_slots_log2;  /* spatial hash table size = 2^hash_slots_log2 */
  float hash_cell;           /* spatial hash cell edge (0 = 2 * margin) */
  uint32_t contact_capacity; /* contacts kept per vertex in the fused path (0 = default 16) */
} ysim_sim_params;

/* Stages timed by ysim_hip_scene_stage_timings, in execution order. */
enum {
  YSIM_STAGE_BOX = 0,        /* S0 scene box */
  YSIM_STAGE_MORTON = 1,     /* S1 Morton codes (+ sort digit histograms) */
  YSIM_STAGE_SORT = 2,       /* S2 LSD radix sort (all passes) */
  YSIM_STAGE_HIERARCHY = 3,  /* S3 Karras hierarchy */
  YSIM_STAGE_REFIT = 4,      /* S4+S5 swept leaves + atomic bottom-up */
  YSIM_STAGE_QUERY = 5,      /* S6+S7 traversal fused with the narrow phase */
  YSIM_STAGE_INTEGRATE = 6,  /* S9 springs + integrate + contact projection */
  YSIM_STAGE_COUNT = 7
};

typedef struct ysim_ctx ysim_ctx;
typedef struct ysim_scene ysim_scene;
typedef struct ysim_batch ysim_batch;

/* ------------------------------------------------------------------ runtime
 * Replaces MetalGlobalContext / GlobalAutoAllocator
 * (include/sim/core_types.hpp:83-187, 497-529). */
/* ctx flags: YSIM_CTX_FORMAL_REFIT publishes bottom-up boxes with acq_rel
 * device-scope atomics (L2 writeback/invalidate) instead of device-coherent
 * stores + wait; both give identical results, the default is faster. */
#define YSIM_CTX_FORMAL_REFIT 1u
const char* ysim_hip_last_error(void);
int ysim_hip_version(void);
int ysim_hip_device_count(int* out);
int ysim_hip_ctx_create(int device, uint32_t flags, ysim_ctx** out);
/* Destroy calls wait for the stream, then release everything; a failure the
 * stream reports is returned after the release. */
int ysim_hip_ctx_destroy(ysim_ctx* ctx);
int ysim_hip_sync(ysim_ctx* ctx);
int ysim_hip_alloc(ysim_ctx* ctx, size_t bytes, void** d_out);
int ysim_hip_free(ysim_ctx* ctx, void* d_ptr);
int ysim_hip_upload(ysim_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int ysim_hip_download(ysim_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
int ysim_hip_memset(ysim_ctx* ctx, void* d_dst, int value, size_t bytes);

/* ------------------------------------------------------------------ stages
 * Scene box over packed float3 positions: BVH::build's CPU loop
 * (include/sim/bvh.hpp:1229-1231).  Writes one ysim_aabb4 with i0 = num_prims. */
int ysim_hip_scene_box(ysim_ctx* ctx, const float* d_x, uint32_t num_vertices,
                       uint32_t num_prims, ysim_aabb4* d_box);

/* Morton codes of triangle centroids: fillMortons_Tri (src/metal/bvh.metal:26-63).
 * bits == 30: d_keys is uint32_t[n]; bits == 63: d_keys is uint64_t[n].
 * d_idx[i] = i. */
int ysim_hip_morton(ysim_ctx* ctx, const float* d_x, const uint32_t* d_facets,
                    uint32_t num_facets, const ysim_aabb4* d_box, uint32_t bits,
                    void* d_keys, uint32_t* d_idx);

/* Stable LSD radix sort of (key, value) pairs in place: RadixSorter<METAL>::sort
 * (include/sim/spatial_hash.hpp:48-70, src/metal/radixsort.metal:21-112).
 * key_bits <= 32: uint32_t keys; key_bits <= 64: uint64_t keys.  Ties keep
 * the input order. */
int ysim_hip_sort_pairs(ysim_ctx* ctx, void* d_keys, uint32_t* d_vals, uint32_t n,
                        uint32_t key_bits);
/* The same sort under an explicit configuration (results identical for all):
 * digit_bits 8 or 11 and items (keys per thread per tile) 8 or 16 for the LSD
 * onesweep; hybrid = 1 selects one 11-bit MSD pass + per-bucket LSD in LDS. */
int ysim_hip_sort_pairs_ex(ysim_ctx* ctx, void* d_keys, uint32_t* d_vals, uint32_t n, uint32_t key_bits,
                           uint32_t digit_bits, uint32_t items, uint32_t hybrid);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  ysim_hip_sort_pairs (same plan, same kernels, same
 * permutation) with its scratch in a static device array of
 * YSIM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: it calls no
 * allocator.  Inputs needing more scratch than that (about 2^20 keys of
 * 64 bits, more of 32) return YSIM_ERR_CAPACITY before any launch.  All calls
 * on one GPU share the array, so calls that overlap on one GPU (two host
 * threads, two contexts of the same device) race: serialize them. */
#define YSIM_NOALLOC_SCRATCH_BYTES (24u << 20)
int ysim_hip_sort_pairs_noalloc(ysim_ctx* ctx, void* d_keys, uint32_t* d_vals, uint32_t n,
                                uint32_t key_bits);

/* Karras hierarchy over sorted codes: buildTree_Tri (src/metal/bvh.metal:197-258,
 * 352-393).  Writes child links of the 2n-1 node slots and d_parent
 * (root parent = -1); leaf slots get child_a = -1, child_b = facet id.
 * Boxes are left for ysim_hip_lbvh_refit. */
int ysim_hip_lbvh_hierarchy(ysim_ctx* ctx, const void* d_sorted_keys,
                            const uint32_t* d_sorted_idx, uint32_t n, uint32_t bits,
                            ysim_bvh_node* d_tree, int32_t* d_parent);

/* Agglomerative (Apetrei 2014) build over sorted codes: agglomerativeBuild_Tri +
 * agglomerativeSwapRoot (src/metal/bvh.metal:884-1011, 1135-1199), host
 * BVH::agglomerativeBuildGPU / agglomerativeSwapRootGPU (include/sim/bvh.hpp:
 * 649-694).  Child links, d_parent and the boxes of the 2n-1 slots in one pass
 * (leaf box = facet hull of d_x); the same tree as lbvh_hierarchy + lbvh_refit
 * with internal nodes numbered by their split and the root moved to slot 0. */
int ysim_hip_lbvh_agglomerative(ysim_ctx* ctx, const void* d_sorted_keys,
                                const uint32_t* d_sorted_idx, uint32_t n, uint32_t bits,
                                const float* d_x, const uint32_t* d_facets,
                                ysim_bvh_node* d_tree, int32_t* d_parent);

/* Leaf boxes + atomic bottom-up union: buildLeaf_Tri / buildSweptLeaf_Tri +
 * bottomUpBoxes (src/metal/bvh.metal:260-351, 464-586).  swept != 0 unions each
 * vertex with x + v*dt (enlargeTrajectory, bvh.hpp:1608-1682). */
int ysim_hip_lbvh_refit(ysim_ctx* ctx, const float* d_x, const float* d_v, float dt,
                        uint32_t swept, const uint32_t* d_facets, uint32_t num_facets,
                        ysim_bvh_node* d_tree, const int32_t* d_parent);

/* Vertex-vs-triangle broad query of one (query mesh, target tree) pair:
 * queryPoints/queryAABB (src/metal/bvh.metal:1285-1412), host BVH::queryPoints
 * (include/sim/bvh.hpp:1803-1841).  Appends to d_out starting at *d_count;
 * never truncates silently: if the total exceeds capacity, returns
 * YSIM_ERR_CAPACITY and *h_required receives the needed size. */
int ysim_hip_broad_query(ysim_ctx* ctx, const float* d_xq, uint32_t num_query,
                         uint32_t query_obj, uint32_t query_behavior, uint32_t query_shape,
                         const ysim_bvh_node* d_tree, uint32_t num_target_facets,
                         const uint32_t* d_target_facets, uint32_t target_obj,
                         uint32_t target_behavior, uint32_t target_shape,
                         float margin, ysim_broad_collision* d_out, uint32_t capacity,
                         uint32_t* d_count, uint32_t* h_required);

/* The same query with the segmented emission of queryPointsSegmented +
 * scanReserveSegmented + compactSegmented (src/metal/bvh.metal:1732-1827,
 * host BVH::queryPointsSegmented): each workgroup stages its pairs in LDS and
 * reserves its output range with one atomic.  Same pair set and count; the
 * records are grouped by workgroup. */
int ysim_hip_broad_query_segmented(ysim_ctx* ctx, const float* d_xq, uint32_t num_query,
                                   uint32_t query_obj, uint32_t query_behavior, uint32_t query_shape,
                                   const ysim_bvh_node* d_tree, uint32_t num_target_facets,
                                   const uint32_t* d_target_facets, uint32_t target_obj,
                                   uint32_t target_behavior, uint32_t target_shape,
                                   float margin, ysim_broad_collision* d_out, uint32_t capacity,
                                   uint32_t* d_count, uint32_t* h_required);

/* Swept point-triangle narrow phase: narrow_pt_tri (src/metal/bruteforce.metal:9-178),
 * host BruteForce::narrow (include/sim/bruteforce.hpp:114-153).  Positions are
 * the packed scene arrays indexed through d_states_offsets / d_facets_offsets. */
int ysim_hip_narrow(ysim_ctx* ctx, const ysim_broad_collision* d_pairs, uint32_t num_pairs,
                    const float* d_x, const float* d_xprev, const uint32_t* d_states_offsets,
                    const uint32_t* d_facets, const uint32_t* d_facets_offsets,
                    const uint32_t* d_adj_offsets, const uint32_t* d_adj_facets,
                    float radius, float thickness,
                    ysim_narrow_collision* d_out, uint32_t capacity, uint32_t* d_count,
                    uint32_t* h_required);

/* Per-vertex contact CSR: narrow_bucket_zero/count/scan/scatter
 * (src/metal/physics.metal:546-616), CPU twin bruteforce.hpp:358-383.  Buckets
 * are ordered by (target_obj, triangle) so the response is deterministic. */
int ysim_hip_contact_csr(ysim_ctx* ctx, const ysim_narrow_collision* d_contacts,
                         uint32_t num_contacts, const uint32_t* d_states_offsets,
                         uint32_t num_points_total, uint32_t* d_offsets /* V+1 */,
                         ysim_narrow_collision* d_sorted);

/* FastGrid springs + symplectic integrate + contact projection for one cloth
 * mesh: compute_cloth_grid_forces_fast + integrate_cloth_grid
 * (src/metal/physics.metal:146-312).  Reads (d_x, d_v), writes (d_x_out, d_v_out)
 * (distinct buffers), d_f_out optional (may be NULL).  d_anomaly: uint32 flag set
 * when the world-bounds guard fires. */
int ysim_hip_cloth_step(ysim_ctx* ctx, const ysim_fastgrid_params* h_params,
                        uint32_t num_vertices, float subh, float kair, float kd,
                        const float* d_x, const float* d_v, const float* d_ext,
                        const float* d_mass, const float* d_fixed,
                        const uint32_t* d_csr_offsets, const ysim_narrow_collision* d_csr,
                        float* d_x_out, float* d_v_out, float* d_f_out,
                        uint32_t* d_anomaly);

/* ------------------------------------------------------------------ scene
 * The Simulator::initialize / update substep loop (include/sim/simulator.hpp:
 * 2942-3115, 3512-3815) over a packed scene resident in HBM. */
int ysim_hip_scene_create(ysim_ctx* ctx, uint32_t num_meshes, const ysim_mesh_desc* h_meshes,
                          const float* h_x, const float* h_v, const float* h_mass,
                          const float* h_ext, const float* h_fixed, const uint32_t* h_facets,
                          const ysim_sim_params* h_params, ysim_scene** out);
int ysim_hip_scene_destroy(ysim_scene* scene);
/* Run n substeps (eager launches, or the captured graph when use_graph != 0). */
int ysim_hip_scene_step(ysim_scene* scene, uint32_t n, uint32_t use_graph);
/* The substep as its three reference stages, for callers that drive them one by
 * one (BVH::refit/refitSwept, detectCollisions + narrowAndSortByVertices,
 * SymplecticSystem::update; simulator.hpp:3624-3805).  They must be called in
 * this order once per substep (YSIM_ERR_STATE otherwise); scene_step runs all
 * three.  refit rebuilds the dynamic trees when the rebuild cadence says so;
 * query runs the traversal fused with the narrow phase and leaves the
 * per-vertex contacts; integrate runs springs + integrate + contacts and
 * advances the substep (xPrev := x is implicit). */
int ysim_hip_scene_refit(ysim_scene* scene);
int ysim_hip_scene_query(ysim_scene* scene);
int ysim_hip_scene_integrate(ysim_scene* scene);
/* Copy current x, v (packed float3) and xPrev to host; any pointer may be NULL. */
int ysim_hip_scene_get_state(ysim_scene* scene, float* h_x, float* h_v, float* h_xprev);
int ysim_hip_scene_set_state(ysim_scene* scene, const float* h_x, const float* h_v,
                             const float* h_xprev);
/* Contacts of the last substep as a CSR ordered by (vertex, target_obj, triangle). */
int ysim_hip_scene_get_contacts(ysim_scene* scene, uint32_t* h_offsets /* Vtot+1 */,
                                ysim_narrow_collision* h_contacts, uint32_t capacity,
                                uint32_t* h_count);
/* Broad candidate pairs at the current state (collision pipeline only, no step),
 * unordered; caller sorts to compare as a set. */
int ysim_hip_scene_get_pairs(ysim_scene* scene, ysim_broad_collision* h_pairs,
                             uint32_t capacity, uint32_t* h_count);
/* Tree of one mesh (2T-1 nodes) and the sorted Morton keys (uint64 widened).
 * Keys and order are kept for the last dynamic build group only (its meshes:
 * every dynamic mesh when one group holds them); for a mesh of another group
 * asking for them returns YSIM_ERR_STATE. */
int ysim_hip_scene_get_tree(ysim_scene* scene, uint32_t mesh, ysim_bvh_node* h_nodes,
                            int32_t* h_parent, uint64_t* h_sorted_keys, uint32_t* h_sorted_idx);
/* Per-stage average device time (ms) over the last n_steps eager substeps run
 * with timing enabled; out has YSIM_STAGE_COUNT entries.  With one dynamic
 * mesh (tile query, no analytic colliders) one kernel runs the query and the
 * integrate: the query entry covers both and the integrate entry is ~0; the
 * box entry is ~0 too while the cloth step supplies the Morton box. */
int ysim_hip_scene_stage_timings(ysim_scene* scene, uint32_t n_steps, float* h_ms);
/* Sticky flags: bit0 anomaly (world-bounds guard), bit1 contact overflow,
 * bit2 traversal guard. */
int ysim_hip_scene_flags(ysim_scene* scene, uint32_t* h_flags);
/* Device-side snapshot of the dynamic state (x, v, xPrev, substep counter)
 * and its restore: episodes of a fixed-length run without host transfers.
 * Both are stream-ordered device copies. */
int ysim_hip_scene_snapshot(ysim_scene* scene);
int ysim_hip_scene_restore(ysim_scene* scene);
/* Counters since the last reset: out[0] candidate pairs tested by the fused
 * query, out[1] substeps run, out[2] contacts of the last substep. */
int ysim_hip_scene_counters(ysim_scene* scene, uint64_t* h_out, uint32_t reset);
/* The per-call collision arguments of a substep: BVH::detectCollisions(margin,
 * self, ...) (include/sim/bvh.hpp:2853) and
 * BruteForce::narrowAndSortByVertices(radius, thickness, ...)
 * (include/sim/bruteforce.hpp:311).  Sets the scene's margin, radius,
 * narrow_thickness and self_collision (ysim_sim_params meanings; margin <= 0
 * keeps each cloth's own thickness) for this and later substeps; callable
 * before the substep's query (YSIM_ERR_STATE after it).  Unchanged values
 * cost nothing; new values are written into device arrays that a captured
 * graph reads, and a change of self_collision rebuilds the target lists and
 * drops the graph.  Hash scenes reject a margin above 7 cells. */
int ysim_hip_scene_set_collision(ysim_scene* scene, float margin, float radius, float narrow_thickness,
                                 uint32_t self_collision);
/* No counterpart in the reference (diagnostics).  Graph replays
 * (scene_step with use_graph) leave out the incremental sort's fallback
 * kernels: a rebuild whose order check fails keeps the previous topology,
 * refit to the new positions (the candidate sets, contacts and states are
 * unchanged; scene_get_tree shows that topology), and the next scene_step
 * call runs its first substep eagerly, with the fallback.  A state change
 * through this API (set_state, restore, tuning) does the same without a
 * failed check.  *h_out: how many replays kept a previous topology. */
int ysim_hip_scene_deferred_builds(ysim_scene* scene, uint64_t* h_out);
/* No counterpart in the reference (parity audit of the kernel a substep
 * actually runs).  With capacity > 0, every later substep's query (the tile
 * walk fused with the cloth step, the per-vertex walk or the hash query,
 * whichever the tuning selects; eager or graph replay) also records each
 * candidate pair it reports, into a device buffer of `capacity` records
 * (16 B each) that each substep's query clears first.  capacity 0 turns the
 * audit off and frees the buffer.  Turning it on or off drops the captured
 * graph. */
int ysim_hip_scene_set_audit(ysim_scene* scene, uint32_t capacity);
/* The candidate pairs recorded by the last audited query, as BroadCollision
 * records sorted by (query_obj, point, target_obj, triangle), duplicates kept
 * (a correct walk reports each pair once).  The records of a tile walk that
 * overflowed and was redone per lane are dropped (*h_redo, optional, counts
 * those walks).  YSIM_ERR_CAPACITY when the device buffer overflowed
 * (*h_count = records needed: set_audit with that and run the substep again)
 * or when `capacity` is smaller than *h_count; h_pairs may be null to ask for
 * the count.  YSIM_ERR_STATE when the audit is off. */
int ysim_hip_scene_get_audit(ysim_scene* scene, ysim_broad_collision* h_pairs, uint32_t capacity,
                             uint32_t* h_count, uint32_t* h_redo);
/* No counterpart in the reference (the tile walk's start frontier).  The tree
 * cut of one mesh that its last Karras build listed: *h_count nodes (0 when
 * no cut was listed; the walk then starts from the root), written to h_nodes
 * when capacity allows (YSIM_CUT_CAP at most). */
#define YSIM_CUT_CAP 512u
int ysim_hip_scene_get_cut(ysim_scene* scene, uint32_t mesh, int32_t* h_nodes, uint32_t capacity,
                           uint32_t* h_count);

/* ---- analytic colliders -----------------------------------------------------
 * ColliderKind (include/sim/collision_pod.hpp:32, common_types.metalh:80-84):
 * a mesh whose collider kind is not Mesh answers contacts from its shape
 * parameters instead of its triangles. */
enum {
  YSIM_COLLIDER_MESH = 0,
  YSIM_COLLIDER_SPHERE = 1,    /* ellipsoid with semi-axes half_ext_height.xyz (a sphere reduces exactly) */
  YSIM_COLLIDER_BOX = 2,       /* half-extents half_ext_height.xyz */
  YSIM_COLLIDER_CYLINDER = 3,  /* local +Y axis, radius center_radius.w, half-height half_ext_height.w */
  YSIM_COLLIDER_PLANE = 4      /* infinite half-space: outside is local +Y */
};
/* AnalyticShape (include/sim/collision_pod.hpp:285-310, the metal mirror
 * common_types.metalh:42-52), 96 B.  Quaternions are (w, x, y, z), local to
 * world; prev_* hold the transform at the start of the substep. */
typedef struct ysim_analytic_shape {
  float center_radius[4];    /* xyz world centre, w radius (cylinder) */
  float half_ext_height[4];  /* xyz half-extents / semi-axes, w half-height (cylinder) */
  float rot_quat[4];
  float prev_center_pad[4];  /* xyz previous centre, w unused */
  float prev_rot_quat[4];
  uint32_t kind;             /* YSIM_COLLIDER_SPHERE .. YSIM_COLLIDER_PLANE */
  uint32_t obj_index;        /* the collider's mesh (array index); must be a FLOAT mesh whose
                              * desc.shape equals kind; at most one shape per mesh */
  uint32_t behavior_type;    /* must equal the mesh's behavior */
  uint32_t flags;            /* bit0 collidable: 0 keeps the mesh out of the triangle path
                              * without giving contacts */
} ysim_analytic_shape;
/* Install (or replace) the scene's analytic colliders: the BVH broad phase's
 * analytic branch (BroadPhase::detectCollisions, include/sim/bvh.hpp:2854-2949,
 * objectBroadAABB 2486-2518) and BruteForce::narrowAnalytic /
 * narrow_pt_analytic (include/sim/bruteforce.hpp:192-308,
 * src/metal/bruteforce.metal:198-499).  Each substep a cloth (query mesh of
 * the same group, collidable) is paired with each shape whose mesh is
 * collidable when the cloth's tree root box meets the shape's swept box padded
 * by the cloth's narrow band plus the rotation band (a plane always pairs),
 * or when it did at the previous substep; every vertex of a paired cloth then
 * runs the swept motion-inverted test against the shape and keeps one contact
 * (target_obj = obj_index, triangle = 0).  The shapes' meshes leave the
 * triangle traversal.  A call with the same obj_index sequence updates poses
 * (moving colliders: pass the previous pose in prev_*) and keeps the
 * previous-substep pairing; another sequence resets it; n = 0 removes them
 * all.  With broad_mode 1 (spatial hash) the shapes are stored but inactive
 * and the meshes collide through their triangles, as in the reference.  In a
 * staged substep it may be called until scene_query (YSIM_ERR_STATE after). */
int ysim_hip_scene_set_analytic(ysim_scene* scene, const ysim_analytic_shape* h_shapes, uint32_t n);

/* ---- measurement ------------------------------------------------------------
 * No counterpart in the reference.  Achievable HBM bandwidth: a plain
 * 16-byte-vector copy kernel ("k_copy_u4") of `bytes` from one device buffer to
 * another, `reps` times after one warm-up; *h_ms = mean duration per copy
 * (HIP events).  Also the calibration kernel of known traffic (bytes read and
 * bytes written) for hardware counter runs. */
int ysim_hip_copy_bandwidth(ysim_ctx* ctx, size_t bytes, uint32_t reps, float* h_ms);

/* ---- tuning -----------------------------------------------------------------
 * No counterpart in the reference (its kernels use fixed threadgroup sizes and
 * 8-bit digits, radixsort.metal:21-112).  Launch parameters of the scene's
 * kernels; every setting gives bit-identical results. */
typedef struct ysim_tuning {
  uint32_t sort_digit_bits;  /* radix digit width: 8 or 11 */
  uint32_t sort_items;       /* keys per thread of a sort tile (tile = 256 x items): 8 or 16 */
  uint32_t query_block;      /* traversal + narrow threads per block: 64, 128 or 256 */
  uint32_t cloth_block;      /* springs + integrate threads per block: 128, 256 or 512 */
  float ms_per_substep;      /* autotune: the chosen setting's measured time; else 0 */
  float ms_per_substep_default; /* autotune: the default setting's measured time */
  uint32_t refit_block;      /* bottom-up refit leaves per block (LDS levels): 256, 512 or 1024 */
  uint32_t sort_hybrid;      /* 1: one 11-bit MSD pass + per-bucket LSD in LDS (sort_digit_bits unused) */
  uint32_t sort_incremental; /* 1: keys in the last build's sorted order, LDS window sorts and an exact
                              * check; the hybrid sort runs only when the check fails */
  uint32_t query_waves;      /* traversal occupancy target, waves per SIMD: 5, 6 or 8 */
  uint32_t query_tile;       /* 1: one wave per 64-vertex tile walks the tree breadth-first with the
                              * tile's union box (default); 0: one thread per vertex (query_block and
                              * query_waves apply); 2: as 1 with 128-leaf and 32-node tile buffers, so
                              * that the tiles' buffer flushes and per-vertex redos run (tests those) */
  uint32_t build_mode;       /* rebuilds: 0 Karras hierarchy + refit (default); 1 agglomerative (Apetrei)
                              * build with the leaf boxes fused in (one kernel), refits between rebuilds
                              * through parent links.  Same tree, nodes numbered by their split */
} ysim_tuning;
int ysim_hip_scene_get_tuning(ysim_scene* scene, ysim_tuning* h_out);
int ysim_hip_scene_set_tuning(ysim_scene* scene, const ysim_tuning* h_in);
/* Time `trial_substeps` substeps (hipGraph) of the scene from its current state
 * under each candidate setting (coordinate descent over the four parameters),
 * apply the fastest, and restore the state (positions, velocities, step count,
 * flags and counters) so the results of later substeps are unchanged. */
int ysim_hip_scene_autotune(ysim_scene* scene, uint32_t trial_substeps, ysim_tuning* h_out);

/* ---- differentiation ------------------------------------------------------
 * No counterpart in the reference (it has no adjoint).  The differentiated map
 * is its substep (Simulator::update, include/sim/simulator.hpp:3512-3815) for a
 * scene with exactly one FastGrid cloth whose contacts are against static
 * meshes (self_collision = 0): springs with one stiffness per spring
 * (k = exp(log_k), spring order struct-x, struct-y, shear-a, shear-b, bend-x,
 * bend-y, each row-major over its first endpoint), damping kd, symplectic
 * integration and contact projection (src/metal/physics.metal:171-312).  The
 * derivative holds each step's contact set and its gates (distance <
 * thickness, v.n < 0) fixed.  Device pointers throughout; x/v arrays are the
 * cloth's own 3 * num_vertices floats. */
int ysim_hip_adj_num_springs(ysim_scene* scene, uint32_t* h_out);
/* Run `steps` substeps from the scene's current state with k = exp(d_log_k)
 * and damping kd (both stay in effect for later scene_step calls), keeping
 * (x, v, xPrev) checkpoints every `checkpoint_every` substeps on the device. */
int ysim_hip_adj_forward(ysim_scene* scene, const float* d_log_k, float kd, uint32_t steps,
                         uint32_t checkpoint_every);
/* Given dL/dx_N (and dL/dv_N, may be NULL = 0) at the end of the last
 * adj_forward, write dL/dlog_k (num_springs floats), dL/dkd (1 float),
 * dL/dx_0 and dL/dv_0 (3 * num_vertices floats each); any output may be NULL.
 * Each segment is recomputed from its checkpoint; the scene is left in the
 * state the forward ended in. */
int ysim_hip_adj_backward(ysim_scene* scene, const float* d_x_bar, const float* d_v_bar, float* d_g_log_k,
                          float* d_g_kd, float* d_g_x0, float* d_g_v0);

#ifdef __cplusplus
}
#endif
#endif /* YSIM_HIP_H */
