// kernels.hpp — device-side tables and host launchers of the LBVH, query and
// cloth kernels.  Every kernel is "segmented": one launch covers all meshes of
// a scene (one segment per mesh), so the plain single-cloth substep and the
// batched multi-instance substep run the same code.
#pragma once
#include "audit.hpp"
#include "common.hpp"
#include "sort.hpp"

namespace ysim {

// A ragged range table: items [off[k], off[k+1]) belong to segment k, whose
// mesh is mesh[k].
struct SegTable {
  uint32_t nseg = 0;
  const uint32_t* off = nullptr;   // device, nseg + 1
  const uint32_t* mesh = nullptr;  // device, nseg
  // One-segment tables filled by the host (one_mesh != ~0u): the segment's
  // mesh, first item, item count and the mesh's node / facet / vertex offsets,
  // passed with the kernel arguments so no table load precedes a thread's
  // first data load.
  uint32_t one_mesh = ~0u, one_s = 0, one_n = 0, one_noff = 0, one_foff = 0, one_voff = 0;
};

__device__ __forceinline__ uint32_t seg_of(const SegTable& s, uint32_t j) {
  if (s.nseg <= 1) return 0;
  uint32_t lo = 0, hi = s.nseg;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (s.off[mid] <= j) lo = mid; else hi = mid;
  }
  return lo;
}

// Per-mesh offsets into the packed scene arrays (device, nm + 1 each).
struct MeshTables {
  const uint32_t* voff = nullptr;  // vertices
  const uint32_t* foff = nullptr;  // facets
  const uint32_t* noff = nullptr;  // BVH node windows (2T - 1 per mesh)
};

// Item j's segment with its mesh and that mesh's offsets: from the kernel
// arguments for a host-filled one-segment table (one cloth), else a per-lane
// binary search, then the mesh and its offsets.
struct SegInfo {
  uint32_t sg, m, s, n;       // segment, mesh, first item, items
  uint32_t noff, foff, voff;  // the mesh's node window, facets, vertices
};
__device__ __forceinline__ SegInfo seg_info(const SegTable& t, const MeshTables& mt, uint32_t j) {
  if (t.nseg <= 1 && t.one_mesh != ~0u)
    return {0u, t.one_mesh, t.one_s, t.one_n, t.one_noff, t.one_foff, t.one_voff};
  if (t.nseg <= 1) {
    const uint32_t m = t.mesh[0], s = t.off[0];
    return {0u, m, s, t.off[1] - s, mt.noff[m], mt.foff[m], mt.voff[m]};
  }
  const uint32_t sg = seg_of(t, j);
  const uint32_t m = t.mesh[sg], s = t.off[sg];
  return {sg, m, s, t.off[sg + 1] - s, mt.noff[m], mt.foff[m], mt.voff[m]};
}

// Box of the triangle at f (three vertex ids from vb on); when swept, also of
// x + v dt per vertex.  The ids load as one 12-byte record and every position
// (and velocity) load is issued before the first use, so the gather takes two
// dependent rounds.  Union order p0, q0, p1, q1, p2, q2 (q: swept point).
struct __attribute__((packed, aligned(4))) pu3 {
  uint32_t a, b, c;
};
// tri_box_ids: the same from the three vertex ids.
__device__ __forceinline__ void tri_box_ids(const float* __restrict__ x, const float* __restrict__ v, float dt,
                                            bool swept, uint32_t a, uint32_t b, uint32_t c, f3& lo, f3& hi) {
  const f3 p0 = ld3(x, a), p1 = ld3(x, b), p2 = ld3(x, c);
  if (swept) {
    const f3 v0 = ld3(v, a), v1 = ld3(v, b), v2 = ld3(v, c);
    const f3 q0 = p0 + v0 * dt, q1 = p1 + v1 * dt, q2 = p2 + v2 * dt;
    lo = vmin(vmin(vmin(vmin(vmin(p0, q0), p1), q1), p2), q2);
    hi = vmax(vmax(vmax(vmax(vmax(p0, q0), p1), q1), p2), q2);
  } else {
    lo = vmin(vmin(p0, p1), p2);
    hi = vmax(vmax(p0, p1), p2);
  }
}
__device__ __forceinline__ void tri_box(const float* __restrict__ x, const float* __restrict__ v, float dt,
                                        bool swept, const uint32_t* f, uint32_t vb, f3& lo, f3& hi) {
  const pu3 t = *reinterpret_cast<const pu3*>(f);
  tri_box_ids(x, v, dt, swept, vb + t.a, vb + t.b, vb + t.c, lo, hi);
}

// ---------------------------------------------------------------- lbvh.hip
// S0: per-segment exact min/max of vertex positions into box_acc[seg*6] (ordered keys).
// single_mesh: the mesh id when vsegs is that one mesh (a kernel with fewer
// atomics), else 0xffffffff.  parts (single mesh only): box_parts_words(n)
// words, an arrival counter (word 0, zero before the first use) and per-block
// boxes;
// the last block reduces them (one same-address atomic per block).
uint32_t box_parts_words(uint32_t nverts);
int launch_box(hipStream_t s, SegTable vsegs, MeshTables mt, const float* x, uint32_t nitems,
               uint32_t* box_acc, uint32_t single_mesh = 0xffffffffu, uint32_t* parts = nullptr);
// S1: Morton keys of every triangle of tsegs (segment id prefixed above the code
// bits when nseg > 1), values = mesh-local facet id, plus the digit histograms
// of all sort passes of `plan`; zeroes the sort's block counters and advances
// its look-back epoch (either may be null when no sort follows).
// run_flag (optional): the kernel does nothing while *run_flag == 0.
// cut_reset (optional, per mesh): zeroed for the group's meshes; cut_fill
// (optional, the cut lists at CUT_CAP per mesh): a one-segment group's list
// is filled with -1, so readers may load entries before the count.
int launch_morton(hipStream_t s, SegTable tsegs, MeshTables mt, const float* x,
                  const uint32_t* facets, const uint32_t* box_acc, uint32_t code_bits,
                  bool key64, void* keys, uint32_t* vals, uint32_t nitems, const SortPlan& plan,
                  uint32_t* hist, uint32_t* counters, uint32_t* epoch, const uint32_t* run_flag = nullptr,
                  uint32_t* cut_reset = nullptr, int32_t* cut_fill = nullptr);
// Incremental S1: keys in the order perm (previous sorted facets), each
// 2048-key window sorted by (key, facet); zeroes *need_full.
// With nparts > 0 (one segment only) the box is the reduction of nparts
// 6-word records at box_parts (the cloth kernel's per-block boxes); every
// block reduces them and block 0 writes the box to box_acc for the build's
// later kernels.  Otherwise box_acc holds the box per segment.
// cut_reset (per mesh): zeroed for the group's meshes (their build starts);
// cut_fill as for launch_morton.
int launch_morton_inc(hipStream_t s, SegTable tsegs, MeshTables mt, const float* x, const uint32_t* facets,
                      uint32_t* box_acc, uint32_t code_bits, bool key64, const uint32_t* perm, void* keys,
                      uint32_t* vals, uint32_t nitems, uint32_t* need_full, const uint32_t* box_parts = nullptr,
                      uint32_t nparts = 0, uint32_t* cut_reset = nullptr, int32_t* cut_fill = nullptr);
// Extra links of a scene build for the LDS refit, written by the hierarchy
// kernel (null: not written, and the refit takes the device-wide path only).
// With fid set, the hierarchy writes no tree nodes: the LDS refit that follows
// writes each node whole (box and child links) in one pass, so no node is
// written in part twice per rebuild.
struct RefitLinks {
  // [node window + node] parent's leaf count << 2 | (the parent's index is
  // the high end of its leaf range) << 1 | (node is child_b): with the node's
  // own range this gives the parent's range and index (a Karras node is an
  // endpoint of its range)
  uint32_t* up = nullptr;
  uint32_t* fid = nullptr;  // [node window + leaf slot] the leaf's facet id
  // Tree cut per mesh: the nodes of at most S = ceil(leaves / 128) leaves whose
  // parent has more (a frontier that covers every leaf once), appended in any
  // order to cut + CUT_CAP mesh, counted in cut_cnt[mesh] (past CUT_CAP: not
  // listed, and readers walk from the root).  The build's Morton kernel zeroes
  // the group's counts (and fills a one-segment group's list with -1), so a
  // build that writes no cut leaves them at zero.
  uint32_t* cut_cnt = nullptr;
  int32_t* cut = nullptr;
  // Nodes spanning refit blocks of span_rb sorted leaves (span_rb 0: not
  // recorded).  Position space: the leaf positions of the launch's segment
  // table, which the refit that reads them must share.  span_mask: one bit
  // per position (a node at its own index's position), one word per 64;
  // span_rec[position]: (lo, hi, child_a, child_b), mesh-relative, written for
  // spanning nodes only.
  uint32_t span_rb = 0;
  uint64_t* span_mask = nullptr;
  int4* span_rec = nullptr;
  // [node window + internal node] the other end of the node's leaf range and
  // its split at the last hierarchy build of these links, or (-1, -1), which
  // never passes: checked before the searches, and a node whose hint passes
  // has the children it had then, so it rewrites neither the hint nor its
  // children's parent and `up` links.  Any other writer of the parent links
  // (the agglomerative build) resets the hints to (-1, -1).
  int2* hint = nullptr;
};
constexpr uint32_t CUT_CAP = 512;
// Partial boxes of the two-kernel refit (span_rb as above): part[position]
// two float4 per leaf position (only positions that start or end a spanning
// node's range are written), blk[block] the union of a block's leaves.
struct RefitSpan {
  const uint64_t* mask = nullptr;
  const int4* rec = nullptr;
  float4* part = nullptr;
  float4* blk = nullptr;
};
// S3: Karras hierarchy per segment; child and parent links (and rl.up when
// set); zeroes the refit visit counters; resets box_acc (nbox segments) and the
// sort histogram for the next build.
// Optional gate of the hierarchy: when *flag is set (the incremental sort's
// order check failed and its fallback was left out) the kernel only does its
// resets and keeps the previous topology, sorted order and links, and writes 1
// to *note (host-visible, system scope).
struct HierarchySkip {
  const uint32_t* flag = nullptr;
  uint32_t* note = nullptr;
};
int launch_hierarchy(hipStream_t s, SegTable tsegs, MeshTables mt, const void* keys, bool key64,
                     const uint32_t* vals, uint32_t code_bits, ysim_bvh_node* tree, int32_t* parent,
                     uint32_t* visit, uint32_t nitems, uint32_t* box_acc, uint32_t nbox,
                     uint32_t* hist, uint32_t hist_words, uint32_t* perm_out = nullptr,
                     RefitLinks rl = RefitLinks{}, HierarchySkip skip = HierarchySkip{});
// S3 alternative: agglomerative (Apetrei) hierarchy fused with the leaf boxes
// (swept optional) and their bottom-up union, root swapped into slot 0;
// writes child and parent links (no refit links), the sorted facet order and
// the box / histogram resets of launch_hierarchy.  range: per node window, two
// ints of scratch.  Bit 2 of *err on a corrupt key order.
int launch_agglo(hipStream_t s, SegTable tsegs, MeshTables mt, const void* keys, bool key64, const uint32_t* vals,
                 uint32_t code_bits, const float* x, const float* v, float dt, bool swept, const uint32_t* facets,
                 ysim_bvh_node* tree, int32_t* parent, uint32_t* visit, int2* range, uint32_t nitems,
                 uint32_t* box_acc, uint32_t nbox, uint32_t* hist, uint32_t hist_words, uint32_t* perm_out,
                 uint32_t* err, bool formal);
// S4+S5: leaf boxes (swept optional) + atomic bottom-up union per segment.
// With rl (from launch_hierarchy) the levels inside blocks of `block` (256,
// 512 or 1024) sorted leaves are combined in LDS.  With span (its mask and
// records written by the hierarchy kernel of this topology with span_rb ==
// block, over the same leaf positions) the nodes spanning blocks are
// finished by a second kernel from per-block partial boxes instead of the
// device-wide arrival protocol (not with formal).
bool refit_block_ok(uint32_t b);
int launch_refit(hipStream_t s, SegTable tsegs, MeshTables mt, const float* x, const float* v,
                 float dt, bool swept, const uint32_t* facets, ysim_bvh_node* tree,
                 const int32_t* parent, uint32_t* visit, uint32_t nitems, uint32_t* err, bool formal,
                 const RefitLinks* rl = nullptr, uint32_t block = 256, const RefitSpan* span = nullptr);

// ---------------------------------------------------------------- query.hip
struct QueryArgs {
  SegTable qsegs;                  // query vertex segments (one per dynamic mesh)
  MeshTables mt;
  const uint32_t* tgt_off = nullptr;  // per query segment: [tgt_off[k], tgt_off[k+1]) into tgt
  const uint32_t* tgt = nullptr;      // target mesh ids
  const uint32_t* mesh_nfacets = nullptr;
  const float* mesh_margin = nullptr; // per query mesh broad half-width
  const float* mesh_band = nullptr;   // per query mesh narrow band (radius + thickness)
  const uint32_t* mesh_behavior = nullptr;
  const uint32_t* mesh_shape = nullptr;
  const uint32_t* mesh_grid_n = nullptr; // grid side of FastGrid cloths (queries in tile order), else 0
  const float* x = nullptr;
  const float* xprev = nullptr;
  const uint32_t* facets = nullptr;
  const ysim_bvh_node* tree = nullptr;
  const int32_t* parent = nullptr;
  uint32_t nitems = 0;                // total query vertices
  uint32_t cap = 16;                  // contact slots per vertex
  uint32_t* ccount = nullptr;         // per query vertex contact count
  float4* cnd = nullptr;              // [nitems * cap] normal + distance
  uint2* cid = nullptr;               // [nitems * cap] (target mesh, facet)
  uint32_t* npairs = nullptr;         // pair-emission counter
  uint32_t* wstats = nullptr;         // per-wave running candidate-pair counts (one word per wave)
  uint32_t* flags = nullptr;          // bit1 contact overflow, bit2 traversal guard
  // pair-emission mode (parity / drop-in broad phase)
  uint32_t* pairs = nullptr;          // 4 words per pair (qobj, point, tobj, tri)
  uint32_t pair_cap = 0;
  uint32_t block = 128;               // threads per block: 64, 128 or 256
  uint32_t waves = 5;                 // occupancy target (waves per SIMD): 5, 6 or 8
  uint32_t tile = 1;                  // 1: one wave per 64-vertex tile (k_query_tile); 0: one thread per vertex
  // optional tree cuts (RefitLinks::cut): the tile walk starts from them
  const uint32_t* cut_cnt = nullptr;
  const int32_t* cut = nullptr;
  // candidate audit (off when null): every candidate the walk reports is also
  // recorded as {query item, target mesh, facet, tag} (audit.hpp: tags, and
  // the host filter that drops the records of tile walks redone per lane)
  uint4* audit = nullptr;
  uint32_t* audit_cnt = nullptr;  // records reserved (may exceed audit_cap)
  uint32_t audit_cap = 0;
  // One query segment (one dynamic mesh): its mesh, vertex offset and grid
  // side, and its first target's mesh and tree, given by the host so that
  // the tile kernel's first loads wait on no table lookups (pre_t < 0: not
  // given; the first target is then read from tgt_off / tgt).
  int32_t pre_t = -1;
  uint32_t seg0_mesh = 0, seg0_voff = 0, seg0_grid_n = 0;
  const ysim_bvh_node* pre_tree = nullptr;
};
// Audit record of one candidate found by a per-lane walk (a.audit set).
__device__ __forceinline__ void audit_lane(const QueryArgs& a, uint32_t j, uint32_t t, uint32_t fid) {
  const uint32_t s = atomicAdd(a.audit_cnt, 1u);
  if (s < a.audit_cap) a.audit[s] = make_uint4(j, t, fid, audit_tag(AUDIT_LANE, 0u));
}
bool query_block_ok(uint32_t b);
bool query_waves_ok(uint32_t w);
int launch_query(hipStream_t s, const QueryArgs& a, bool emit_pairs);
struct ClothArgs;
// The tile query fused with the cloth step: each tile's wave, after its
// query, runs cloth_vertex on its 64 vertices (ca.nitems == a.nitems, the
// same dynamic vertices) and writes the box of their new positions to
// ca.box_parts + 6 tile (when set; query_tiles(a.nitems) records).  The
// contact slots stay in a.ccount / a.cnd / a.cid as with launch_query.
int launch_query_cloth(hipStream_t s, const QueryArgs& a, const ClothArgs& ca);
uint32_t query_tiles(uint32_t nitems);
// Drop-in broad query of one (query mesh, target tree): BroadCollision records;
// segmented: pairs staged per workgroup in LDS, one output reservation each.
int launch_broad_pairs(hipStream_t s, const float* xq, uint32_t nq, uint32_t qobj, uint32_t qbeh,
                       uint32_t qshape, const ysim_bvh_node* tree, const int32_t* parent, uint32_t nT,
                       const uint32_t* tfacets, uint32_t tobj, uint32_t tbeh, uint32_t tshape,
                       float margin, ysim_broad_collision* out, uint32_t cap, uint32_t* count,
                       uint32_t* flags, bool segmented = false);
// Drop-in narrow phase over broad pairs (NarrowCollision records, atomic append).
int launch_narrow_pairs(hipStream_t s, const ysim_broad_collision* pairs, uint32_t np, const float* x,
                        const float* xprev, const uint32_t* states_off, const uint32_t* facets,
                        const uint32_t* facets_off, const uint32_t* adj_off, const uint32_t* adj,
                        float band, ysim_narrow_collision* out, uint32_t cap, uint32_t* count);
// Contact CSR: per-vertex buckets ordered by (target_obj, triangle).
int launch_contact_csr(hipStream_t s, const ysim_narrow_collision* c, uint32_t nc,
                       const uint32_t* states_off, uint32_t nv, uint32_t* offsets,
                       ysim_narrow_collision* sorted, uint32_t* scratch /* nv + 1024 words */);

// ---------------------------------------------------------------- cloth.hip
// Per-spring index of a FastGrid cloth of side N, families in the order
// struct-x (r,c)-(r,c+1), struct-y (r,c)-(r+1,c), shear-a (r,c)-(r+1,c+1),
// shear-b (r,c+1)-(r+1,c), bend-x (r,c)-(r,c+2), bend-y (r,c)-(r+2,c), each
// row-major over its first endpoint.
struct SpringIdx {
  uint32_t N, SY, SA, SB, BX, BY, total;
  YSIM_HD explicit SpringIdx(uint32_t n) : N(n) {
    SY = N * (N - 1);
    SA = SY + (N - 1) * N;
    SB = SA + (N - 1) * (N - 1);
    BX = SB + (N - 1) * (N - 1);
    BY = BX + N * (N - 2);
    total = BY + (N - 2) * N;
  }
};

// Visit the (up to) 12 grid springs of local vertex (row, col) in the gather
// order of compute_cloth_grid_forces_fast (src/metal/physics.metal:195-218):
// fn(local neighbour, row offset, column offset, rest length, family k,
// spring index).
template <class Fn>
__device__ __forceinline__ void for_grid_springs_rc(const ysim_fastgrid_params& P, uint32_t row, uint32_t col,
                                                    Fn&& fn) {
  const uint32_t N = P.particle_num_1d;
  const SpringIdx S(N);
  const uint32_t i = row * N + col;
  if (col > 0) fn(i - 1, 0, -1, P.stretch_rest_x, P.kstretch, row * (N - 1) + col - 1);
  if (col < N - 1) fn(i + 1, 0, 1, P.stretch_rest_x, P.kstretch, row * (N - 1) + col);
  if (row > 0) fn(i - N, -1, 0, P.stretch_rest_y, P.kstretch, S.SY + (row - 1) * N + col);
  if (row < N - 1) fn(i + N, 1, 0, P.stretch_rest_y, P.kstretch, S.SY + row * N + col);
  if (col > 0 && row > 0) fn(i - N - 1, -1, -1, P.shear_rest_a, P.kshear, S.SA + (row - 1) * (N - 1) + col - 1);
  if (col < N - 1 && row > 0) fn(i - N + 1, -1, 1, P.shear_rest_b, P.kshear, S.SB + (row - 1) * (N - 1) + col);
  if (col > 0 && row < N - 1) fn(i + N - 1, 1, -1, P.shear_rest_b, P.kshear, S.SB + row * (N - 1) + col - 1);
  if (col < N - 1 && row < N - 1) fn(i + N + 1, 1, 1, P.shear_rest_a, P.kshear, S.SA + row * (N - 1) + col);
  if (col > 1) fn(i - 2, 0, -2, P.bend_rest_x, P.kbend, S.BX + row * (N - 2) + col - 2);
  if (col < N - 2) fn(i + 2, 0, 2, P.bend_rest_x, P.kbend, S.BX + row * (N - 2) + col);
  if (row > 1) fn(i - 2 * N, -2, 0, P.bend_rest_y, P.kbend, S.BY + (row - 2) * N + col);
  if (row < N - 2) fn(i + 2 * N, 2, 0, P.bend_rest_y, P.kbend, S.BY + row * N + col);
}
// The same for local vertex i: fn(local neighbour, rest length, family k,
// spring index).
template <class Fn>
__device__ __forceinline__ void for_grid_springs(const ysim_fastgrid_params& P, uint32_t i, Fn&& fn) {
  const uint32_t N = P.particle_num_1d;
  for_grid_springs_rc(P, i / N, i % N,
                      [&](uint32_t q, int, int, float rest, float k, uint32_t sidx) { fn(q, rest, k, sidx); });
}

// Neighbour states for grid_force_rc: vertex vb + q of X and V.
struct NbrGlobal {
  const float* X;
  const float* V;
  uint32_t vb;
  __device__ __forceinline__ void operator()(uint32_t q, int, int, f3& x, f3& v) const {
    x = ld3(X, vb + q);
    v = ld3(V, vb + q);
  }
};

// Spring + damping + air force on the vertex at (row, col): ext + kair v + the
// 12 calc_spring terms in gather order, neighbour states from nbr.  KSPRING:
// KS holds one stiffness per spring and replaces the family constants (a
// template switch, so the family constants stay in registers).
template <bool KSPRING, class Nbr>
__device__ __forceinline__ f3 grid_force_rc(const ysim_fastgrid_params& P, uint32_t row, uint32_t col, const Nbr& nbr,
                                            f3 pos0, f3 vel0, f3 ext, float kair, float kd, const float* KS) {
  f3 force = ext + vel0 * kair;
  for_grid_springs_rc(P, row, col, [&](uint32_t q, int dr, int dc, float rest, float k, uint32_t sidx) {
    f3 xq, vq;
    nbr(q, dr, dc, xq, vq);
    force = force + calc_spring(pos0, vel0, xq, vq, rest, KSPRING ? KS[sidx] : k, kd);
  });
  return force;
}

// grid_force_rc on local vertex i (global offset vb), neighbours from X and V;
// KS (optional) holds one stiffness per spring.
__device__ __forceinline__ f3 grid_force(const ysim_fastgrid_params& P, uint32_t i, uint32_t vb, const float* X,
                                         const float* V, f3 pos0, f3 vel0, f3 ext, float kair, float kd,
                                         const float* KS) {
  const uint32_t N = P.particle_num_1d, row = i / N, col = i % N;
  const NbrGlobal nbr{X, V, vb};
  return KS ? grid_force_rc<true>(P, row, col, nbr, pos0, vel0, ext, kair, kd, KS)
            : grid_force_rc<false>(P, row, col, nbr, pos0, vel0, ext, kair, kd, KS);
}

// compute_tri_spring_forces (src/metal/physics.metal:319-378) on global vertex
// g: ext + kair v, then the stretch partners [off[g], mid[g]) with k_stretch,
// then the bend partners [mid[g], off[g+1]) with k_bend; each entry is
// (global partner, rest length bits).
__device__ __forceinline__ f3 tri_force(const ysim_fastgrid_params& P, uint32_t g, const uint32_t* off,
                                        const uint32_t* mid, const uint2* ent, const float* X, const float* V,
                                        f3 pos0, f3 vel0, f3 ext, float kair, float kd) {
  f3 force = ext + vel0 * kair;
  const uint32_t b = off[g], mdl = mid[g], e = off[g + 1];
  for (uint32_t k = b; k < mdl; ++k) {
    const uint2 q = ent[k];
    force = force + calc_spring(pos0, vel0, ld3(X, q.x), ld3(V, q.x), __uint_as_float(q.y), P.kstretch, kd);
  }
  for (uint32_t k = mdl; k < e; ++k) {
    const uint2 q = ent[k];
    force = force + calc_spring(pos0, vel0, ld3(X, q.x), ld3(V, q.x), __uint_as_float(q.y), P.kbend, kd);
  }
  return force;
}


struct ClothArgs {
  SegTable dsegs;                     // dynamic vertex segments (== query segments)
  MeshTables mt;
  const ysim_fastgrid_params* cloth = nullptr;  // per mesh
  const uint32_t* behavior = nullptr; // per mesh (null: every mesh is a FastGrid cloth)
  const uint32_t* tri_off = nullptr;  // TriangularCloth springs per global vertex (tri_force)
  const uint32_t* tri_mid = nullptr;
  const uint2* tri_ent = nullptr;
  const float* kspring = nullptr;     // optional per-spring k of the first cloth
  float subh = 0, kair = 0, kd = 0;
  const float* x = nullptr;
  const float* v = nullptr;
  const float* ext = nullptr;
  const float* mass = nullptr;
  const float* fixed = nullptr;
  const uint32_t* ccount = nullptr;
  const float4* cnd = nullptr;
  const uint2* cid = nullptr;
  uint32_t cap = 16;
  // alternative contact source: a CSR ordered by (target mesh, facet) per
  // vertex (the drop-in cloth_step entry point); used when csr_off != nullptr
  const uint32_t* csr_off = nullptr;
  const ysim_narrow_collision* csr = nullptr;
  float* x_out = nullptr;
  float* v_out = nullptr;
  float* f_out = nullptr;             // optional
  uint32_t nitems = 0;
  uint32_t* flags = nullptr;          // bit0 anomaly
  uint32_t block = 256;               // threads per block: 128, 256 or 512
  // optional: per block, the box of its x_out (6 ordered keys) at
  // box_parts + 6 block; cloth_box_parts(nitems, block) records
  uint32_t* box_parts = nullptr;
};
bool cloth_block_ok(uint32_t b);
uint32_t cloth_box_parts_words(uint32_t nitems);  // enough for every block size
uint32_t cloth_box_parts(uint32_t nitems, uint32_t block);

// Where cloth_vertex reads the states of a vertex and its spring neighbours:
// x and v in device memory (NoTile), or a GridTile staged in LDS by the
// caller: the 12 x 12 vertices (rows r0-2 .. r0+9, columns c0-2 .. c0+9) of
// an 8 x 8 tile of an N x N FastGrid cloth with its two-vertex halo, as
// float4 (xyz, unused w); entry (r - r0 + 2) * 12 + (c - c0 + 2) holds
// vertex (r, c), h is the caller's own entry and (row, col) its vertex, ext
// its external force.
// Entries outside the cloth are never read: the springs to them do not exist.
struct NoTile {
  static constexpr bool staged = false;
};
struct GridTile {
  static constexpr bool staged = true;
  static constexpr int W = 12;  // halo row length
  const float4* X;
  const float4* V;
  uint32_t row, col, h;
  f3 ext;  // the caller's external force, loaded with the halo
};
// Neighbour states for grid_force_rc from a GridTile, relative to the
// caller's entry (X, V point at it).
struct NbrTile {
  const float4* X;
  const float4* V;
  __device__ __forceinline__ void operator()(uint32_t, int dr, int dc, f3& x, f3& v) const {
    const int o = dr * GridTile::W + dc;
    const float4 a = X[o], b = V[o];
    x = {a.x, a.y, a.z};
    v = {b.x, b.y, b.z};
  }
};

// The state and total force of local vertex i of mesh m (global vertex vb +
// i, dynamic segment sg): the first half of cloth_vertex.  A GridTile (the
// vertex is (tile.row, tile.col) of a FastGrid cloth) supplies the states of
// the vertex and its neighbours; the arithmetic is the same either way.
struct ClothForce {
  f3 pos0, vel0, force;
};
template <class Tile = NoTile>
__device__ __forceinline__ ClothForce cloth_force_at(const ClothArgs& a, uint32_t sg, uint32_t m, uint32_t i,
                                                     uint32_t vb, const Tile& tile = Tile{}) {
  const uint32_t g = vb + i;
  const ysim_fastgrid_params P = a.cloth[m];
  const float* KS = (a.kspring && sg == 0) ? a.kspring : nullptr;
  ClothForce cf;
  if constexpr (Tile::staged) {
    const float4 xs = tile.X[tile.h], vs = tile.V[tile.h];
    cf.pos0 = {xs.x, xs.y, xs.z};
    cf.vel0 = {vs.x, vs.y, vs.z};
    const NbrTile nbr{tile.X + tile.h, tile.V + tile.h};
    cf.force = KS ? grid_force_rc<true>(P, tile.row, tile.col, nbr, cf.pos0, cf.vel0, tile.ext, a.kair, a.kd, KS)
                  : grid_force_rc<false>(P, tile.row, tile.col, nbr, cf.pos0, cf.vel0, tile.ext, a.kair, a.kd, KS);
  } else {
    cf.pos0 = ld3(a.x, g);
    cf.vel0 = ld3(a.v, g);
    cf.force = (a.behavior && a.behavior[m] == YSIM_BEHAVIOR_TRIANGULAR_CLOTH)
                   ? tri_force(P, g, a.tri_off, a.tri_mid, a.tri_ent, a.x, a.v, cf.pos0, cf.vel0, ld3(a.ext, g),
                               a.kair, a.kd)
                   : grid_force(P, i, vb, a.x, a.v, cf.pos0, cf.vel0, ld3(a.ext, g), a.kair, a.kd, KS);
  }
  return cf;
}

// The rest of vertex j's substep (global vertex g of mesh m) from its state
// and force: integrate, contacts, guard; stores v_out and x_out (and f_out)
// and returns the new position.  count >= 0: the vertex's contact count,
// known to the caller (else read from a.ccount).
__device__ __forceinline__ f3 cloth_apply_at(const ClothArgs& a, uint32_t j, uint32_t m, uint32_t g, int32_t count,
                                             const ClothForce& cf) {
  const f3 pos0 = cf.pos0, vel0 = cf.vel0, force = cf.force;
  if (a.f_out) st3(a.f_out, g, force);

  // integrate_cloth_grid
  const float mask = a.fixed[g];
  const float mass = a.mass[g];
  const float subh = a.subh;
  f3 pos = pos0, vel = vel0;
  vel.x = vel.x + ((subh * force.x) / mass) * mask;
  vel.y = vel.y + ((subh * force.y) / mass) * mask;
  vel.z = vel.z + ((subh * force.z) / mass) * mask;

  const float thickness = a.cloth[m].thickness;
  auto apply = [&](float4 nd) {
    f3 n = {nd.x, nd.y, nd.z};
    float nlen2 = dot(n, n);
    if (nlen2 < 1e-12f) return;
    float inv = 1.0f / sqrt_len(nlen2);
    n = n * inv;
    float vn = dot(vel, n);
    float dist = nd.w;
    if (dist < thickness) {
      if (vn < 0.0f) vel = vel - vn * n;
      pos = pos + (thickness - dist) * n;
    }
  };
  if (a.csr_off) {
    // bucket already ordered; skip repeats of a (target, facet) (physics.metal:257-271)
    const uint32_t b = a.csr_off[g], e = a.csr_off[g + 1];
    for (uint32_t c = b; c < e; ++c) {
      bool dup = false;
      for (uint32_t d = b; d < c; ++d)
        if (a.csr[d].target_obj == a.csr[c].target_obj && a.csr[d].triangle == a.csr[c].triangle) {
          dup = true;
          break;
        }
      if (dup) continue;
      const float* q = a.csr[c].normal_distance;
      apply(make_float4(q[0], q[1], q[2], q[3]));
    }
  } else {
    uint32_t cnt = count >= 0 ? (uint32_t)count : (a.ccount ? a.ccount[j] : 0u);
    if (cnt > a.cap) cnt = a.cap;
    const float4* ND = a.cnd + (size_t)j * a.cap;
    const uint2* ID = a.cid + (size_t)j * a.cap;
    // visit contacts in ascending (target mesh, facet) order; equal keys are
    // duplicates and are applied once (physics.metal:257-271)
    uint64_t last = 0;
    bool first = true;
    for (uint32_t it = 0; it < cnt; ++it) {
      uint64_t best = ~0ull;
      uint32_t bi = 0;
      for (uint32_t c = 0; c < cnt; ++c) {
        uint2 id = ID[c];
        uint64_t key = ((uint64_t)id.x << 32) | id.y;
        if ((first || key > last) && key < best) {
          best = key;
          bi = c;
        }
      }
      if (best == ~0ull) break;
      first = false;
      last = best;
      apply(ND[bi]);
    }
  }
  pos = pos + (subh * vel) * mask;
  if (sanitize(pos, vel, pos0)) atomicOr(a.flags, 1u);
  st3(a.v_out, g, vel);
  st3(a.x_out, g, pos);
  return pos;
}

// One vertex j < a.nitems: forces, integrate, contacts, guard; stores v_out
// and x_out and returns the new position (count as for cloth_apply_at).
__device__ __forceinline__ f3 cloth_vertex(const ClothArgs& a, uint32_t j, int32_t count = -1) {
  const uint32_t sg = seg_of(a.dsegs, j);
  const uint32_t m = a.dsegs.mesh[sg], vb = a.mt.voff[m], i = j - a.dsegs.off[sg];
  return cloth_apply_at(a, j, m, vb + i, count, cloth_force_at(a, sg, m, i, vb));
}

int launch_cloth(hipStream_t s, const ClothArgs& a);

// ---------------------------------------------------------------- analytic.hip
// One analytic collider on the device: the caller's record plus constants the
// host computes once per install (the swept box before padding, the broad
// rotation band, the narrow rotation angle (0 below 1e-9 and for planes) and
// radius).
struct AShape {
  ysim_analytic_shape s;
  float lo[3], rot_band;
  float hi[3], pad0;
  double theta, rmax;
};
static_assert(sizeof(AShape) == 144, "AShape layout");
struct AnalyticArgs {
  SegTable dsegs;                  // dynamic vertex segments (one per cloth)
  MeshTables mt;
  const AShape* shapes = nullptr;
  const uint32_t* pair_off = nullptr;    // per cloth segment: its (cloth, shape) pairs, nseg + 1
  const uint32_t* pair_shape = nullptr;  // shape index of each pair
  const float* x = nullptr;
  const float* xprev = nullptr;
  const ysim_bvh_node* tree = nullptr;
  const float* mesh_band = nullptr;      // radius + narrow thickness per mesh
  const float* mesh_thick = nullptr;     // narrow thickness per mesh
  float radius = 0.f;
  const uint32_t* mark_prev = nullptr;   // pairing at the previous substep, per pair
  uint32_t* mark_cur = nullptr;          // pairing at this substep
  uint32_t* ccount = nullptr;
  float4* cnd = nullptr;
  uint2* cid = nullptr;
  uint32_t cap = 16;
  uint32_t* flags = nullptr;
  uint32_t nitems = 0;
};
// After the query: per dynamic vertex, the contacts with the analytic
// colliders its cloth is paired with, appended to the vertex's slots.
int launch_analytic(hipStream_t s, const AnalyticArgs& a);

// ---------------------------------------------------------------- adjoint.hip
// Adjoint of springs + integrate + contact projection for one FastGrid cloth
// whose contacts are against static meshes (contact slots indexed by the
// cloth-local vertex).  Adjoint arrays (xb, vb, fb, kdv) are cloth-local.
struct AdjArgs {
  ysim_fastgrid_params P{};
  uint32_t V = 0, voff = 0;
  const float* x = nullptr;     // scene arrays (global vertex index)
  const float* v = nullptr;
  const float* ext = nullptr;
  const float* mass = nullptr;
  const float* fixed = nullptr;
  const float* ks = nullptr;    // per-spring stiffness
  float subh = 0, kair = 0, kd = 0;
  const uint32_t* ccount = nullptr;
  const float4* cnd = nullptr;
  const uint2* cid = nullptr;
  uint32_t cap = 16;
  float* xb = nullptr;
  float* vb = nullptr;
  float* fb = nullptr;
  float* gk = nullptr;          // per-spring d/dk accumulator
  float* kdv = nullptr;         // per-vertex d/dkd accumulator
};
int launch_adj_exp(hipStream_t s, const float* theta, float* k, uint32_t n);
int launch_adj_step(hipStream_t s, const AdjArgs& a);
int launch_adj_finish(hipStream_t s, const float* gk, const float* k, float* g_log_k, uint32_t n,
                      const float* kdv, uint32_t V, float* g_kd);

}  // namespace ysim
