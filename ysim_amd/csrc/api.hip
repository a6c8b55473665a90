// api.hip — the C ABI (include/ysim_hip.h): runtime, stage entry points and the
// resident scene with its substep loop.
//
// The scene mirrors Simulator::initialize / update (include/sim/simulator.hpp:
// 2942-3115, 3512-3815) for FastGrid cloth + static (Float) obstacle meshes:
//   [rebuild: box -> Morton -> sort -> Karras] -> leaf/bottom-up refit
//   -> traversal fused with the narrow phase -> springs + integrate + contacts.
// Positions and velocities are double-buffered: substep s reads buffer s%2 and
// writes buffer (s+1)%2, whose old content (x before the previous integrate)
// is exactly the xPrev the swept narrow phase needs.  A hipGraph captures a
// whole period of substeps (even length, so buffer roles repeat).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "hash.hpp"
#include "kernels.hpp"

namespace ysim {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int Ctx::ensure_scratch(size_t bytes) {
  if (scratch_bytes >= bytes) return YSIM_OK;
  if (scratch) YSIM_HIP_TRY(hipFree(scratch));
  scratch = nullptr;
  scratch_bytes = 0;
  YSIM_HIP_TRY(hipMalloc(&scratch, bytes));
  scratch_bytes = bytes;
  return YSIM_OK;
}

}  // namespace ysim

using namespace ysim;

#define YSIM_TRY(expr)              \
  do {                              \
    int rc_ = (expr);               \
    if (rc_ != YSIM_OK) return rc_; \
  } while (0)

#define YSIM_ARG(cond, msg) \
  do {                      \
    if (!(cond)) return fail(YSIM_ERR_INVALID, msg); \
  } while (0)

struct ysim_ctx : ysim::Ctx {};

namespace {

uint32_t bits_for(uint32_t n) {  // ceil(log2(n)) for n >= 1
  uint32_t b = 0;
  while ((1ull << b) < n) ++b;
  return b;
}

// Springs of a TriangularCloth, restating MeshAdjacencyInitializer::initialize
// (include/sim/mesh_state.hpp:236-454).  Every facet (a, b, c) gives the edge
// records {a,b} opposite c, {b,c} opposite a, {c,a} opposite b; the records
// sorted by (min, max) index make the edge list.  A vertex's stretch partners
// are the other ends of its edges in edge order; an edge of two facets makes
// its two opposite vertices bend partners of each other, in edge order.  Rest
// lengths are tinym's norm of x[hi] - x[lo] (sqrt((x*x + y*y) + z*z)), and the
// bend length is symmetric in its ends.  Fails on a facet with a repeated
// vertex and on an edge of three or more facets.
struct TriSprings {
  std::vector<uint32_t> first, mid;  // per local vertex: stretch [first, mid), bend [mid, first[+1])
  std::vector<uint32_t> other;       // local partner
  std::vector<float> rest;
};

static uint32_t float_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

static float tri_length(const float* x, uint32_t p, uint32_t q) {
  const float dx = x[3 * q] - x[3 * p], dy = x[3 * q + 1] - x[3 * p + 1], dz = x[3 * q + 2] - x[3 * p + 2];
  return std::sqrt((dx * dx + dy * dy) + dz * dz);
}

static bool tri_springs(const uint32_t* F, uint32_t nf, const float* x, uint32_t nv, TriSprings& ts) {
  std::vector<std::pair<uint64_t, uint32_t>> rec;  // ((lo << 32) | hi, opposite vertex)
  rec.reserve(3 * (size_t)nf);
  for (uint32_t f = 0; f < nf; ++f) {
    const uint32_t c[3] = {F[3 * f], F[3 * f + 1], F[3 * f + 2]};
    if (c[0] == c[1] || c[1] == c[2] || c[2] == c[0]) return false;
    for (int k = 0; k < 3; ++k) {
      const uint32_t p = c[k], q = c[(k + 1) % 3];
      rec.push_back({((uint64_t)std::min(p, q) << 32) | std::max(p, q), c[(k + 2) % 3]});
    }
  }
  std::stable_sort(rec.begin(), rec.end(),
                   [](const std::pair<uint64_t, uint32_t>& u, const std::pair<uint64_t, uint32_t>& w) {
                     return u.first < w.first;
                   });
  struct Edge { uint32_t lo, hi, o0, o1; bool inner; };
  std::vector<Edge> edges;
  for (size_t k = 0; k < rec.size();) {
    size_t e = k + 1;
    while (e < rec.size() && rec[e].first == rec[k].first) ++e;
    if (e - k > 2) return false;
    edges.push_back({(uint32_t)(rec[k].first >> 32), (uint32_t)rec[k].first, rec[k].second,
                     e - k == 2 ? rec[k + 1].second : 0u, e - k == 2});
    k = e;
  }
  std::vector<uint32_t> ns(nv, 0), nb(nv, 0);
  for (const Edge& e : edges) {
    ns[e.lo]++;
    ns[e.hi]++;
    if (e.inner) {
      nb[e.o0]++;
      nb[e.o1]++;
    }
  }
  ts.first.assign(nv + 1, 0);
  ts.mid.assign(nv, 0);
  for (uint32_t i = 0; i < nv; ++i) {
    ts.mid[i] = ts.first[i] + ns[i];
    ts.first[i + 1] = ts.mid[i] + nb[i];
  }
  ts.other.assign(ts.first[nv], 0);
  ts.rest.assign(ts.first[nv], 0.f);
  std::vector<uint32_t> ps(ts.first.begin(), ts.first.end() - 1), pb(ts.mid);
  for (const Edge& e : edges) {
    const float r = tri_length(x, e.lo, e.hi);
    ts.other[ps[e.lo]] = e.hi;
    ts.rest[ps[e.lo]++] = r;
    ts.other[ps[e.hi]] = e.lo;
    ts.rest[ps[e.hi]++] = r;
    if (e.inner) {
      const float rb = tri_length(x, e.o0, e.o1);
      ts.other[pb[e.o0]] = e.o1;
      ts.rest[pb[e.o0]++] = rb;
      ts.other[pb[e.o1]] = e.o0;
      ts.rest[pb[e.o1]++] = rb;
    }
  }
  return true;
}

struct BuildGroup {
  std::vector<uint32_t> meshes;
  SegTable tsegs, vsegs;
  uint32_t T = 0, V = 0;
  uint32_t code_bits = 30;
  bool key64 = false;
  SortPlan plan;
  uint32_t* perm = nullptr;  // dynamic groups: sorted facet order of the last build (incremental sort)
};

constexpr int NEV = YSIM_STAGE_COUNT + 1;

}  // namespace

struct ysim_scene {
  ysim_ctx* ctx = nullptr;
  hipStream_t st = nullptr;
  ysim_sim_params P{};
  uint32_t nm = 0, Vtot = 0, Ttot = 0, Ntot = 0;
  std::vector<ysim_mesh_desc> meshes;
  std::vector<uint32_t> voff, foff, noff;
  std::vector<uint32_t> h_tgt_off, h_tgt, h_grid_n;  // host copies of d_tgt_off, d_tgt, d_grid_n
  std::vector<uint32_t> dyn, stat;
  std::vector<uint32_t> dyn_voff;        // host copy of dyn_vsegs offsets
  uint32_t Vdyn = 0, Tdyn = 0;
  float* x[2] = {nullptr, nullptr};
  float* v[2] = {nullptr, nullptr};
  float *mass = nullptr, *ext = nullptr, *fixed = nullptr, *kspring = nullptr;
  uint32_t* facets = nullptr;
  MeshTables mt;
  ysim_fastgrid_params* d_cloth = nullptr;
  uint32_t *d_nfacets = nullptr, *d_behavior = nullptr, *d_shape = nullptr, *d_grid_n = nullptr;
  uint32_t *tri_off = nullptr, *tri_mid = nullptr;  // TriangularCloth springs (tri_force)
  uint2* tri_ent = nullptr;
  float *d_margin = nullptr, *d_band = nullptr;
  ysim_bvh_node* tree = nullptr;
  int32_t* parent = nullptr;
  uint32_t* visit = nullptr;
  RefitLinks links;             // per node: parent's leaf count and side, leaf facet ids (LDS refit)
  std::vector<BuildGroup> dyn_groups, stat_groups;
  SegTable dyn_tsegs, dyn_vsegs;
  uint32_t *d_tgt_off = nullptr, *d_tgt = nullptr;
  void* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t *hist = nullptr, *counters = nullptr, *box_acc = nullptr;
  uint64_t* status = nullptr;   // sort look-back words (epoch-tagged, never cleared)
  uint32_t* epoch = nullptr;    // [0] look-back epoch of the current sort
  uint32_t* need_full = nullptr;  // [0] incremental sort: the window passes left the keys unsorted
  uint32_t* inc_bcnt = nullptr;   // incremental sort: window-boundary arrival counters
  uint32_t* box_parts = nullptr;  // one-mesh box: per-block boxes + arrival counter
  // graph replays leave out the incremental sort's fallback kernels: a build
  // whose order check fails keeps the previous (valid) topology, refit to the
  // new positions, and raises h_stale (pinned host word written by the GPU);
  // the host then runs the next substep eagerly, where the fallback sorts
  bool capturing = false;
  bool full_next = true;          // the next substep runs eagerly (state jumped)
  // One dynamic mesh: the cloth kernel also writes per block the box of its
  // new positions (cloth_parts); while box_fresh (no host change of positions
  // since) the next incremental build reduces them in its Morton kernel and
  // launches no box kernel.
  bool fuse_box = false, box_fresh = false;
  uint32_t* cloth_parts = nullptr;
  uint32_t parts_count = 0;  // box records the last cloth step wrote (per block or per tile)
  uint32_t* h_stale = nullptr;    // pinned host word
  uint32_t* d_stale = nullptr;    // its device address
  uint64_t deferred_builds = 0;
  uint64_t* inc_edges = nullptr;  // incremental sort: each window's first and last (key, facet)
  uint32_t cap = 16;
  uint32_t* ccount = nullptr;
  float4* cnd = nullptr;
  uint2* cid = nullptr;
  uint32_t* npairs = nullptr;  // [1] pair-emission counter
  uint32_t* wstats = nullptr;  // per query wave: running candidate-pair count
  uint32_t wstats_words = 0;
  std::vector<uint32_t> tune_wstats;
  uint32_t* flags = nullptr;   // [0] sticky flags
  uint32_t* pairs = nullptr;
  uint32_t pair_cap = 0;
  // candidate audit (scene_set_audit): records of the last audited query
  uint4* audit = nullptr;
  uint32_t* audit_cnt = nullptr;
  uint32_t audit_cap = 0;
  std::vector<void*> allocs;
  uint32_t parity = 0;
  uint64_t step = 0;
  bool formal_refit = false;
  uint32_t last_sorted = 0;
  hipGraphExec_t graph = nullptr;
  uint32_t graph_period = 0;
  hipEvent_t ev[NEV] = {};
  uint64_t pairs_hi = 0;        // reserved for a widened pair counter
  float *snap_x = nullptr, *snap_v = nullptr, *snap_xp = nullptr;
  uint64_t snap_step = 0;
  bool snap_valid = false;
  uint64_t counter_step0 = 0;
  uint64_t steps_run = 0;       // substeps executed (monotonic, unaffected by restore)
  int stage_phase = 0;          // 0: next is refit, 1: query, 2: integrate (staged driving)
  bool hash_mode = false;       // broad_mode 1: spatial hash instead of LBVH traversal
  HashArgs hargs;
  bool timing = false;
  double stage_ms[YSIM_STAGE_COUNT] = {0};
  // differentiation: checkpoints of the cloth's (x, v, xPrev) every K substeps
  // of the last adjoint forward, the per-segment tape and adjoint buffers
  struct Adj {
    bool valid = false;
    uint32_t mesh = 0, V = 0, voff = 0, S = 0;
    uint32_t steps = 0, K = 0, nseg = 0;
    uint64_t step0 = 0;
    uint32_t par_end = 0;
    float kd = 0.f;
    float *ck = nullptr, *tape = nullptr, *xb = nullptr, *vb = nullptr, *fb = nullptr, *gk = nullptr,
          *kdv = nullptr;
    size_t ck_floats = 0, tape_floats = 0;
  } adj;
  float graph_kd = 0.f;
  const float* graph_ks = nullptr;
  ysim_tuning tune{8, 8, 128, 256, 0.f, 0.f, 256, 0, 0, 5, 1, 0};
  float* tune_buf = nullptr;   // autotune state save: x, v, xPrev of the whole scene
  // analytic colliders: the installed shapes, their device copies, the
  // (cloth, shape) pairs per dynamic segment and each pair's pairing flag of
  // the last two substeps (amark[par * napairs + pair] is written by the
  // substep of parity par)
  std::vector<ysim_analytic_shape> ashapes;
  std::vector<int> shape_of;    // per mesh: its shape index, -1 none
  AShape* d_ashapes = nullptr;
  uint32_t ashape_cap = 0;
  uint32_t *d_apair_off = nullptr, *d_apair_shape = nullptr;
  uint32_t napairs = 0;
  uint32_t* amark = nullptr;
  std::vector<uint32_t> snap_amark;
  std::vector<uint32_t> tune_amark;
  float* d_thick = nullptr;
  bool analytic_on() const { return !ashapes.empty() && !hash_mode; }
  // agglomerative builds: range scratch (two ints per node) and whether the
  // dynamic trees were last built by it (their refit links are then absent)
  int2* agg_range = nullptr;
  bool tree_agglo = false;
  // Two-kernel refit: spanning-node records of the hierarchy kernel (in
  // links), partial boxes, and the refit block size the dynamic group's
  // records were written for (0: none; the refit then climbs device-wide).
  RefitSpan span;
  uint32_t span_rb_dyn = 0;
  // the same view after a replay of the captured graph
  bool graph_tree_agglo = false;
  uint32_t graph_last_sorted = 0, graph_span_rb = 0;
  // The links a group's hierarchy kernel writes: spanning nodes are recorded
  // when the refit that follows shares the group's leaf positions (a static
  // group's own refit at creation, or the only dynamic group).
  bool records_span(const BuildGroup& g) const {
    const bool dyn_only = dyn_groups.size() == 1 && &g == &dyn_groups[0];
    const bool stat = &g >= stat_groups.data() && &g < stat_groups.data() + stat_groups.size();
    return !formal_refit && (dyn_only || stat);
  }
  RefitLinks links_for(const BuildGroup& g) {
    RefitLinks l = links;
    if (records_span(g)) {
      l.span_rb = tune.refit_block;
      if (dyn_groups.size() == 1 && &g == &dyn_groups[0]) span_rb_dyn = tune.refit_block;
    }
    return l;
  }
  const RefitSpan* span_for_dyn() const {
    return (dyn_groups.size() == 1 && span_rb_dyn == tune.refit_block && !formal_refit) ? &span : nullptr;
  }

  int grow(float** p, size_t& have, size_t need) {
    if (need <= have && *p) return YSIM_OK;
    if (*p) {
      YSIM_HIP_TRY(hipStreamSynchronize(st));
      allocs.erase(std::find(allocs.begin(), allocs.end(), (void*)*p));
      YSIM_HIP_TRY(hipFree(*p));
      *p = nullptr;
    }
    YSIM_TRY(alloc(p, need));
    have = need;
    return YSIM_OK;
  }
  // copy the cloth range of (x[par], v[par], x[par ^ 1]) to/from 3 consecutive
  // 3V-float blocks at dst
  int save_state(float* dst, uint32_t par) {
    const size_t n = 3 * (size_t)adj.V, o = 3 * (size_t)adj.voff;
    YSIM_HIP_TRY(hipMemcpyAsync(dst, x[par] + o, 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(dst + n, v[par] + o, 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(dst + 2 * n, x[par ^ 1] + o, 4 * n, hipMemcpyDeviceToDevice, st));
    return YSIM_OK;
  }
  int load_state(const float* src, uint32_t par) {
    full_next = true;
    box_fresh = false;
    const size_t n = 3 * (size_t)adj.V, o = 3 * (size_t)adj.voff;
    YSIM_HIP_TRY(hipMemcpyAsync(x[par] + o, src, 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(v[par] + o, src + n, 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(x[par ^ 1] + o, src + 2 * n, 4 * n, hipMemcpyDeviceToDevice, st));
    return YSIM_OK;
  }
  void drop_graph() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }

  int apply_tuning(const ysim_tuning& t) {
    if (!sort_config_ok(t.sort_digit_bits, t.sort_items))
      return fail(YSIM_ERR_INVALID, "tuning: sort_digit_bits must be 8 or 11 and sort_items 8 or 16");
    if (!query_block_ok(t.query_block)) return fail(YSIM_ERR_INVALID, "tuning: query_block must be 64, 128 or 256");
    if (!cloth_block_ok(t.cloth_block)) return fail(YSIM_ERR_INVALID, "tuning: cloth_block must be 128, 256 or 512");
    if (!refit_block_ok(t.refit_block)) return fail(YSIM_ERR_INVALID, "tuning: refit_block must be 256, 512 or 1024");
    if (t.sort_hybrid > 1) return fail(YSIM_ERR_INVALID, "tuning: sort_hybrid must be 0 or 1");
    if (t.sort_incremental > 1) return fail(YSIM_ERR_INVALID, "tuning: sort_incremental must be 0 or 1");
    if (!query_waves_ok(t.query_waves)) return fail(YSIM_ERR_INVALID, "tuning: query_waves must be 5, 6 or 8");
    if (t.query_tile > 2) return fail(YSIM_ERR_INVALID, "tuning: query_tile must be 0, 1 or 2");
    if (t.build_mode > 1) return fail(YSIM_ERR_INVALID, "tuning: build_mode must be 0 or 1");
    if (t.build_mode && !agg_range) YSIM_TRY(alloc(&agg_range, Ntot));
    const bool replan = t.sort_digit_bits != tune.sort_digit_bits || t.sort_items != tune.sort_items ||
                        t.sort_hybrid != tune.sort_hybrid || t.sort_incremental != tune.sort_incremental;
    tune.sort_digit_bits = t.sort_digit_bits;
    tune.sort_items = t.sort_items;
    tune.query_block = t.query_block;
    tune.cloth_block = t.cloth_block;
    tune.refit_block = t.refit_block;
    tune.sort_hybrid = t.sort_hybrid;
    tune.sort_incremental = t.sort_incremental;
    tune.query_waves = t.query_waves;
    tune.query_tile = t.query_tile;
    tune.build_mode = t.build_mode;
    if (replan) {
      for (auto& g : dyn_groups)
        g.plan = (t.sort_hybrid || t.sort_incremental)
                     ? sort_plan_hybrid(g.T, g.plan.key_bits, t.sort_items)
                     : sort_plan(g.T, g.plan.key_bits, t.sort_digit_bits, t.sort_items);
      YSIM_TRY(zero(hist, SORT_HIST_WORDS));
    }
    drop_graph();
    full_next = true;
    box_fresh = false;
    return YSIM_OK;
  }

  // whole-scene state save/restore for autotune trials
  int tune_save(uint32_t* fl_np) {
    const size_t n = 3 * (size_t)Vtot;
    if (!tune_buf) YSIM_TRY(alloc(&tune_buf, 3 * n + 4));
    YSIM_HIP_TRY(hipMemcpyAsync(tune_buf, x[parity], 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(tune_buf + n, v[parity], 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(tune_buf + 2 * n, x[parity ^ 1], 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(fl_np, flags, 4, hipMemcpyDeviceToHost, st));
    YSIM_HIP_TRY(hipMemcpyAsync(fl_np + 1, npairs, 4, hipMemcpyDeviceToHost, st));
    tune_wstats.resize(wstats_words);
    if (wstats_words)
      YSIM_HIP_TRY(hipMemcpyAsync(tune_wstats.data(), wstats, 4 * (size_t)wstats_words, hipMemcpyDeviceToHost, st));
    YSIM_HIP_TRY(hipStreamSynchronize(st));
    return save_amark(tune_amark);
  }
  int tune_restore(uint32_t par, uint64_t stp, const uint32_t* fl_np) {
    const size_t n = 3 * (size_t)Vtot;
    YSIM_HIP_TRY(hipMemcpyAsync(x[par], tune_buf, 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(v[par], tune_buf + n, 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(x[par ^ 1], tune_buf + 2 * n, 4 * n, hipMemcpyDeviceToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(flags, fl_np, 4, hipMemcpyHostToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(npairs, fl_np + 1, 4, hipMemcpyHostToDevice, st));
    if (wstats_words && tune_wstats.size() == wstats_words)
      YSIM_HIP_TRY(hipMemcpyAsync(wstats, tune_wstats.data(), 4 * (size_t)wstats_words, hipMemcpyHostToDevice, st));
    YSIM_HIP_TRY(hipStreamSynchronize(st));
    YSIM_TRY(load_amark(tune_amark, par));
    parity = par;
    step = stp;
    full_next = true;
    box_fresh = false;
    return YSIM_OK;
  }

  int autotune(uint32_t trial, ysim_tuning* out) {
    if (stage_phase != 0) return fail(YSIM_ERR_STATE, "autotune: a staged substep is in progress");
    if (trial == 0) trial = 10;
    uint32_t saved[2] = {0, 0};
    YSIM_TRY(tune_save(saved));
    const uint32_t par0 = parity;
    const uint64_t step0 = step;
    const uint64_t run0 = steps_run;
    const bool timing0 = timing;
    timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    YSIM_HIP_TRY(hipEventCreate(&e0));
    YSIM_HIP_TRY(hipEventCreate(&e1));
    auto measure = [&](const ysim_tuning& t, float& ms) -> int {
      YSIM_TRY(apply_tuning(t));
      YSIM_TRY(tune_restore(par0, step0, saved));
      YSIM_TRY(steps(2, true));  // capture + warm
      YSIM_TRY(tune_restore(par0, step0, saved));
      YSIM_HIP_TRY(hipEventRecord(e0, st));
      YSIM_TRY(steps(trial, true));
      YSIM_HIP_TRY(hipEventRecord(e1, st));
      YSIM_HIP_TRY(hipEventSynchronize(e1));
      YSIM_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      ms /= (float)trial;
      return YSIM_OK;
    };
    // a setting replaces the current best only when it is at least 1% faster,
    // so timing noise does not churn the choice
    constexpr float kGain = 0.99f;
    ysim_tuning best = tune;
    float best_ms = 0.f;
    int rc = measure(best, best_ms);
    const float default_ms = best_ms;
    const bool rebuilds = P.rebuild_every > 0 && !hash_mode;
    // coordinate descent: sort (only when the substep rebuilds), query block, cloth block
    if (rc == YSIM_OK && rebuilds) {
      struct SortCand { uint32_t db, items, hybrid, inc; };
      for (SortCand c : {SortCand{8, 8, 0, 0}, SortCand{8, 16, 0, 0}, SortCand{11, 8, 0, 0}, SortCand{11, 16, 0, 0},
                         SortCand{8, 8, 1, 0}, SortCand{8, 16, 1, 0}, SortCand{8, 16, 0, 1}}) {
        ysim_tuning t = best;
        t.sort_digit_bits = c.db;
        t.sort_items = c.items;
        t.sort_hybrid = c.hybrid;
        t.sort_incremental = c.inc;
        if (t.sort_digit_bits == best.sort_digit_bits && t.sort_items == best.sort_items &&
            t.sort_hybrid == best.sort_hybrid && t.sort_incremental == best.sort_incremental)
          continue;
        float ms = 0.f;
        if ((rc = measure(t, ms)) != YSIM_OK) break;
        if (ms < best_ms * kGain) { best_ms = ms; best = t; }
      }
    }
    if (rc == YSIM_OK && rebuilds) {
      ysim_tuning t = best;
      t.build_mode = best.build_mode ? 0u : 1u;
      float ms = 0.f;
      if ((rc = measure(t, ms)) == YSIM_OK && ms < best_ms * kGain) {
        best_ms = ms;
        best = t;
      }
    }
    if (rc == YSIM_OK && !hash_mode)
      for (uint32_t qt : {1u, 0u}) {
        if (qt == best.query_tile) continue;
        ysim_tuning t = best;
        t.query_tile = qt;
        float ms = 0.f;
        if ((rc = measure(t, ms)) != YSIM_OK) break;
        if (ms < best_ms * kGain) { best_ms = ms; best = t; }
      }
    // block size and occupancy apply to the per-vertex kernel only
    if (rc == YSIM_OK && !hash_mode && !best.query_tile)
      for (uint32_t qb : {64u, 128u, 256u}) {
        if (qb == best.query_block) continue;
        ysim_tuning t = best;
        t.query_block = qb;
        float ms = 0.f;
        if ((rc = measure(t, ms)) != YSIM_OK) break;
        if (ms < best_ms * kGain) { best_ms = ms; best = t; }
      }
    if (rc == YSIM_OK && !hash_mode && !best.query_tile)
      for (uint32_t qw : {5u, 6u, 8u}) {
        if (qw == best.query_waves) continue;
        ysim_tuning t = best;
        t.query_waves = qw;
        float ms = 0.f;
        if ((rc = measure(t, ms)) != YSIM_OK) break;
        if (ms < best_ms * kGain) { best_ms = ms; best = t; }
      }
    if (rc == YSIM_OK && !hash_mode)
      for (uint32_t rb : {256u, 512u, 1024u}) {
        if (rb == best.refit_block) continue;
        ysim_tuning t = best;
        t.refit_block = rb;
        float ms = 0.f;
        if ((rc = measure(t, ms)) != YSIM_OK) break;
        if (ms < best_ms * kGain) { best_ms = ms; best = t; }
      }
    if (rc == YSIM_OK)
      for (uint32_t cb : {128u, 256u, 512u}) {
        if (cb == best.cloth_block) continue;
        ysim_tuning t = best;
        t.cloth_block = cb;
        float ms = 0.f;
        if ((rc = measure(t, ms)) != YSIM_OK) break;
        if (ms < best_ms * kGain) { best_ms = ms; best = t; }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    YSIM_TRY(rc);
    YSIM_TRY(apply_tuning(best));
    YSIM_TRY(tune_restore(par0, step0, saved));
    steps_run = run0;
    timing = timing0;
    tune.ms_per_substep = best_ms;
    tune.ms_per_substep_default = default_ms;
    if (out) *out = tune;
    return YSIM_OK;
  }

  template <typename T>
  int alloc(T** p, size_t n) {
    *p = nullptr;
    YSIM_HIP_TRY(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
    allocs.push_back((void*)*p);
    return YSIM_OK;
  }
  // free one allocation of this scene (stream drained first)
  int release(void* p) {
    if (!p) return YSIM_OK;
    YSIM_HIP_TRY(hipStreamSynchronize(st));
    allocs.erase(std::remove(allocs.begin(), allocs.end(), p), allocs.end());
    YSIM_HIP_TRY(hipFree(p));
    return YSIM_OK;
  }
  template <typename T>
  int put(T** p, const T* h, size_t n) {
    YSIM_TRY(alloc(p, n));
    if (n) YSIM_HIP_TRY(hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
    return YSIM_OK;
  }
  template <typename T>
  int put(T** p, const std::vector<T>& h) {
    return put(p, h.data(), h.size());
  }
  int fill(void* p, size_t words, uint32_t v) {  // a kernel: graph captures record no memset nodes
    if (!words) return YSIM_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>(4096, ceil_div(words, 256))), dim3(256), 0, st,
                       (uint32_t*)p, (uint32_t)words, v);
    YSIM_CHECK_LAUNCH();
    return YSIM_OK;
  }
  int zero(void* p, size_t words) {
    if (!words) return YSIM_OK;
    hipLaunchKernelGGL(k_zero_u32, dim3(std::min<uint32_t>(4096, ceil_div(words, 256))), dim3(256), 0, st,
                       (uint32_t*)p, (uint32_t)words);
    YSIM_CHECK_LAUNCH();
    return YSIM_OK;
  }
  bool dynamic(uint32_t m) const { return meshes[m].behavior != YSIM_BEHAVIOR_FLOAT; }

  // Triangle targets per query (dynamic) mesh: the collidable meshes of its
  // group, itself only with self collision, analytic colliders' meshes never
  // while the analytic path is on (BroadPhase::analyticSkip, bvh.hpp:2520-2523).
  int build_targets() {
    std::vector<uint32_t> toff(1, 0), tg;
    for (uint32_t q : dyn) {
      const ysim_mesh_desc& Q = meshes[q];
      if (Q.collidable) {
        for (uint32_t t = 0; t < nm; ++t) {
          const ysim_mesh_desc& T = meshes[t];
          if (!T.collidable || T.group != Q.group || T.num_facets == 0) continue;
          if (t == q && !P.self_collision) continue;
          if (analytic_on() && shape_of[t] >= 0) continue;
          tg.push_back(t);
        }
      }
      toff.push_back((uint32_t)tg.size());
    }
    YSIM_TRY(release(d_tgt_off));
    YSIM_TRY(release(d_tgt));
    d_tgt_off = d_tgt = nullptr;
    YSIM_TRY(put(&d_tgt_off, toff));
    YSIM_TRY(put(&d_tgt, tg));
    h_tgt_off = toff;
    h_tgt = tg;
    return YSIM_OK;
  }

  int set_analytic(const ysim_analytic_shape* h, uint32_t n);

  // Per-query-mesh broad margin, narrow band and contact thickness from the
  // scene parameters (margin <= 0: each cloth's own thickness).
  void collision_arrays(std::vector<float>& margin, std::vector<float>& band, std::vector<float>& thick) const {
    margin.resize(nm);
    band.resize(nm);
    thick.resize(nm);
    for (uint32_t m = 0; m < nm; ++m) {
      margin[m] = P.margin > 0.f ? P.margin : meshes[m].cloth.thickness;
      band[m] = P.radius + (P.margin > 0.f ? P.narrow_thickness : meshes[m].cloth.thickness);
      thick[m] = P.margin > 0.f ? P.narrow_thickness : meshes[m].cloth.thickness;
    }
  }

  // detectCollisions(margin, self) / narrowAndSortByVertices(radius, thickness)
  // arguments of one substep: new values are written into the device arrays
  // the query reads (a captured graph stays valid); a change of self
  // collision rebuilds the target lists and drops the graph.
  int set_collision(float margin, float radius, float thick, uint32_t self) {
    if (stage_phase == 2) return fail(YSIM_ERR_STATE, "scene_set_collision: the substep's query already ran");
    YSIM_ARG(std::isfinite(margin) && std::isfinite(radius) && std::isfinite(thick) && radius >= 0.f && thick >= 0.f,
             "scene_set_collision: margin, radius and thickness must be finite, radius and thickness >= 0");
    YSIM_ARG(self <= 1, "scene_set_collision: self_collision must be 0 or 1");
    if (margin == P.margin && radius == P.radius && thick == P.narrow_thickness && self == P.self_collision)
      return YSIM_OK;
    if (hash_mode) {
      float mx = 0.f;
      for (uint32_t m : dyn) mx = std::max(mx, margin > 0.f ? margin : meshes[m].cloth.thickness);
      YSIM_ARG(7.f / hargs.inv_h >= mx, "scene_set_collision: margin exceeds 7 hash cells");
    }
    const bool self_changed = self != P.self_collision;
    P.margin = margin;
    P.radius = radius;
    P.narrow_thickness = thick;
    P.self_collision = self;
    std::vector<float> mg, bd, th;
    collision_arrays(mg, bd, th);
    YSIM_HIP_TRY(hipMemcpyAsync(d_margin, mg.data(), 4 * (size_t)nm, hipMemcpyHostToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(d_band, bd.data(), 4 * (size_t)nm, hipMemcpyHostToDevice, st));
    YSIM_HIP_TRY(hipMemcpyAsync(d_thick, th.data(), 4 * (size_t)nm, hipMemcpyHostToDevice, st));
    YSIM_HIP_TRY(hipStreamSynchronize(st));
    if (self_changed) {
      hargs.self_collision = self;
      YSIM_TRY(build_targets());
      drop_graph();
    }
    return YSIM_OK;
  }

  int enqueue_analytic(uint32_t par) {
    if (!analytic_on() || !napairs) return YSIM_OK;
    AnalyticArgs a;
    a.dsegs = dyn_vsegs;
    a.mt = mt;
    a.shapes = d_ashapes;
    a.pair_off = d_apair_off;
    a.pair_shape = d_apair_shape;
    a.x = x[par];
    a.xprev = x[par ^ 1];
    a.tree = tree;
    a.mesh_band = d_band;
    a.mesh_thick = d_thick;
    a.radius = P.radius;
    a.mark_prev = amark + (size_t)(par ^ 1) * napairs;
    a.mark_cur = amark + (size_t)par * napairs;
    a.ccount = ccount;
    a.cnd = cnd;
    a.cid = cid;
    a.cap = cap;
    a.flags = flags;
    a.nitems = Vdyn;
    return launch_analytic(st, a);
  }
  // previous-substep pairing flags (those the next substep reads) to / from host
  int save_amark(std::vector<uint32_t>& out) {
    out.assign(napairs, 0u);
    if (!napairs) return YSIM_OK;
    YSIM_HIP_TRY(hipMemcpyAsync(out.data(), amark + (size_t)(parity ^ 1) * napairs, 4 * (size_t)napairs,
                                hipMemcpyDeviceToHost, st));
    YSIM_HIP_TRY(hipStreamSynchronize(st));
    return YSIM_OK;
  }
  int load_amark(const std::vector<uint32_t>& in, uint32_t par) {
    if (!napairs || in.size() != napairs) return YSIM_OK;
    YSIM_HIP_TRY(hipMemcpyAsync(amark + (size_t)(par ^ 1) * napairs, in.data(), 4 * (size_t)napairs,
                                hipMemcpyHostToDevice, st));
    YSIM_HIP_TRY(hipStreamSynchronize(st));
    return YSIM_OK;
  }

  int make_seg(const std::vector<uint32_t>& ms, bool tris, SegTable& out) {
    std::vector<uint32_t> off(1, 0);
    for (uint32_t m : ms) off.push_back(off.back() + (tris ? meshes[m].num_facets : meshes[m].num_vertices));
    uint32_t *d_off = nullptr, *d_mesh = nullptr;
    YSIM_TRY(put(&d_off, off));
    YSIM_TRY(put(&d_mesh, ms));
    out.nseg = (uint32_t)ms.size();
    out.off = d_off;
    out.mesh = d_mesh;
    if (ms.size() == 1) {  // the one-segment values travel in the kernel arguments
      const uint32_t m = ms[0];
      out.one_mesh = m;
      out.one_s = 0;
      out.one_n = off[1];
      out.one_noff = noff[m];
      out.one_foff = foff[m];
      out.one_voff = voff[m];
    }
    return YSIM_OK;
  }

  // Groups of meshes whose keys (segment id + Morton code) fit in 64 bits.
  int make_groups(const std::vector<uint32_t>& ms, std::vector<BuildGroup>& out) {
    const uint32_t cb = P.morton_bits;
    std::vector<std::vector<uint32_t>> lists;
    std::vector<uint32_t> nonempty;
    for (uint32_t m : ms)
      if (meshes[m].num_facets) nonempty.push_back(m);
    if (nonempty.empty()) return YSIM_OK;
    if (nonempty.size() == 1 || bits_for((uint32_t)nonempty.size()) + cb <= 64) {
      lists.push_back(nonempty);
    } else {
      for (uint32_t m : nonempty) lists.push_back({m});
    }
    for (auto& l : lists) {
      BuildGroup g;
      g.meshes = l;
      YSIM_TRY(make_seg(l, true, g.tsegs));
      YSIM_TRY(make_seg(l, false, g.vsegs));
      for (uint32_t m : l) {
        g.T += meshes[m].num_facets;
        g.V += meshes[m].num_vertices;
      }
      g.code_bits = cb;
      uint32_t key_bits = cb + (l.size() > 1 ? bits_for((uint32_t)l.size()) : 0);
      g.key64 = key_bits > 32;
      g.plan = tune.sort_hybrid ? sort_plan_hybrid(g.T, key_bits, tune.sort_items)
                                : sort_plan(g.T, key_bits, tune.sort_digit_bits, tune.sort_items);
      out.push_back(g);
    }
    return YSIM_OK;
  }

  // agglo: the group's build is the fused agglomerative pass with leaf boxes
  // (x + v dt when swept), so the substep's refit is not needed after it
  int enqueue_build(BuildGroup& g, const float* xc, const float* vc = nullptr, bool agglo = false) {
    if (!g.T) return YSIM_OK;
    // the incremental Morton kernel reduces the cloth kernel's block boxes
    const bool parts_box = fuse_box && box_fresh && &g == &dyn_groups[0] && g.perm && tune.sort_incremental;
    if (!parts_box)
      YSIM_TRY(launch_box(st, g.vsegs, mt, xc, g.V, box_acc, g.meshes.size() == 1 ? g.meshes[0] : 0xffffffffu,
                          box_parts));
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[1], st));
    if (g.perm && tune.sort_incremental) {
      // keys in the last build's sorted order + LDS window sorts + exact check;
      // the full build (Morton in facet order + hybrid sort) runs only if the
      // check failed (its kernels return at once otherwise)
      YSIM_TRY(launch_morton_inc(st, g.tsegs, mt, xc, facets, box_acc, g.code_bits, g.key64, g.perm, keys[0],
                                 vals[0], g.T, need_full, parts_box ? cloth_parts : nullptr,
                                 parts_box ? parts_count : 0u, links.cut_cnt, links.cut));
      if (timing) YSIM_HIP_TRY(hipEventRecord(ev[2], st));
      YSIM_TRY(launch_incremental_fix(st, g.key64, keys[0], vals[0], g.T, need_full, inc_bcnt, inc_edges));
      const bool defer = capturing && !agglo;
      if (!defer) {
        YSIM_TRY(launch_morton(st, g.tsegs, mt, xc, facets, box_acc, g.code_bits, g.key64, keys[0], vals[0], g.T,
                               g.plan, hist, counters, epoch, need_full));
        YSIM_TRY(launch_sort_hybrid(st, g.plan, keys[0], vals[0], keys[1], vals[1], hist, status, counters, epoch,
                                    flags, need_full));
      }
      if (timing) YSIM_HIP_TRY(hipEventRecord(ev[3], st));
      if (defer) {
        YSIM_TRY(launch_hierarchy(st, g.tsegs, mt, keys[0], g.key64, vals[0], g.code_bits, tree, parent, visit, g.T,
                                  box_acc, g.tsegs.nseg, hist, (uint32_t)g.plan.hist_words, g.perm, links_for(g),
                                  HierarchySkip{need_full, d_stale}));
        tree_agglo = false;
      } else {
        YSIM_TRY(enqueue_hierarchy(g, 0, xc, vc, agglo));
      }
      last_sorted = 0;
      return YSIM_OK;
    }
    YSIM_TRY(launch_morton(st, g.tsegs, mt, xc, facets, box_acc, g.code_bits, g.key64, keys[0], vals[0], g.T,
                           g.plan, hist, counters, epoch, nullptr, links.cut_cnt, links.cut));
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[2], st));
    if (g.plan.hybrid)
      YSIM_TRY(launch_sort_hybrid(st, g.plan, keys[0], vals[0], keys[1], vals[1], hist, status, counters, epoch,
                                  flags));
    else
      YSIM_TRY(launch_sort_passes(st, g.plan, keys[0], vals[0], keys[1], vals[1], hist, status, counters, epoch,
                                  flags));
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[3], st));
    const uint32_t sorted = (g.plan.hybrid || g.plan.passes % 2 == 0) ? 0 : 1;
    YSIM_TRY(enqueue_hierarchy(g, sorted, xc, vc, agglo));
    last_sorted = sorted;
    return YSIM_OK;
  }

  int enqueue_hierarchy(BuildGroup& g, uint32_t sorted, const float* xc, const float* vc, bool agglo) {
    if (agglo) {
      YSIM_TRY(launch_agglo(st, g.tsegs, mt, keys[sorted], g.key64, vals[sorted], g.code_bits, xc, vc, P.subh,
                            P.swept != 0 && vc, facets, tree, parent, visit, agg_range, g.T, box_acc, g.tsegs.nseg,
                            hist, (uint32_t)g.plan.hist_words, g.perm, flags, formal_refit));
      tree_agglo = true;
      // its parent links are not the Karras topology the hints describe
      YSIM_TRY(fill(links.hint, 2 * (size_t)Ntot, 0xffffffffu));
      if (dyn_groups.size() == 1 && &g == &dyn_groups[0]) span_rb_dyn = 0;
      return YSIM_OK;
    }
    YSIM_TRY(launch_hierarchy(st, g.tsegs, mt, keys[sorted], g.key64, vals[sorted], g.code_bits, tree, parent,
                              visit, g.T, box_acc, g.tsegs.nseg, hist, (uint32_t)g.plan.hist_words, g.perm,
                              links_for(g)));
    tree_agglo = false;
    return YSIM_OK;
  }

  QueryArgs query_args(uint32_t par) {
    QueryArgs a;
    a.qsegs = dyn_vsegs;
    a.mt = mt;
    a.tgt_off = d_tgt_off;
    a.tgt = d_tgt;
    a.mesh_nfacets = d_nfacets;
    a.mesh_margin = d_margin;
    a.mesh_band = d_band;
    a.mesh_behavior = d_behavior;
    a.mesh_shape = d_shape;
    a.mesh_grid_n = d_grid_n;
    a.cut_cnt = links.cut_cnt;
    a.cut = links.cut;
    a.x = x[par];
    a.xprev = x[par ^ 1];
    a.facets = facets;
    a.tree = tree;
    a.parent = parent;
    a.nitems = Vdyn;
    a.cap = cap;
    a.ccount = ccount;
    a.cnd = cnd;
    a.cid = cid;
    a.npairs = npairs;
    a.wstats = wstats;
    a.flags = flags;
    a.block = tune.query_block;
    a.waves = tune.query_waves;
    a.tile = tune.query_tile;
    if (dyn.size() == 1 && h_tgt_off.size() == 2 && h_tgt_off[1] > h_tgt_off[0]) {
      const uint32_t m = dyn[0], t = h_tgt[h_tgt_off[0]];
      a.pre_t = (int32_t)t;
      a.seg0_mesh = m;
      a.seg0_voff = voff[m];
      a.seg0_grid_n = h_grid_n[m];
      a.pre_tree = tree + noff[t];
    }
    return a;
  }

  // The audit fields of a query launch; clears the record count first (a
  // kernel, so graph replays clear it too).
  int audited(QueryArgs& a) {
    if (!audit) return YSIM_OK;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, audit_cnt, 1u);
    YSIM_CHECK_LAUNCH();
    a.audit = audit;
    a.audit_cnt = audit_cnt;
    a.audit_cap = audit_cap;
    return YSIM_OK;
  }

  int enqueue_collision(uint32_t par, bool rebuild) {
    const float* xc = x[par];
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[0], st));
    if (hash_mode) {
      if (timing)
        for (int e = 1; e <= 4; ++e) YSIM_HIP_TRY(hipEventRecord(ev[e], st));
      HashArgs h = hargs;
      h.x = xc;
      h.v = v[par];
      YSIM_TRY(launch_hash_build(st, h));
      if (timing) YSIM_HIP_TRY(hipEventRecord(ev[5], st));
      return YSIM_OK;
    }
    const bool agglo = rebuild && tune.build_mode == 1;
    if (rebuild) {
      for (auto& g : dyn_groups) YSIM_TRY(enqueue_build(g, xc, v[par], agglo));
    } else if (timing) {
      for (int e = 1; e <= 3; ++e) YSIM_HIP_TRY(hipEventRecord(ev[e], st));
    }
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[4], st));
    // the agglomerative build already wrote the swept boxes; a tree built by
    // it has no refit links, so its refits climb through parent links
    if (!agglo)
      YSIM_TRY(launch_refit(st, dyn_tsegs, mt, xc, v[par], P.subh, P.swept != 0, facets, tree, parent, visit, Tdyn,
                            flags, formal_refit, tree_agglo ? nullptr : &links, tune.refit_block,
                            tree_agglo ? nullptr : span_for_dyn()));
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[5], st));
    return YSIM_OK;
  }

  // With one dynamic mesh the tile query runs each tile's cloth step itself
  // (one kernel for query and integrate) unless analytic contacts join the
  // slots after it.  Scenes of many meshes keep the cloth kernel: its
  // occupancy is higher than the query's, which matters when many waves
  // take the per-lane walk (measured on the batched scene).
  bool fuse_query_cloth() const {
    return fuse_box && !hash_mode && tune.query_tile && !analytic_on() && Vdyn > 0;
  }
  // box records the cloth step writes in the current configuration
  uint32_t parts_for_config() const {
    return fuse_query_cloth() ? query_tiles(Vdyn) : cloth_box_parts(Vdyn, tune.cloth_block);
  }

  int enqueue_substep(uint32_t par, bool rebuild) {
    YSIM_TRY(enqueue_collision(par, rebuild));
    if (fuse_query_cloth()) {
      ClothArgs ca = cloth_args(par);
      if (fuse_box) ca.box_parts = cloth_parts;
      QueryArgs qa = query_args(par);
      YSIM_TRY(audited(qa));
      YSIM_TRY(launch_query_cloth(st, qa, ca));
      parts_count = query_tiles(Vdyn);
      if (fuse_box) box_fresh = true;
      if (timing) {
        YSIM_HIP_TRY(hipEventRecord(ev[6], st));
        YSIM_HIP_TRY(hipEventRecord(ev[7], st));
      }
      return YSIM_OK;
    }
    YSIM_TRY(enqueue_query(par));
    return enqueue_integrate(par);
  }

  int enqueue_query(uint32_t par) {
    QueryArgs qa = query_args(par);
    YSIM_TRY(audited(qa));
    if (hash_mode) {
      HashArgs h = hargs;
      h.x = x[par];
      YSIM_TRY(launch_hash_query(st, h, qa, false));
    } else {
      YSIM_TRY(launch_query(st, qa, false));
      YSIM_TRY(enqueue_analytic(par));
    }
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[6], st));
    return YSIM_OK;
  }

  int enqueue_integrate(uint32_t par) {
    ClothArgs ca = cloth_args(par);
    if (fuse_box) ca.box_parts = cloth_parts;
    YSIM_TRY(launch_cloth(st, ca));
    parts_count = cloth_box_parts(Vdyn, tune.cloth_block);
    if (fuse_box) box_fresh = true;
    if (timing) YSIM_HIP_TRY(hipEventRecord(ev[7], st));
    return YSIM_OK;
  }

  ClothArgs cloth_args(uint32_t par) const {
    ClothArgs ca;
    ca.dsegs = dyn_vsegs;
    ca.mt = mt;
    ca.cloth = d_cloth;
    ca.behavior = d_behavior;
    ca.tri_off = tri_off;
    ca.tri_mid = tri_mid;
    ca.tri_ent = tri_ent;
    ca.kspring = kspring;
    ca.subh = P.subh;
    ca.kair = P.kair;
    ca.kd = P.kd;
    ca.x = x[par];
    ca.v = v[par];
    ca.ext = ext;
    ca.mass = mass;
    ca.fixed = fixed;
    ca.ccount = ccount;
    ca.cnd = cnd;
    ca.cid = cid;
    ca.cap = cap;
    ca.x_out = x[par ^ 1];
    ca.v_out = v[par ^ 1];
    ca.nitems = Vdyn;
    ca.flags = flags;
    ca.block = tune.cloth_block;
    return ca;
  }

  bool rebuild_at(uint64_t s) const { return P.rebuild_every > 0 && (s % P.rebuild_every) == 0; }

  int eager_steps(uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
      YSIM_TRY(enqueue_substep(parity, rebuild_at(step)));
      if (timing) {
        YSIM_HIP_TRY(hipEventSynchronize(ev[7]));
        for (int k = 0; k < YSIM_STAGE_COUNT; ++k) {
          float ms = 0.f;
          YSIM_HIP_TRY(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
          stage_ms[k] += ms;
        }
      }
      parity ^= 1u;
      ++step;
      ++steps_run;
    }
    return YSIM_OK;
  }

  int capture_graph() {
    if (graph) return YSIM_OK;
    uint32_t period = 2;
    if (P.rebuild_every > 1) period = (P.rebuild_every % 2 == 0) ? P.rebuild_every : 2 * P.rebuild_every;
    hipGraph_t g = nullptr;
    YSIM_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = YSIM_OK;
    // a replay always follows a substep (box_fresh), so its first build
    // skips the box kernel too; the host flag is left as it was
    const bool fresh = box_fresh;
    const uint32_t parts = parts_count;
    // the host's view of the tree in memory (which builder made it, which
    // key buffer holds its order, the block size its spanning nodes were
    // recorded for) changes as builds are enqueued; captured builds run only
    // when the graph is replayed, so the view is kept as it was and the
    // view after a replay is applied at each launch
    const bool agglo0 = tree_agglo;
    const uint32_t sorted0 = last_sorted, span0 = span_rb_dyn;
    box_fresh = true;
    parts_count = parts_for_config();
    capturing = true;
    for (uint32_t i = 0; i < period && rc == YSIM_OK; ++i)
      rc = enqueue_substep(i % 2, rebuild_at(i));
    capturing = false;
    box_fresh = fresh;
    parts_count = parts;
    graph_tree_agglo = tree_agglo;
    graph_last_sorted = last_sorted;
    graph_span_rb = span_rb_dyn;
    tree_agglo = agglo0;
    last_sorted = sorted0;
    span_rb_dyn = span0;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != YSIM_OK) return rc;
    if (e != hipSuccess) return fail(YSIM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    YSIM_HIP_TRY(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
    YSIM_HIP_TRY(hipGraphDestroy(g));
    graph_period = period;
    graph_kd = P.kd;
    graph_ks = kspring;
    return YSIM_OK;
  }

  int steps(uint32_t n, bool use_graph) {
    if (!use_graph) {
      full_next = false;
      return eager_steps(n);
    }
    YSIM_TRY(capture_graph());
    while (n) {
      // a jumped state, or a replay that kept a previous topology, takes one
      // eager substep (its build runs the fallback sort when needed)
      if (*(volatile uint32_t*)h_stale) {
        *(volatile uint32_t*)h_stale = 0u;
        ++deferred_builds;
        full_next = true;
      }
      const bool box_ok = !fuse_box || (box_fresh && parts_count == parts_for_config());
      if (full_next || !box_ok || (step % graph_period) != 0 || n < graph_period) {
        YSIM_TRY(eager_steps(1));
        full_next = false;
        --n;
        continue;
      }
      YSIM_HIP_TRY(hipGraphLaunch(graph, st));
      tree_agglo = graph_tree_agglo;
      last_sorted = graph_last_sorted;
      span_rb_dyn = graph_span_rb;
      step += graph_period;
      steps_run += graph_period;
      n -= graph_period;
    }
    return YSIM_OK;
  }

  // Teardown is best-effort: every handle is released whatever an earlier
  // release returned (the destroy entry point reports the stream's state).
  ~ysim_scene() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (h_stale) (void)hipHostFree(h_stale);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
  }
};

// ====================================================================== analytic colliders
namespace {

// Hamilton product and rotateVector (include/sim/material_quat.hpp:28-35,
// 59-63): q * (0, v) * conj(q), each component summed left to right.
struct HQuat {
  float w, x, y, z;
};
HQuat hmul(HQuat a, HQuat b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
f3 rotate_vec(HQuat q, f3 v) {
  const HQuat r = hmul(hmul(q, HQuat{0.f, v.x, v.y, v.z}), HQuat{q.w, -q.x, -q.y, -q.z});
  return {r.x, r.y, r.z};
}

// World half-extents of a shape's bounding box at orientation q (w, x, y, z):
// BroadPhase::analyticWorldHalf (include/sim/bvh.hpp:2431-2460).  A sphere
// takes its largest semi-axis on every axis; box and cylinder (r, hh, r) go
// through the |R| h sum.
f3 world_half(const ysim_analytic_shape& s, const float* q) {
  const float* H = s.half_ext_height;
  f3 h{0.f, 0.f, 0.f};
  if (s.kind == YSIM_COLLIDER_SPHERE) {
    const float r = std::max(std::max(H[0], H[1]), H[2]);
    return {r, r, r};
  }
  if (s.kind == YSIM_COLLIDER_BOX)
    h = {H[0], H[1], H[2]};
  else if (s.kind == YSIM_COLLIDER_CYLINDER)
    h = {s.center_radius[3], H[3], s.center_radius[3]};
  else
    return h;
  const HQuat Q{q[0], q[1], q[2], q[3]};
  const f3 cx = rotate_vec(Q, {1.f, 0.f, 0.f}), cy = rotate_vec(Q, {0.f, 1.f, 0.f}),
           cz = rotate_vec(Q, {0.f, 0.f, 1.f});
  return {std::fabs(cx.x) * h.x + std::fabs(cy.x) * h.y + std::fabs(cz.x) * h.z,
          std::fabs(cx.y) * h.x + std::fabs(cy.y) * h.y + std::fabs(cz.y) * h.z,
          std::fabs(cx.z) * h.x + std::fabs(cy.z) * h.y + std::fabs(cz.z) * h.z};
}

// The per-shape constants of the device record: swept box before padding
// (objectBroadAABB, bvh.hpp:2505-2517), broad rotation band (analyticRotBand,
// bvh.hpp:2465-2480) and the narrow rotation margin's angle and radius
// (BruteForce::analyticRotMargin, bruteforce.hpp:201-242), all from the
// geodesic angle 2 acos(|q . q_prev|) taken in double.
AShape analytic_consts(const ysim_analytic_shape& sh) {
  AShape a{};
  a.s = sh;
  const float* H = sh.half_ext_height;
  const float* C = sh.center_radius;
  const float* Pc = sh.prev_center_pad;
  const f3 hc = world_half(sh, sh.rot_quat), hp = world_half(sh, sh.prev_rot_quat);
  const float hcv[3] = {hc.x, hc.y, hc.z}, hpv[3] = {hp.x, hp.y, hp.z};
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::min(C[k] - hcv[k], Pc[k] - hpv[k]);
    a.hi[k] = std::max(C[k] + hcv[k], Pc[k] + hpv[k]);
  }
  double dq = (double)sh.rot_quat[0] * (double)sh.prev_rot_quat[0] +
              (double)sh.rot_quat[1] * (double)sh.prev_rot_quat[1] +
              (double)sh.rot_quat[2] * (double)sh.prev_rot_quat[2] +
              (double)sh.rot_quat[3] * (double)sh.prev_rot_quat[3];
  dq = std::fabs(dq);
  if (dq > 1.0) dq = 1.0;
  const double theta = 2.0 * std::acos(dq);
  const double hn = std::sqrt((double)H[0] * H[0] + (double)H[1] * H[1] + (double)H[2] * H[2]);
  a.rot_band = theta <= 0.0 ? 0.f : (float)(theta * (hn + (double)C[3]));
  a.theta = (theta > 1e-9 && sh.kind != YSIM_COLLIDER_PLANE) ? theta : 0.0;
  if (sh.kind == YSIM_COLLIDER_SPHERE)
    a.rmax = std::max(std::max((double)H[0], (double)H[1]), (double)H[2]);
  else if (sh.kind == YSIM_COLLIDER_BOX)
    a.rmax = hn;
  else if (sh.kind == YSIM_COLLIDER_CYLINDER)
    a.rmax = std::sqrt((double)C[3] * C[3] + (double)H[3] * H[3]);
  return a;
}

}  // namespace

int ysim_scene::set_analytic(const ysim_analytic_shape* h, uint32_t n) {
  // shapes feed the query, not the refit: a staged substep may set them until its query
  if (stage_phase == 2) return fail(YSIM_ERR_STATE, "scene_set_analytic: the substep's query already ran");
  YSIM_ARG(n == 0 || h, "scene_set_analytic: null shapes");
  std::vector<int> so(nm, -1);
  for (uint32_t k = 0; k < n; ++k) {
    const ysim_analytic_shape& sh = h[k];
    YSIM_ARG(sh.kind >= YSIM_COLLIDER_SPHERE && sh.kind <= YSIM_COLLIDER_PLANE,
             "scene_set_analytic: kind must be sphere, box, cylinder or plane");
    YSIM_ARG(sh.obj_index < nm, "scene_set_analytic: obj_index out of range");
    const ysim_mesh_desc& d = meshes[sh.obj_index];
    YSIM_ARG(d.behavior == YSIM_BEHAVIOR_FLOAT, "scene_set_analytic: a collider's mesh must be a static (FLOAT) mesh");
    YSIM_ARG(d.shape == sh.kind, "scene_set_analytic: the mesh's shape lane must equal the collider kind");
    YSIM_ARG(sh.behavior_type == d.behavior, "scene_set_analytic: behavior_type must equal the mesh's behavior");
    YSIM_ARG(so[sh.obj_index] < 0, "scene_set_analytic: two shapes for one mesh");
    so[sh.obj_index] = (int)k;
  }
  bool same = n == ashapes.size();
  for (uint32_t k = 0; same && k < n; ++k) same = h[k].obj_index == ashapes[k].obj_index;
  std::vector<AShape> hs(n);
  for (uint32_t k = 0; k < n; ++k) hs[k] = analytic_consts(h[k]);
  ashapes.assign(h, h + n);
  shape_of = so;
  if (!same) {
    std::vector<uint32_t> off(1, 0), ps;
    for (uint32_t q : dyn) {
      const ysim_mesh_desc& Q = meshes[q];
      if (Q.collidable)
        for (uint32_t k = 0; k < n; ++k) {
          const ysim_mesh_desc& T = meshes[h[k].obj_index];
          if (T.collidable && T.group == Q.group) ps.push_back(k);
        }
      off.push_back((uint32_t)ps.size());
    }
    YSIM_TRY(release(d_apair_off));
    YSIM_TRY(release(d_apair_shape));
    YSIM_TRY(release(amark));
    d_apair_off = d_apair_shape = amark = nullptr;
    YSIM_TRY(put(&d_apair_off, off));
    YSIM_TRY(put(&d_apair_shape, ps));
    napairs = (uint32_t)ps.size();
    YSIM_TRY(alloc(&amark, 2 * (size_t)napairs));
    YSIM_TRY(zero(amark, 2 * (size_t)napairs));
    snap_amark.clear();
    tune_amark.clear();
    YSIM_TRY(build_targets());
    drop_graph();
    box_fresh = false;  // the cloth step may change kernels (and box records)
  }
  if (n > ashape_cap) {
    YSIM_TRY(release(d_ashapes));
    d_ashapes = nullptr;
    YSIM_TRY(alloc(&d_ashapes, n));
    ashape_cap = n;
    drop_graph();
  }
  if (n) YSIM_HIP_TRY(hipMemcpyAsync(d_ashapes, hs.data(), n * sizeof(AShape), hipMemcpyHostToDevice, st));
  YSIM_HIP_TRY(hipStreamSynchronize(st));
  return YSIM_OK;
}

// ====================================================================== C ABI
extern "C" {

const char* ysim_hip_last_error(void) { return g_err.c_str(); }
int ysim_hip_version(void) { return 10000; }

int ysim_hip_device_count(int* out) {
  YSIM_ARG(out, "null out");
  YSIM_HIP_TRY(hipGetDeviceCount(out));
  return YSIM_OK;
}

// ------------------------------------------------------------------ static scratch
namespace {
// The allocation-free sort's scratch (histograms, look-back status words and
// the ping-pong copy of keys and values): a file-scope device array, so the
// entry point that uses it calls no allocator.  Each GPU the library runs on
// holds its own copy, at its own address, until the process ends.  In host
// code the array's name is only a stand-in: kernels and copies get the address
// hipGetSymbolAddress returned for the current device (noalloc_ptr).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)YSIM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char noalloc_scratch[NOALLOC_SCRATCH_BYTES];
std::mutex noalloc_mu;
constexpr int NOALLOC_DEVICES = 64;
void* noalloc_ptr[NOALLOC_DEVICES] = {};
hipError_t noalloc_rc[NOALLOC_DEVICES] = {};

// Looks up this device's copy once (the device must be current).
void noalloc_setup(ysim_ctx* c) {
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (c->device >= NOALLOC_DEVICES) {
    c->noalloc_err = (int)hipErrorInvalidDevice;
    return;
  }
  if (!noalloc_ptr[c->device] && noalloc_rc[c->device] == hipSuccess) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e == hipSuccess && p)
      noalloc_ptr[c->device] = p;
    else
      noalloc_rc[c->device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
  }
  c->noalloc = noalloc_ptr[c->device];
  c->noalloc_err = (int)noalloc_rc[c->device];
}
}  // namespace

int ysim_hip_ctx_create(int device, uint32_t flags, ysim_ctx** out) {
  YSIM_ARG(out, "null out");
  *out = nullptr;
  int n = 0;
  YSIM_HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(YSIM_ERR_INVALID, "no such HIP device");
  YSIM_HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  YSIM_HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    return fail(YSIM_ERR_UNSUPPORTED, std::string("library built for gfx950, device is ") + prop.gcnArchName);
  ysim_ctx* c = new ysim_ctx();
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  c->formal_refit = (flags & YSIM_CTX_FORMAL_REFIT) != 0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(YSIM_ERR_HIP, "hipStreamCreate failed");
  }
  noalloc_setup(c);
  *out = c;
  return YSIM_OK;
}

int ysim_hip_ctx_destroy(ysim_ctx* ctx) {
  if (!ctx) return YSIM_OK;
  // Everything is released even when the stream reports an earlier failure;
  // that failure is then the return value.
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  if (e != hipSuccess) return fail(YSIM_ERR_HIP, std::string("ctx_destroy: ") + hipGetErrorString(e));
  return YSIM_OK;
}

int ysim_hip_sync(ysim_ctx* ctx) {
  YSIM_ARG(ctx, "null ctx");
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}

int ysim_hip_alloc(ysim_ctx* ctx, size_t bytes, void** d_out) {
  YSIM_ARG(ctx && d_out, "null argument");
  YSIM_HIP_TRY(hipMalloc(d_out, std::max<size_t>(bytes, 1)));
  return YSIM_OK;
}
int ysim_hip_free(ysim_ctx* ctx, void* d_ptr) {
  YSIM_ARG(ctx, "null ctx");
  if (d_ptr) {
    YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    YSIM_HIP_TRY(hipFree(d_ptr));
  }
  return YSIM_OK;
}
int ysim_hip_upload(ysim_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
  YSIM_ARG(ctx && (bytes == 0 || (d_dst && h_src)), "null argument");
  if (bytes) YSIM_HIP_TRY(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}
int ysim_hip_download(ysim_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
  YSIM_ARG(ctx && (bytes == 0 || (d_src && h_dst)), "null argument");
  if (bytes) YSIM_HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}
int ysim_hip_memset(ysim_ctx* ctx, void* d_dst, int value, size_t bytes) {
  YSIM_ARG(ctx && (bytes == 0 || d_dst), "null argument");
  if (bytes) YSIM_HIP_TRY(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
  return YSIM_OK;
}

// ------------------------------------------------------------------ stages
namespace {
// single-segment tables for stage entry points: off = {0, n}, mesh = {0},
// voff = foff = noff = {0, ...}
struct Unit {
  uint32_t* d = nullptr;  // [0]=0 [1]=n [2]=0 (mesh id) [3]=0 [4]=0 [5]=0
  SegTable seg() const { return SegTable{1, d, d + 2}; }
  MeshTables mt() const { return MeshTables{d + 3, d + 4, d + 5}; }
};
int unit_table(ysim_ctx* ctx, uint32_t n, Unit& u) {
  YSIM_TRY(ctx->ensure_scratch(1 << 20));
  u.d = (uint32_t*)ctx->scratch;
  uint32_t h[8] = {0, n, 0, 0, 0, 0, 0, 0};
  YSIM_HIP_TRY(hipMemcpyAsync(u.d, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  return YSIM_OK;
}
uint32_t* scratch_words(ysim_ctx* ctx, size_t off_words) { return (uint32_t*)ctx->scratch + off_words; }
}  // namespace

int ysim_hip_scene_box(ysim_ctx* ctx, const float* d_x, uint32_t num_vertices, uint32_t num_prims,
                       ysim_aabb4* d_box) {
  YSIM_ARG(ctx && d_x && d_box && num_vertices > 0, "scene_box: bad argument");
  Unit u;
  YSIM_TRY(unit_table(ctx, num_vertices, u));
  uint32_t* acc = scratch_words(ctx, 64);
  uint32_t init[6] = {~0u, ~0u, ~0u, 0u, 0u, 0u};
  YSIM_HIP_TRY(hipMemcpyAsync(acc, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  YSIM_TRY(launch_box(ctx->stream, u.seg(), u.mt(), d_x, num_vertices, acc, 0u));
  uint32_t h[6];
  YSIM_HIP_TRY(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  ysim_aabb4 b;
  auto un = [](uint32_t k) {
    uint32_t uu = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &uu, 4);
    return f;
  };
  for (int k = 0; k < 3; ++k) {
    b.min[k] = un(h[k]);
    b.max[k] = un(h[3 + k]);
  }
  b.i0 = (int32_t)num_prims;
  b.i1 = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(d_box, &b, sizeof(b), hipMemcpyHostToDevice, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}

int ysim_hip_morton(ysim_ctx* ctx, const float* d_x, const uint32_t* d_facets, uint32_t num_facets,
                    const ysim_aabb4* d_box, uint32_t bits, void* d_keys, uint32_t* d_idx) {
  YSIM_ARG(ctx && d_x && d_facets && d_box && d_keys && d_idx, "morton: null argument");
  YSIM_ARG(bits == 30 || bits == 63, "morton: bits must be 30 or 63");
  if (!num_facets) return YSIM_OK;
  Unit u;
  YSIM_TRY(unit_table(ctx, num_facets, u));
  // box in ordered-key form
  ysim_aabb4 hb;
  YSIM_HIP_TRY(hipMemcpyAsync(&hb, d_box, sizeof(hb), hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  auto key = [](float f) {
    uint32_t uu;
    std::memcpy(&uu, &f, 4);
    return (uu & 0x80000000u) ? ~uu : (uu | 0x80000000u);
  };
  uint32_t acc_h[6] = {key(hb.min[0]), key(hb.min[1]), key(hb.min[2]), key(hb.max[0]), key(hb.max[1]), key(hb.max[2])};
  uint32_t* acc = scratch_words(ctx, 64);
  uint32_t* hist = scratch_words(ctx, 128);  // unused: no histogram passes
  YSIM_HIP_TRY(hipMemcpyAsync(acc, acc_h, sizeof(acc_h), hipMemcpyHostToDevice, ctx->stream));
  SortPlan none = sort_plan(0, bits);
  none.passes = 0;
  // vertex count is not needed: facets index into d_x directly
  YSIM_TRY(launch_morton(ctx->stream, u.seg(), u.mt(), d_x, d_facets, acc, bits, bits > 32, d_keys, d_idx,
                         num_facets, none, hist, nullptr, nullptr));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}

int ysim_hip_sort_pairs(ysim_ctx* ctx, void* d_keys, uint32_t* d_vals, uint32_t n, uint32_t key_bits) {
  return ysim_hip_sort_pairs_ex(ctx, d_keys, d_vals, n, key_bits, 8, 8, 0);
}

namespace {
// Scratch layout of one sort of plan p:
// [4096 words misc][hist][counters][err 16][epoch 16][status (64-bit words)][pad to 256 B]
// [keys_b, 256-B rounded][vals_b]
size_t sort_scratch_bytes(const SortPlan& p) {
  const size_t kb = (p.wide ? 8 : 4);
  const size_t tmp_bytes = 256 + (((size_t)p.n * kb + 255) & ~(size_t)255) + (size_t)p.n * 4 + 256;
  return 4 * (4096 + p.hist_words + p.counter_words + 32 + 2 * p.status_words + tmp_bytes / 4 + 2);
}

// The sort of plan p on d_keys / d_vals with its scratch at base (at least
// sort_scratch_bytes(p), any contents: the sort clears what must start at zero).
int sort_in(ysim_ctx* ctx, void* base_, const SortPlan& p, void* d_keys, uint32_t* d_vals) {
  const uint32_t n = p.n;
  const size_t kb = (p.wide ? 8 : 4);
  uint32_t* base = (uint32_t*)base_;
  uint32_t* hist = base + 4096;
  uint32_t* counters = hist + p.hist_words;
  uint32_t* err = counters + p.counter_words;
  uint32_t* epoch = err + 16;
  uint64_t* status = (uint64_t*)(((uintptr_t)(epoch + 16) + 7) & ~(uintptr_t)7);
  char* tmp = (char*)(status + p.status_words);
  tmp = (char*)(((uintptr_t)tmp + 255) & ~(uintptr_t)255);
  void* keys_b = tmp;
  uint32_t* vals_b = (uint32_t*)(tmp + ((n * kb + 255) & ~(size_t)255));
  YSIM_TRY(launch_sort_reset(ctx->stream, p, hist, status, counters, epoch, true));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, err, 16u);
  YSIM_TRY(launch_sort_histogram(ctx->stream, p, d_keys, hist));
  if (p.hybrid)
    YSIM_TRY(launch_sort_hybrid(ctx->stream, p, d_keys, d_vals, keys_b, vals_b, hist, status, counters, epoch, err));
  else
    YSIM_TRY(launch_sort_passes(ctx->stream, p, d_keys, d_vals, keys_b, vals_b, hist, status, counters, epoch, err));
  if (!p.hybrid && p.passes % 2 == 1) {
    YSIM_HIP_TRY(hipMemcpyAsync(d_keys, keys_b, n * kb, hipMemcpyDeviceToDevice, ctx->stream));
    YSIM_HIP_TRY(hipMemcpyAsync(d_vals, vals_b, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  uint32_t herr = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (herr) return fail(YSIM_ERR_HIP, "sort_pairs: look-back guard fired");
  return YSIM_OK;
}
}  // namespace

int ysim_hip_sort_pairs_ex(ysim_ctx* ctx, void* d_keys, uint32_t* d_vals, uint32_t n, uint32_t key_bits,
                           uint32_t digit_bits, uint32_t items, uint32_t hybrid) {
  YSIM_ARG(ctx && (n == 0 || (d_keys && d_vals)), "sort_pairs: null argument");
  YSIM_ARG(key_bits >= 1 && key_bits <= 64, "sort_pairs: key_bits in [1, 64]");
  YSIM_ARG(hybrid ? (items == 8 || items == 16) : sort_config_ok(digit_bits, items),
           "sort_pairs: digit_bits 8 or 11, items 8 or 16");
  if (n == 0) return YSIM_OK;
  YSIM_ARG(n < (1u << 30), "sort_pairs: n must be < 2^30");
  const SortPlan p = hybrid ? sort_plan_hybrid(n, key_bits, items) : sort_plan(n, key_bits, digit_bits, items);
  YSIM_TRY(ctx->ensure_scratch(sort_scratch_bytes(p)));
  return sort_in(ctx, ctx->scratch, p, d_keys, d_vals);
}

int ysim_hip_sort_pairs_noalloc(ysim_ctx* ctx, void* d_keys, uint32_t* d_vals, uint32_t n, uint32_t key_bits) {
  YSIM_ARG(ctx && (n == 0 || (d_keys && d_vals)), "sort_pairs_noalloc: null argument");
  YSIM_ARG(key_bits >= 1 && key_bits <= 64, "sort_pairs_noalloc: key_bits in [1, 64]");
  if (n == 0) return YSIM_OK;
  YSIM_ARG(n < (1u << 30), "sort_pairs_noalloc: n must be < 2^30");
  if (!ctx->noalloc)
    return fail(YSIM_ERR_HIP, std::string("sort_pairs_noalloc: the static scratch of this device has no address (") +
                                  hipGetErrorString((hipError_t)ctx->noalloc_err) + ")");
  // the same plan as ysim_hip_sort_pairs (8-bit digits, 8 keys per thread)
  const SortPlan p = sort_plan(n, key_bits, 8, 8);
  const size_t need = sort_scratch_bytes(p);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(YSIM_ERR_CAPACITY, "sort_pairs_noalloc: " + std::to_string(n) + " keys of " +
                                       std::to_string(key_bits) + " bits need " + std::to_string(need) +
                                       " bytes of scratch; the static scratch holds " +
                                       std::to_string(NOALLOC_SCRATCH_BYTES) + " (use ysim_hip_sort_pairs)");
  return sort_in(ctx, ctx->noalloc, p, d_keys, d_vals);
}

int ysim_hip_lbvh_hierarchy(ysim_ctx* ctx, const void* d_sorted_keys, const uint32_t* d_sorted_idx, uint32_t n,
                            uint32_t bits, ysim_bvh_node* d_tree, int32_t* d_parent) {
  YSIM_ARG(ctx && d_sorted_keys && d_sorted_idx && d_tree && d_parent, "lbvh_hierarchy: null argument");
  YSIM_ARG(bits == 30 || bits == 63 || bits == 32 || bits == 64, "lbvh_hierarchy: bits");
  if (!n) return YSIM_OK;
  YSIM_TRY(ctx->ensure_scratch((size_t)(1 << 20) + (size_t)n * 8 + 4096));
  Unit u;
  YSIM_TRY(unit_table(ctx, n, u));
  uint32_t* visit = (uint32_t*)((char*)ctx->scratch + (1 << 20));
  YSIM_TRY(launch_hierarchy(ctx->stream, u.seg(), u.mt(), d_sorted_keys, bits > 32, d_sorted_idx, bits, d_tree,
                            d_parent, visit, n, nullptr, 0, nullptr, 0));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}

int ysim_hip_lbvh_agglomerative(ysim_ctx* ctx, const void* d_sorted_keys, const uint32_t* d_sorted_idx, uint32_t n,
                                uint32_t bits, const float* d_x, const uint32_t* d_facets, ysim_bvh_node* d_tree,
                                int32_t* d_parent) {
  YSIM_ARG(ctx && d_sorted_keys && d_sorted_idx && d_x && d_facets && d_tree && d_parent,
           "lbvh_agglomerative: null argument");
  YSIM_ARG(bits == 30 || bits == 63 || bits == 32 || bits == 64, "lbvh_agglomerative: bits");
  if (!n) return YSIM_OK;
  YSIM_TRY(ctx->ensure_scratch((size_t)(1 << 20) + (size_t)n * 24 + 4096));
  Unit u;
  YSIM_TRY(unit_table(ctx, n, u));
  uint32_t* visit = (uint32_t*)((char*)ctx->scratch + (1 << 20));
  int2* range = (int2*)((char*)ctx->scratch + (1 << 20) + (((size_t)n * 8 + 255) & ~(size_t)255));
  uint32_t* err = scratch_words(ctx, 64);
  hipLaunchKernelGGL(k_zero_u32, dim3(ceil_div(2 * (size_t)n, 256)), dim3(256), 0, ctx->stream, visit, 2 * n);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, err, 1u);
  YSIM_TRY(launch_agglo(ctx->stream, u.seg(), u.mt(), d_sorted_keys, bits > 32, d_sorted_idx, bits, d_x, nullptr, 0.f,
                        false, d_facets, d_tree, d_parent, visit, range, n, nullptr, 0, nullptr, 0, nullptr, err,
                        ctx->formal_refit));
  uint32_t herr = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (herr) return fail(YSIM_ERR_INVALID, "lbvh_agglomerative: keys not sorted (guard fired)");
  return YSIM_OK;
}

int ysim_hip_lbvh_refit(ysim_ctx* ctx, const float* d_x, const float* d_v, float dt, uint32_t swept,
                        const uint32_t* d_facets, uint32_t num_facets, ysim_bvh_node* d_tree,
                        const int32_t* d_parent) {
  YSIM_ARG(ctx && d_x && d_facets && d_tree && d_parent, "lbvh_refit: null argument");
  YSIM_ARG(!swept || d_v, "lbvh_refit: swept needs velocities");
  if (!num_facets) return YSIM_OK;
  YSIM_TRY(ctx->ensure_scratch((size_t)(1 << 20) + (size_t)num_facets * 4 + 4096));
  Unit u;
  YSIM_TRY(unit_table(ctx, num_facets, u));
  uint32_t* visit = (uint32_t*)((char*)ctx->scratch + (1 << 20));
  uint32_t* err = scratch_words(ctx, 64);
  hipLaunchKernelGGL(k_zero_u32, dim3(std::max<uint32_t>(1, ceil_div(num_facets, 256))), dim3(256), 0, ctx->stream,
                     visit, num_facets);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, err, 1u);
  YSIM_TRY(launch_refit(ctx->stream, u.seg(), u.mt(), d_x, d_v ? d_v : d_x, dt, swept != 0, d_facets, d_tree,
                        d_parent, visit, num_facets, err, ctx->formal_refit));
  uint32_t herr = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (herr) return fail(YSIM_ERR_INVALID, "lbvh_refit: corrupted topology (guard fired)");
  return YSIM_OK;
}

namespace {
int broad_query(ysim_ctx* ctx, const float* d_xq, uint32_t num_query, uint32_t query_obj, uint32_t query_behavior,
                uint32_t query_shape, const ysim_bvh_node* d_tree, uint32_t num_target_facets,
                const uint32_t* d_target_facets, uint32_t target_obj, uint32_t target_behavior,
                uint32_t target_shape, float margin, ysim_broad_collision* d_out, uint32_t capacity,
                uint32_t* d_count, uint32_t* h_required, bool segmented) {
  YSIM_ARG(ctx && d_xq && d_tree && d_target_facets && d_count, "broad_query: null argument");
  YSIM_ARG(capacity == 0 || d_out, "broad_query: null output");
  if (!num_query || !num_target_facets) return YSIM_OK;
  // parent links are needed by the traversal: rebuild them from child links
  const uint32_t nn = 2 * num_target_facets - 1;
  YSIM_TRY(ctx->ensure_scratch((size_t)(1 << 20) + (size_t)nn * 4 + 4096));
  int32_t* par = (int32_t*)((char*)ctx->scratch + (1 << 20));
  uint32_t* flags = scratch_words(ctx, 64);
  std::vector<ysim_bvh_node> h(nn);
  YSIM_HIP_TRY(hipMemcpyAsync(h.data(), d_tree, nn * sizeof(ysim_bvh_node), hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  std::vector<int32_t> hp(nn, -1);
  for (uint32_t i = 0; i + 1 < num_target_facets; ++i) {
    int a = h[i].child_a, b = h[i].child_b;
    if (a < 0 || b < 0 || a >= (int)nn || b >= (int)nn) return fail(YSIM_ERR_INVALID, "broad_query: bad tree");
    hp[a] = (int32_t)i;
    hp[b] = (int32_t)i;
  }
  YSIM_HIP_TRY(hipMemcpyAsync(par, hp.data(), nn * 4, hipMemcpyHostToDevice, ctx->stream));
  uint32_t start = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&start, d_count, 4, hipMemcpyDeviceToHost, ctx->stream));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, flags, 1u);
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  YSIM_TRY(launch_broad_pairs(ctx->stream, d_xq, num_query, query_obj, query_behavior, query_shape, d_tree, par,
                              num_target_facets, d_target_facets, target_obj, target_behavior, target_shape, margin,
                              d_out, capacity, d_count, flags, segmented));
  uint32_t res[2] = {0, 0};
  YSIM_HIP_TRY(hipMemcpyAsync(&res[0], d_count, 4, hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipMemcpyAsync(&res[1], flags, 4, hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (h_required) *h_required = res[0];
  if (res[1]) return fail(YSIM_ERR_INVALID, "broad_query: traversal guard fired");
  if (res[0] > capacity) return fail(YSIM_ERR_CAPACITY, "broad_query: pair capacity exceeded");
  (void)start;
  return YSIM_OK;
}
}  // namespace

int ysim_hip_broad_query(ysim_ctx* ctx, const float* d_xq, uint32_t num_query, uint32_t query_obj,
                         uint32_t query_behavior, uint32_t query_shape, const ysim_bvh_node* d_tree,
                         uint32_t num_target_facets, const uint32_t* d_target_facets, uint32_t target_obj,
                         uint32_t target_behavior, uint32_t target_shape, float margin,
                         ysim_broad_collision* d_out, uint32_t capacity, uint32_t* d_count, uint32_t* h_required) {
  return broad_query(ctx, d_xq, num_query, query_obj, query_behavior, query_shape, d_tree, num_target_facets,
                     d_target_facets, target_obj, target_behavior, target_shape, margin, d_out, capacity, d_count,
                     h_required, false);
}

int ysim_hip_broad_query_segmented(ysim_ctx* ctx, const float* d_xq, uint32_t num_query, uint32_t query_obj,
                                   uint32_t query_behavior, uint32_t query_shape, const ysim_bvh_node* d_tree,
                                   uint32_t num_target_facets, const uint32_t* d_target_facets, uint32_t target_obj,
                                   uint32_t target_behavior, uint32_t target_shape, float margin,
                                   ysim_broad_collision* d_out, uint32_t capacity, uint32_t* d_count,
                                   uint32_t* h_required) {
  return broad_query(ctx, d_xq, num_query, query_obj, query_behavior, query_shape, d_tree, num_target_facets,
                     d_target_facets, target_obj, target_behavior, target_shape, margin, d_out, capacity, d_count,
                     h_required, true);
}

int ysim_hip_narrow(ysim_ctx* ctx, const ysim_broad_collision* d_pairs, uint32_t num_pairs, const float* d_x,
                    const float* d_xprev, const uint32_t* d_states_offsets, const uint32_t* d_facets,
                    const uint32_t* d_facets_offsets, const uint32_t* d_adj_offsets, const uint32_t* d_adj_facets,
                    float radius, float thickness, ysim_narrow_collision* d_out, uint32_t capacity,
                    uint32_t* d_count, uint32_t* h_required) {
  YSIM_ARG(ctx && d_x && d_xprev && d_states_offsets && d_facets && d_facets_offsets && d_count,
           "narrow: null argument");
  YSIM_ARG(num_pairs == 0 || d_pairs, "narrow: null pairs");
  YSIM_ARG(capacity == 0 || d_out, "narrow: null output");
  YSIM_TRY(launch_narrow_pairs(ctx->stream, d_pairs, num_pairs, d_x, d_xprev, d_states_offsets, d_facets,
                               d_facets_offsets, d_adj_offsets, d_adj_facets, radius + thickness, d_out, capacity,
                               d_count));
  uint32_t cnt = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&cnt, d_count, 4, hipMemcpyDeviceToHost, ctx->stream));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (h_required) *h_required = cnt;
  if (cnt > capacity) return fail(YSIM_ERR_CAPACITY, "narrow: contact capacity exceeded");
  return YSIM_OK;
}

int ysim_hip_contact_csr(ysim_ctx* ctx, const ysim_narrow_collision* d_contacts, uint32_t num_contacts,
                         const uint32_t* d_states_offsets, uint32_t num_points_total, uint32_t* d_offsets,
                         ysim_narrow_collision* d_sorted) {
  YSIM_ARG(ctx && d_states_offsets && d_offsets, "contact_csr: null argument");
  YSIM_ARG(num_contacts == 0 || (d_contacts && d_sorted), "contact_csr: null contacts");
  YSIM_TRY(ctx->ensure_scratch(((size_t)num_points_total + 2048) * 4));
  YSIM_TRY(launch_contact_csr(ctx->stream, d_contacts, num_contacts, d_states_offsets, num_points_total, d_offsets,
                              d_sorted, (uint32_t*)ctx->scratch));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}

int ysim_hip_cloth_step(ysim_ctx* ctx, const ysim_fastgrid_params* h_params, uint32_t num_vertices, float subh,
                        float kair, float kd, const float* d_x, const float* d_v, const float* d_ext,
                        const float* d_mass, const float* d_fixed, const uint32_t* d_csr_offsets,
                        const ysim_narrow_collision* d_csr, float* d_x_out, float* d_v_out, float* d_f_out,
                        uint32_t* d_anomaly) {
  YSIM_ARG(ctx && h_params && d_x && d_v && d_ext && d_mass && d_fixed && d_x_out && d_v_out && d_anomaly,
           "cloth_step: null argument");
  YSIM_ARG(d_x_out != d_x && d_v_out != d_v, "cloth_step: outputs must not alias inputs");
  YSIM_ARG(h_params->particle_num_1d >= 2 &&
               (uint64_t)h_params->particle_num_1d * h_params->particle_num_1d == num_vertices,
           "cloth_step: num_vertices must be particle_num_1d^2");
  Unit u;
  YSIM_TRY(unit_table(ctx, num_vertices, u));
  ysim_fastgrid_params* dp = (ysim_fastgrid_params*)scratch_words(ctx, 64);
  YSIM_HIP_TRY(hipMemcpyAsync(dp, h_params, sizeof(*h_params), hipMemcpyHostToDevice, ctx->stream));
  ClothArgs a;
  a.dsegs = u.seg();
  a.mt = u.mt();
  a.cloth = dp;
  a.subh = subh;
  a.kair = kair;
  a.kd = kd;
  a.x = d_x;
  a.v = d_v;
  a.ext = d_ext;
  a.mass = d_mass;
  a.fixed = d_fixed;
  a.csr_off = d_csr_offsets;
  a.csr = d_csr;
  a.x_out = d_x_out;
  a.v_out = d_v_out;
  a.f_out = d_f_out;
  a.nitems = num_vertices;
  a.flags = d_anomaly;
  if (!d_csr_offsets) a.cap = 0;
  YSIM_TRY(launch_cloth(ctx->stream, a));
  YSIM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return YSIM_OK;
}

// ------------------------------------------------------------------ scene
int ysim_hip_scene_create(ysim_ctx* ctx, uint32_t num_meshes, const ysim_mesh_desc* h_meshes, const float* h_x,
                          const float* h_v, const float* h_mass, const float* h_ext, const float* h_fixed,
                          const uint32_t* h_facets, const ysim_sim_params* h_params, ysim_scene** out) {
  YSIM_ARG(ctx && h_meshes && h_x && h_v && h_mass && h_ext && h_fixed && h_facets && h_params && out,
           "scene_create: null argument");
  YSIM_ARG(num_meshes > 0, "scene_create: no meshes");
  const ysim_sim_params& P = *h_params;
  YSIM_ARG(P.morton_bits == 30 || P.morton_bits == 63, "scene_create: morton_bits must be 30 or 63");
  YSIM_ARG(P.broad_mode == 0 || P.broad_mode == 1, "scene_create: broad_mode must be 0 (LBVH) or 1 (hash)");
  YSIM_ARG(P.broad_mode == 0 || (P.hash_slots_log2 >= 10 && P.hash_slots_log2 <= 26),
           "scene_create: hash_slots_log2 must be in [10, 26]");
  *out = nullptr;
  std::unique_ptr<ysim_scene> S(new ysim_scene());
  S->ctx = ctx;
  S->st = ctx->stream;
  S->P = P;
  S->formal_refit = ctx->formal_refit;
  S->nm = num_meshes;
  S->meshes.assign(h_meshes, h_meshes + num_meshes);
  S->cap = P.contact_capacity ? P.contact_capacity : 16;
  S->voff.assign(num_meshes + 1, 0);
  S->foff.assign(num_meshes + 1, 0);
  S->noff.assign(num_meshes + 1, 0);
  for (uint32_t m = 0; m < num_meshes; ++m) {
    const ysim_mesh_desc& d = S->meshes[m];
    YSIM_ARG(d.behavior == YSIM_BEHAVIOR_FLOAT || d.behavior == YSIM_BEHAVIOR_FASTGRID_CLOTH ||
                 d.behavior == YSIM_BEHAVIOR_TRIANGULAR_CLOTH,
             "scene_create: unsupported behavior");
    if (d.behavior == YSIM_BEHAVIOR_FASTGRID_CLOTH) {
      YSIM_ARG(d.cloth.particle_num_1d >= 3 &&
                   (uint64_t)d.cloth.particle_num_1d * d.cloth.particle_num_1d == d.num_vertices,
               "scene_create: FastGrid cloth needs num_vertices = N^2, N >= 3");
      S->dyn.push_back(m);
    } else if (d.behavior == YSIM_BEHAVIOR_TRIANGULAR_CLOTH) {
      YSIM_ARG(d.num_facets > 0, "scene_create: a triangular cloth needs facets");
      S->dyn.push_back(m);
    } else {
      S->stat.push_back(m);
    }
    S->voff[m + 1] = S->voff[m] + d.num_vertices;
    S->foff[m + 1] = S->foff[m] + d.num_facets;
    S->noff[m + 1] = S->noff[m] + (d.num_facets ? 2 * d.num_facets - 1 : 0);
  }
  S->Vtot = S->voff[num_meshes];
  S->Ttot = S->foff[num_meshes];
  S->Ntot = S->noff[num_meshes];
  YSIM_ARG((uint64_t)S->Vtot < (1ull << 31) && (uint64_t)S->Ntot < (1ull << 31), "scene_create: scene too large");
  for (uint32_t i = 0; i < 3 * S->Ttot; ++i) {
    uint32_t m = (uint32_t)(std::upper_bound(S->foff.begin(), S->foff.end(), i / 3) - S->foff.begin()) - 1;
    YSIM_ARG(h_facets[i] < S->meshes[m].num_vertices, "scene_create: facet index out of range");
  }
  ysim_scene& s = *S;
  const size_t V3 = 3 * (size_t)s.Vtot;
  for (int b = 0; b < 2; ++b) {
    YSIM_TRY(s.put(&s.x[b], h_x, V3));
    YSIM_TRY(s.put(&s.v[b], h_v, V3));
  }
  YSIM_TRY(s.put(&s.mass, h_mass, s.Vtot));
  YSIM_TRY(s.put(&s.ext, h_ext, V3));
  YSIM_TRY(s.put(&s.fixed, h_fixed, s.Vtot));
  YSIM_TRY(s.put(&s.facets, h_facets, 3 * (size_t)s.Ttot));
  uint32_t *dv, *df, *dn;
  YSIM_TRY(s.put(&dv, s.voff));
  YSIM_TRY(s.put(&df, s.foff));
  YSIM_TRY(s.put(&dn, s.noff));
  s.mt = MeshTables{dv, df, dn};
  std::vector<ysim_fastgrid_params> cloth(num_meshes);
  std::vector<uint32_t> nf(num_meshes), beh(num_meshes), shp(num_meshes);
  std::vector<float> margin, band, thick;
  for (uint32_t m = 0; m < num_meshes; ++m) {
    cloth[m] = s.meshes[m].cloth;
    nf[m] = s.meshes[m].num_facets;
    beh[m] = s.meshes[m].behavior;
    shp[m] = s.meshes[m].shape;
  }
  // margin <= 0: per-mesh margin and band = the cloth's thickness (batching)
  s.collision_arrays(margin, band, thick);
  YSIM_TRY(s.put(&s.d_cloth, cloth));
  YSIM_TRY(s.put(&s.d_nfacets, nf));
  YSIM_TRY(s.put(&s.d_behavior, beh));
  {
    // TriangularCloth springs: per global vertex its stretch partners, then
    // its bend partners (global index, rest length), in the reference's order
    std::vector<uint32_t> toff(s.Vtot + 1, 0u), tmid(s.Vtot, 0u);
    std::vector<uint2> tent;
    for (uint32_t m = 0; m < num_meshes; ++m) {
      const uint32_t vb = s.voff[m], nv = s.meshes[m].num_vertices;
      TriSprings ts;
      if (s.meshes[m].behavior == YSIM_BEHAVIOR_TRIANGULAR_CLOTH) {
        YSIM_ARG(tri_springs(h_facets + 3 * (size_t)s.foff[m], s.meshes[m].num_facets, h_x + 3 * (size_t)vb, nv, ts),
                 "scene_create: a triangular cloth needs a manifold mesh without repeated vertices in a facet");
      }
      for (uint32_t i = 0; i < nv; ++i) {
        toff[vb + i] = (uint32_t)tent.size();
        if (ts.first.empty()) {
          tmid[vb + i] = (uint32_t)tent.size();
          continue;
        }
        for (uint32_t k = ts.first[i]; k < ts.mid[i]; ++k)
          tent.push_back(make_uint2(vb + ts.other[k], float_bits(ts.rest[k])));
        tmid[vb + i] = (uint32_t)tent.size();
        for (uint32_t k = ts.mid[i]; k < ts.first[i + 1]; ++k)
          tent.push_back(make_uint2(vb + ts.other[k], float_bits(ts.rest[k])));
      }
    }
    toff[s.Vtot] = (uint32_t)tent.size();
    if (!tent.empty()) {
      YSIM_TRY(s.put(&s.tri_off, toff));
      YSIM_TRY(s.put(&s.tri_mid, tmid));
      YSIM_TRY(s.put(&s.tri_ent, tent));
    }
  }
  {
    std::vector<uint32_t> gn(num_meshes, 0u);
    for (uint32_t m = 0; m < num_meshes; ++m) {
      const uint32_t N = s.meshes[m].cloth.particle_num_1d;
      if (s.meshes[m].behavior == YSIM_BEHAVIOR_FASTGRID_CLOTH && N >= 8 &&
          (uint64_t)N * N == s.meshes[m].num_vertices)
        gn[m] = N;
    }
    YSIM_TRY(s.put(&s.d_grid_n, gn));
    s.h_grid_n = gn;
  }
  YSIM_TRY(s.put(&s.d_shape, shp));
  YSIM_TRY(s.put(&s.d_margin, margin));
  YSIM_TRY(s.put(&s.d_band, band));
  YSIM_TRY(s.put(&s.d_thick, thick));
  YSIM_TRY(s.alloc(&s.tree, s.Ntot));
  YSIM_TRY(s.alloc(&s.parent, s.Ntot));
  YSIM_TRY(s.alloc(&s.visit, s.Ntot));
  YSIM_TRY(s.alloc(&s.links.up, s.Ntot));
  YSIM_TRY(s.alloc(&s.links.fid, s.Ntot));
  YSIM_TRY(s.alloc(&s.links.hint, s.Ntot));
  YSIM_TRY(s.fill(s.links.hint, 2 * (size_t)s.Ntot, 0xffffffffu));  // (-1, -1): no hint
  YSIM_TRY(s.alloc(&s.links.cut_cnt, num_meshes));
  YSIM_TRY(s.zero(s.links.cut_cnt, num_meshes));
  YSIM_TRY(s.alloc(&s.links.cut, (size_t)num_meshes * CUT_CAP));
  {
    // spanning-node records and partial boxes over the leaf positions of any
    // group (at most every facet of the scene)
    const size_t np = std::max<uint32_t>(s.Ttot, 1u);
    uint64_t* mask = nullptr;
    int4* rec = nullptr;
    float4 *part = nullptr, *blk = nullptr;
    YSIM_TRY(s.alloc(&mask, (np + 63) / 64));
    YSIM_TRY(s.alloc(&rec, np));
    YSIM_TRY(s.alloc(&part, 2 * np));
    YSIM_TRY(s.alloc(&blk, 2 * ((np + 255) / 256)));
    s.links.span_mask = mask;
    s.links.span_rec = rec;
    s.span = RefitSpan{mask, rec, part, blk};
  }
  YSIM_HIP_TRY(hipMemsetAsync(s.links.cut, 0xff, 4 * (size_t)num_meshes * CUT_CAP, s.st));  // -1: no entry
  YSIM_TRY(s.zero(s.tree, 8 * (size_t)s.Ntot));
  YSIM_TRY(s.zero(s.visit, s.Ntot));
  // segment tables
  YSIM_TRY(s.make_groups(s.dyn, s.dyn_groups));
  YSIM_TRY(s.make_groups(s.stat, s.stat_groups));
  YSIM_TRY(s.make_seg(s.dyn, true, s.dyn_tsegs));
  YSIM_TRY(s.make_seg(s.dyn, false, s.dyn_vsegs));
  s.dyn_voff.assign(1, 0);
  for (uint32_t m : s.dyn) {
    s.Vdyn += s.meshes[m].num_vertices;
    s.Tdyn += s.meshes[m].num_facets;
    s.dyn_voff.push_back(s.Vdyn);
  }
  s.shape_of.assign(num_meshes, -1);
  YSIM_TRY(s.build_targets());
  // sort buffers sized for the largest group
  size_t maxT = 0, maxStatus = 0, maxSeg = 1;
  for (auto* gs : {&s.dyn_groups, &s.stat_groups})
    for (auto& g : *gs) {
      maxT = std::max<size_t>(maxT, g.T);
      maxStatus = std::max(maxStatus, sort_status_words_max(g.T, g.plan.key_bits));
      maxSeg = std::max<size_t>(maxSeg, g.tsegs.nseg);
    }
  for (int b = 0; b < 2; ++b) {
    uint64_t* k = nullptr;
    YSIM_TRY(s.alloc(&k, maxT));
    s.keys[b] = k;
    YSIM_TRY(s.alloc(&s.vals[b], maxT));
  }
  YSIM_TRY(s.alloc(&s.hist, SORT_HIST_WORDS));
  YSIM_TRY(s.alloc(&s.status, maxStatus));
  YSIM_TRY(s.alloc(&s.counters, SORT_MAX_PASSES));
  YSIM_TRY(s.alloc(&s.epoch, 16));
  YSIM_TRY(s.alloc(&s.need_full, 16));
  YSIM_TRY(s.zero(s.need_full, 16));
  {
    uint32_t maxV = 0;
    for (uint32_t m = 0; m < num_meshes; ++m) maxV = std::max(maxV, s.meshes[m].num_vertices);
    YSIM_TRY(s.alloc(&s.box_parts, box_parts_words(maxV)));
    YSIM_TRY(s.zero(s.box_parts, box_parts_words(maxV)));
  }
  s.fuse_box = s.dyn_groups.size() == 1 && s.dyn_groups[0].meshes.size() == 1 && s.dyn_groups[0].V == s.Vdyn;
  if (s.fuse_box) YSIM_TRY(s.alloc(&s.cloth_parts, cloth_box_parts_words(s.Vdyn)));
  YSIM_HIP_TRY(hipHostMalloc((void**)&s.h_stale, 64, 0));
  *s.h_stale = 0u;
  YSIM_HIP_TRY(hipHostGetDevicePointer((void**)&s.d_stale, s.h_stale, 0));
  YSIM_TRY(s.alloc(&s.inc_bcnt, incremental_windows((uint32_t)maxT)));
  YSIM_TRY(s.zero(s.inc_bcnt, incremental_windows((uint32_t)maxT)));
  YSIM_TRY(s.alloc(&s.inc_edges, 3 * (size_t)incremental_windows((uint32_t)maxT)));
  // incremental sort: each dynamic group starts from the facet order
  for (auto& g : s.dyn_groups) {
    std::vector<uint32_t> id;
    id.reserve(g.T);
    for (uint32_t m : g.meshes)
      for (uint32_t f = 0; f < s.meshes[m].num_facets; ++f) id.push_back(f);
    YSIM_TRY(s.put(&g.perm, id));
  }
  YSIM_TRY(s.alloc(&s.box_acc, 6 * maxSeg));
  YSIM_TRY(s.zero(s.hist, SORT_HIST_WORDS));
  YSIM_TRY(s.zero(s.status, 2 * maxStatus));
  YSIM_TRY(s.zero(s.epoch, 16));
  {
    std::vector<uint32_t> init(6 * maxSeg);
    for (size_t i = 0; i < init.size(); ++i) init[i] = (i % 6) < 3 ? 0xffffffffu : 0u;
    YSIM_HIP_TRY(hipMemcpyAsync(s.box_acc, init.data(), init.size() * 4, hipMemcpyHostToDevice, s.st));
  }
  YSIM_TRY(s.alloc(&s.ccount, s.Vdyn));
  YSIM_TRY(s.alloc(&s.cnd, (size_t)s.Vdyn * s.cap));
  YSIM_TRY(s.alloc(&s.cid, (size_t)s.Vdyn * s.cap));
  YSIM_TRY(s.zero(s.ccount, s.Vdyn));
  YSIM_TRY(s.alloc(&s.npairs, 4));
  YSIM_TRY(s.alloc(&s.flags, 4));
  YSIM_TRY(s.zero(s.npairs, 4));
  // one statistics word per wave of any query launch (blocks of up to 256 threads)
  s.wstats_words = 4 * ceil_div(s.Vdyn, 256) + 4;
  YSIM_TRY(s.alloc(&s.wstats, s.wstats_words));
  YSIM_TRY(s.zero(s.wstats, s.wstats_words));
  YSIM_TRY(s.zero(s.flags, 4));
  for (auto& e : s.ev) YSIM_HIP_TRY(hipEventCreate(&e));
  if (P.broad_mode == 1) {
    s.hash_mode = true;
    std::vector<uint32_t> targets, dynf(num_meshes), grp(num_meshes);
    float max_margin = 0.f;
    for (uint32_t m = 0; m < num_meshes; ++m) {
      dynf[m] = s.dynamic(m) ? 1u : 0u;
      grp[m] = s.meshes[m].group;
      if (s.meshes[m].collidable && s.meshes[m].num_facets) targets.push_back(m);
      if (s.dynamic(m)) max_margin = std::max(max_margin, margin[m]);
    }
    HashArgs& h = s.hargs;
    YSIM_TRY(s.make_seg(targets, true, h.tsegs));
    h.mt = s.mt;
    h.dt = P.subh;
    h.swept = (int)P.swept;
    h.facets = s.facets;
    uint32_t *dd = nullptr, *dg = nullptr;
    YSIM_TRY(s.put(&dd, dynf));
    YSIM_TRY(s.put(&dg, grp));
    h.mesh_dynamic = dd;
    h.mesh_group = dg;
    for (uint32_t m : targets) h.ntri += s.meshes[m].num_facets;
    const float cell = P.hash_cell > 0.f ? P.hash_cell : 2.f * max_margin;
    YSIM_ARG(cell > 0.f, "scene_create: hash cell must be positive");
    // a query box spans at most 2 margin / cell + 2 cells per axis: keep that small
    YSIM_ARG(cell * 7.f >= max_margin, "scene_create: hash cell must be at least margin / 7");
    h.inv_h = 1.0f / cell;
    h.origin = f3{0.f, 0.f, 0.f};
    h.mask = (1u << P.hash_slots_log2) - 1u;
    h.self_collision = P.self_collision;
    h.flags = s.flags;
    // entry capacity from the initial geometry (cells touched per triangle box),
    // doubled; overflow raises flag bit 3
    uint64_t cells = 0;
    for (uint32_t m : targets)
      for (uint32_t t = 0; t < s.meshes[m].num_facets; ++t) {
        const uint32_t* f = h_facets + 3 * ((size_t)s.foff[m] + t);
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
          lo[k] = hi[k] = h_x[3 * ((size_t)s.voff[m] + f[0]) + k];
          for (int c = 1; c < 3; ++c) {
            float q = h_x[3 * ((size_t)s.voff[m] + f[c]) + k];
            lo[k] = std::min(lo[k], q);
            hi[k] = std::max(hi[k], q);
          }
        }
        uint64_t n = 1;
        for (int k = 0; k < 3; ++k)
          n *= (uint64_t)(std::floor(hi[k] * h.inv_h) - std::floor(lo[k] * h.inv_h) + 2);
        cells += n;
      }
    cells = std::min<uint64_t>(2 * cells + (1u << 20), (1ull << 31) - 1);
    h.entry_cap = (uint32_t)cells;
    YSIM_TRY(s.alloc(&h.entries, h.entry_cap));
    YSIM_TRY(s.alloc(&h.boxes, 6 * (size_t)h.ntri));
    YSIM_TRY(s.alloc(&h.ncells, h.ntri));
    YSIM_TRY(s.alloc(&h.count, (size_t)h.mask + 1));
    YSIM_TRY(s.alloc(&h.cursor, (size_t)h.mask + 1));
    YSIM_TRY(s.alloc(&h.offsets, (size_t)h.mask + 2));
    YSIM_TRY(s.alloc(&h.scan_tmp, 1025));
    YSIM_TRY(s.alloc(&h.big, h.ntri));
    YSIM_TRY(s.alloc(&h.big_cnt, 1));
    YSIM_TRY(s.zero(h.big_cnt, 1));
  }
  // initial build of every tree (unswept), Simulator::initialize -> broadPhase.build
  for (auto* gs : {&s.stat_groups, &s.dyn_groups})
    for (auto& g : *gs) {
      YSIM_TRY(s.enqueue_build(g, s.x[0]));
      // the group's own positions: the records its hierarchy kernel just wrote
      YSIM_TRY(launch_refit(s.st, g.tsegs, s.mt, s.x[0], s.v[0], 0.f, false, s.facets, s.tree, s.parent, s.visit,
                            g.T, s.flags, s.formal_refit, &s.links, s.tune.refit_block,
                            s.records_span(g) ? &s.span : nullptr));
    }
  YSIM_HIP_TRY(hipStreamSynchronize(s.st));
  uint32_t fl = 0;
  YSIM_HIP_TRY(hipMemcpy(&fl, s.flags, 4, hipMemcpyDeviceToHost));
  if (fl) return fail(YSIM_ERR_STATE, "scene_create: initial build raised flags " + std::to_string(fl));
  *out = S.release();
  return YSIM_OK;
}

int ysim_hip_scene_destroy(ysim_scene* scene) {
  if (!scene) return YSIM_OK;
  const hipError_t e = hipStreamSynchronize(scene->st);
  delete scene;
  if (e != hipSuccess) return fail(YSIM_ERR_HIP, std::string("scene_destroy: ") + hipGetErrorString(e));
  return YSIM_OK;
}

int ysim_hip_scene_step(ysim_scene* scene, uint32_t n, uint32_t use_graph) {
  YSIM_ARG(scene, "null scene");
  if (scene->stage_phase != 0) return fail(YSIM_ERR_STATE, "scene_step: a staged substep is in progress");
  scene->timing = false;
  return scene->steps(n, use_graph != 0);
}

int ysim_hip_scene_refit(ysim_scene* s) {
  YSIM_ARG(s, "null scene");
  if (s->stage_phase != 0) return fail(YSIM_ERR_STATE, "scene_refit: previous substep not integrated");
  s->timing = false;
  YSIM_TRY(s->enqueue_collision(s->parity, s->rebuild_at(s->step)));
  s->stage_phase = 1;
  return YSIM_OK;
}

int ysim_hip_scene_query(ysim_scene* s) {
  YSIM_ARG(s, "null scene");
  if (s->stage_phase != 1) return fail(YSIM_ERR_STATE, "scene_query: call scene_refit first");
  YSIM_TRY(s->enqueue_query(s->parity));
  s->stage_phase = 2;
  return YSIM_OK;
}

int ysim_hip_scene_integrate(ysim_scene* s) {
  YSIM_ARG(s, "null scene");
  if (s->stage_phase != 2) return fail(YSIM_ERR_STATE, "scene_integrate: call scene_query first");
  YSIM_TRY(s->enqueue_integrate(s->parity));
  s->parity ^= 1u;
  ++s->step;
  ++s->steps_run;
  s->stage_phase = 0;
  return YSIM_OK;
}

int ysim_hip_scene_get_state(ysim_scene* s, float* h_x, float* h_v, float* h_xprev) {
  YSIM_ARG(s, "null scene");
  const size_t b = 12 * (size_t)s->Vtot;
  if (h_x) YSIM_HIP_TRY(hipMemcpyAsync(h_x, s->x[s->parity], b, hipMemcpyDeviceToHost, s->st));
  if (h_v) YSIM_HIP_TRY(hipMemcpyAsync(h_v, s->v[s->parity], b, hipMemcpyDeviceToHost, s->st));
  if (h_xprev) YSIM_HIP_TRY(hipMemcpyAsync(h_xprev, s->x[s->parity ^ 1], b, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  return YSIM_OK;
}

int ysim_hip_scene_set_state(ysim_scene* s, const float* h_x, const float* h_v, const float* h_xprev) {
  YSIM_ARG(s, "null scene");
  const size_t b = 12 * (size_t)s->Vtot;
  if (h_x) YSIM_HIP_TRY(hipMemcpyAsync(s->x[s->parity], h_x, b, hipMemcpyHostToDevice, s->st));
  if (h_v) YSIM_HIP_TRY(hipMemcpyAsync(s->v[s->parity], h_v, b, hipMemcpyHostToDevice, s->st));
  if (h_xprev) YSIM_HIP_TRY(hipMemcpyAsync(s->x[s->parity ^ 1], h_xprev, b, hipMemcpyHostToDevice, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  s->full_next = true;
  s->box_fresh = false;
  return YSIM_OK;
}

int ysim_hip_scene_get_contacts(ysim_scene* s, uint32_t* h_offsets, ysim_narrow_collision* h_contacts,
                                uint32_t capacity, uint32_t* h_count) {
  YSIM_ARG(s && h_offsets && h_count, "scene_get_contacts: null argument");
  std::vector<uint32_t> cnt(s->Vdyn);
  std::vector<float4> nd((size_t)s->Vdyn * s->cap);
  std::vector<uint2> id((size_t)s->Vdyn * s->cap);
  YSIM_HIP_TRY(hipMemcpyAsync(cnt.data(), s->ccount, 4 * (size_t)s->Vdyn, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipMemcpyAsync(nd.data(), s->cnd, 16 * nd.size(), hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipMemcpyAsync(id.data(), s->cid, 8 * id.size(), hipMemcpyDeviceToHost, s->st));
  uint32_t fl = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&fl, s->flags, 4, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  if (fl & 2u) return fail(YSIM_ERR_CAPACITY, "scene_get_contacts: a vertex exceeded contact_capacity");
  std::fill(h_offsets, h_offsets + s->Vtot + 1, 0u);
  uint32_t total = 0;
  for (size_t k = 0; k < s->dyn.size(); ++k) {
    const uint32_t m = s->dyn[k];
    for (uint32_t p = 0; p < s->meshes[m].num_vertices; ++p) {
      const uint32_t j = s->dyn_voff[k] + p;
      const uint32_t g = s->voff[m] + p;
      h_offsets[g + 1] = std::min(cnt[j], s->cap);
    }
  }
  for (uint32_t g = 0; g < s->Vtot; ++g) h_offsets[g + 1] += h_offsets[g];
  total = h_offsets[s->Vtot];
  *h_count = total;
  if (!h_contacts) return YSIM_OK;
  if (total > capacity) return fail(YSIM_ERR_CAPACITY, "scene_get_contacts: capacity too small");
  for (size_t k = 0; k < s->dyn.size(); ++k) {
    const uint32_t m = s->dyn[k];
    for (uint32_t p = 0; p < s->meshes[m].num_vertices; ++p) {
      const uint32_t j = s->dyn_voff[k] + p;
      const uint32_t g = s->voff[m] + p;
      const uint32_t c = std::min(cnt[j], s->cap);
      std::vector<uint32_t> order(c);
      std::iota(order.begin(), order.end(), 0u);
      std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const uint2 A = id[(size_t)j * s->cap + a], B = id[(size_t)j * s->cap + b];
        return A.x != B.x ? A.x < B.x : A.y < B.y;
      });
      for (uint32_t r = 0; r < c; ++r) {
        const size_t src = (size_t)j * s->cap + order[r];
        ysim_narrow_collision& o = h_contacts[h_offsets[g] + r];
        o.point = p;
        o.triangle = id[src].y;
        o.query_obj = m;
        o.target_obj = id[src].x;
        o.normal_distance[0] = nd[src].x;
        o.normal_distance[1] = nd[src].y;
        o.normal_distance[2] = nd[src].z;
        o.normal_distance[3] = nd[src].w;
        o.query_behavior = s->meshes[m].behavior;
        o.target_behavior = s->meshes[id[src].x].behavior;
        o.query_shape = s->meshes[m].shape;
        o.target_shape = s->meshes[id[src].x].shape;
      }
    }
  }
  return YSIM_OK;
}

int ysim_hip_scene_get_pairs(ysim_scene* s, ysim_broad_collision* h_pairs, uint32_t capacity, uint32_t* h_count) {
  YSIM_ARG(s && h_count, "scene_get_pairs: null argument");
  s->timing = false;
  for (int attempt = 0; attempt < 2; ++attempt) {
    YSIM_TRY(s->enqueue_collision(s->parity, s->rebuild_at(s->step)));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s->st, s->npairs + 1, 1u);
    QueryArgs qa = s->query_args(s->parity);
    qa.npairs = s->npairs + 1;
    qa.pairs = s->pairs;
    qa.pair_cap = s->pair_cap;
    if (s->hash_mode) {
      HashArgs h = s->hargs;
      h.x = s->x[s->parity];
      YSIM_TRY(launch_hash_query(s->st, h, qa, true));
    } else {
      YSIM_TRY(launch_query(s->st, qa, true));
    }
    uint32_t n = 0;
    YSIM_HIP_TRY(hipMemcpyAsync(&n, s->npairs + 1, 4, hipMemcpyDeviceToHost, s->st));
    YSIM_HIP_TRY(hipStreamSynchronize(s->st));
    if (n <= s->pair_cap) {
      *h_count = n;
      if (!h_pairs) return YSIM_OK;
      if (n > capacity) return fail(YSIM_ERR_CAPACITY, "scene_get_pairs: capacity too small");
      std::vector<uint32_t> q(4 * (size_t)n);
      if (n) YSIM_HIP_TRY(hipMemcpy(q.data(), s->pairs, 16 * (size_t)n, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < n; ++i) {
        ysim_broad_collision& r = h_pairs[i];
        r.query_obj = q[4 * i];
        r.point = q[4 * i + 1];
        r.target_obj = q[4 * i + 2];
        r.triangle = q[4 * i + 3];
        r.query_behavior = s->meshes[r.query_obj].behavior;
        r.target_behavior = s->meshes[r.target_obj].behavior;
        r.query_shape = s->meshes[r.query_obj].shape;
        r.target_shape = s->meshes[r.target_obj].shape;
      }
      return YSIM_OK;
    }
    // grow the pair buffer and run again
    if (s->pairs) {
      YSIM_HIP_TRY(hipFree(s->pairs));
      s->allocs.erase(std::remove(s->allocs.begin(), s->allocs.end(), (void*)s->pairs), s->allocs.end());
    }
    s->pair_cap = n + n / 4 + 1024;
    YSIM_TRY(s->alloc(&s->pairs, 4 * (size_t)s->pair_cap));
  }
  return fail(YSIM_ERR_STATE, "scene_get_pairs: pair count unstable");
}

int ysim_hip_scene_get_tree(ysim_scene* s, uint32_t mesh, ysim_bvh_node* h_nodes, int32_t* h_parent,
                            uint64_t* h_sorted_keys, uint32_t* h_sorted_idx) {
  YSIM_ARG(s && mesh < s->nm, "scene_get_tree: bad mesh");
  const uint32_t T = s->meshes[mesh].num_facets;
  if (!T) return YSIM_OK;
  const size_t nn = 2 * (size_t)T - 1;
  if (h_nodes) YSIM_HIP_TRY(hipMemcpyAsync(h_nodes, s->tree + s->noff[mesh], nn * 32, hipMemcpyDeviceToHost, s->st));
  if (h_parent) YSIM_HIP_TRY(hipMemcpyAsync(h_parent, s->parent + s->noff[mesh], nn * 4, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  if (h_sorted_keys || h_sorted_idx) {
    // the sort buffers hold the build of the last dynamic group (groups are
    // built in order and share them)
    if (s->dyn_groups.empty()) return fail(YSIM_ERR_STATE, "scene_get_tree: no dynamic build");
    const BuildGroup& g = s->dyn_groups.back();
    auto it = std::find(g.meshes.begin(), g.meshes.end(), mesh);
    if (it == g.meshes.end()) return fail(YSIM_ERR_STATE, "scene_get_tree: keys of this mesh are not resident");
    uint32_t start = 0;
    for (auto m = g.meshes.begin(); m != it; ++m) start += s->meshes[*m].num_facets;
    const uint64_t mask = (g.code_bits >= 64) ? ~0ull : ((1ull << g.code_bits) - 1);
    if (h_sorted_keys) {
      if (g.key64) {
        YSIM_HIP_TRY(hipMemcpy(h_sorted_keys, (uint64_t*)s->keys[s->last_sorted] + start, 8 * (size_t)T,
                               hipMemcpyDeviceToHost));
      } else {
        std::vector<uint32_t> k(T);
        YSIM_HIP_TRY(hipMemcpy(k.data(), (uint32_t*)s->keys[s->last_sorted] + start, 4 * (size_t)T,
                               hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < T; ++i) h_sorted_keys[i] = k[i];
      }
      for (uint32_t i = 0; i < T; ++i) h_sorted_keys[i] &= mask;
    }
    if (h_sorted_idx)
      YSIM_HIP_TRY(hipMemcpy(h_sorted_idx, s->vals[s->last_sorted] + start, 4 * (size_t)T, hipMemcpyDeviceToHost));
  }
  return YSIM_OK;
}

int ysim_hip_scene_stage_timings(ysim_scene* s, uint32_t n_steps, float* h_ms) {
  YSIM_ARG(s && h_ms, "scene_stage_timings: null argument");
  for (double& d : s->stage_ms) d = 0.0;
  s->timing = true;
  int rc = s->eager_steps(n_steps);
  s->timing = false;
  if (rc != YSIM_OK) return rc;
  for (int k = 0; k < YSIM_STAGE_COUNT; ++k) h_ms[k] = n_steps ? (float)(s->stage_ms[k] / n_steps) : 0.f;
  return YSIM_OK;
}

int ysim_hip_scene_snapshot(ysim_scene* s) {
  YSIM_ARG(s, "null scene");
  const size_t b = 12 * (size_t)s->Vtot;
  if (!s->snap_x) {
    YSIM_TRY(s->alloc(&s->snap_x, 3 * (size_t)s->Vtot));
    YSIM_TRY(s->alloc(&s->snap_v, 3 * (size_t)s->Vtot));
    YSIM_TRY(s->alloc(&s->snap_xp, 3 * (size_t)s->Vtot));
  }
  YSIM_HIP_TRY(hipMemcpyAsync(s->snap_x, s->x[s->parity], b, hipMemcpyDeviceToDevice, s->st));
  YSIM_HIP_TRY(hipMemcpyAsync(s->snap_v, s->v[s->parity], b, hipMemcpyDeviceToDevice, s->st));
  YSIM_HIP_TRY(hipMemcpyAsync(s->snap_xp, s->x[s->parity ^ 1], b, hipMemcpyDeviceToDevice, s->st));
  s->snap_step = s->step;
  YSIM_TRY(s->save_amark(s->snap_amark));
  s->snap_valid = true;
  return YSIM_OK;
}

int ysim_hip_scene_restore(ysim_scene* s) {
  YSIM_ARG(s, "null scene");
  if (!s->snap_valid) return fail(YSIM_ERR_STATE, "scene_restore: no snapshot");
  const size_t b = 12 * (size_t)s->Vtot;
  // restore into the buffers of the snapshot's parity so graph phases line up
  const uint32_t par = (uint32_t)(s->snap_step & 1u);
  YSIM_HIP_TRY(hipMemcpyAsync(s->x[par], s->snap_x, b, hipMemcpyDeviceToDevice, s->st));
  YSIM_HIP_TRY(hipMemcpyAsync(s->v[par], s->snap_v, b, hipMemcpyDeviceToDevice, s->st));
  YSIM_HIP_TRY(hipMemcpyAsync(s->x[par ^ 1], s->snap_xp, b, hipMemcpyDeviceToDevice, s->st));
  YSIM_TRY(s->load_amark(s->snap_amark, par));
  s->parity = par;
  s->step = s->snap_step;
  s->full_next = true;
  s->box_fresh = false;
  return YSIM_OK;
}

int ysim_hip_scene_counters(ysim_scene* s, uint64_t* h_out, uint32_t reset) {
  YSIM_ARG(s && h_out, "scene_counters: null argument");
  std::vector<uint32_t> ws(s->wstats_words);
  if (s->wstats_words)
    YSIM_HIP_TRY(hipMemcpyAsync(ws.data(), s->wstats, 4 * (size_t)s->wstats_words, hipMemcpyDeviceToHost, s->st));
  std::vector<uint32_t> cnt(s->Vdyn);
  if (s->Vdyn) YSIM_HIP_TRY(hipMemcpyAsync(cnt.data(), s->ccount, 4 * (size_t)s->Vdyn, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  uint64_t c = 0;
  for (uint32_t k : cnt) c += std::min(k, s->cap);
  uint64_t np = 0;
  for (uint32_t k : ws) np += k;
  h_out[0] = s->pairs_hi + np;
  h_out[1] = s->steps_run - s->counter_step0;
  h_out[2] = c;
  if (reset) {
    YSIM_TRY(s->zero(s->wstats, s->wstats_words));
    s->pairs_hi = 0;
    s->counter_step0 = s->steps_run;
  }
  return YSIM_OK;
}

int ysim_hip_scene_deferred_builds(ysim_scene* s, uint64_t* h_out) {
  YSIM_ARG(s && h_out, "scene_deferred_builds: null argument");
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  if (*(volatile uint32_t*)s->h_stale) {
    *(volatile uint32_t*)s->h_stale = 0u;
    ++s->deferred_builds;
    s->full_next = true;
  }
  *h_out = s->deferred_builds;
  return YSIM_OK;
}

int ysim_hip_scene_set_audit(ysim_scene* s, uint32_t capacity) {
  YSIM_ARG(s, "null scene");
  if (s->stage_phase != 0) return fail(YSIM_ERR_STATE, "scene_set_audit: a staged substep is in progress");
  YSIM_TRY(s->release(s->audit));
  YSIM_TRY(s->release(s->audit_cnt));
  s->audit = nullptr;
  s->audit_cnt = nullptr;
  s->audit_cap = 0;
  s->drop_graph();
  if (!capacity) return YSIM_OK;
  YSIM_TRY(s->alloc(&s->audit, capacity));
  YSIM_TRY(s->alloc(&s->audit_cnt, 1));
  YSIM_TRY(s->zero(s->audit_cnt, 1));
  s->audit_cap = capacity;
  return YSIM_OK;
}

int ysim_hip_scene_get_audit(ysim_scene* s, ysim_broad_collision* h_pairs, uint32_t capacity, uint32_t* h_count,
                             uint32_t* h_redo) {
  YSIM_ARG(s && h_count, "scene_get_audit: null argument");
  if (!s->audit) return fail(YSIM_ERR_STATE, "scene_get_audit: the audit is off (scene_set_audit)");
  uint32_t n = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&n, s->audit_cnt, 4, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  if (n > s->audit_cap) {
    *h_count = n;
    return fail(YSIM_ERR_CAPACITY, "scene_get_audit: the audit buffer overflowed (records needed in *h_count)");
  }
  std::vector<uint32_t> rec(4 * (size_t)n);
  if (n) YSIM_HIP_TRY(hipMemcpy(rec.data(), s->audit, 16 * (size_t)n, hipMemcpyDeviceToHost));
  uint32_t nr = 0, nb = 0;
  const std::vector<AuditTriple> kept = audit_collect(rec.data(), n, &nr, &nb);
  if (nb) return fail(YSIM_ERR_STATE, "scene_get_audit: records of an unknown kind");
  if (h_redo) *h_redo = nr;
  *h_count = (uint32_t)kept.size();
  if (!h_pairs) return YSIM_OK;
  if (kept.size() > capacity) return fail(YSIM_ERR_CAPACITY, "scene_get_audit: capacity too small");
  // query item j -> (dynamic mesh, mesh-local vertex): dyn_voff holds the
  // query segments' starts
  for (size_t i = 0; i < kept.size(); ++i) {
    const AuditTriple& a = kept[i];
    const size_t k = (size_t)(std::upper_bound(s->dyn_voff.begin(), s->dyn_voff.end(), a.j) - s->dyn_voff.begin()) - 1;
    if (k >= s->dyn.size() || a.t >= s->nm) return fail(YSIM_ERR_STATE, "scene_get_audit: record out of range");
    ysim_broad_collision& r = h_pairs[i];
    r.query_obj = s->dyn[k];
    r.point = a.j - s->dyn_voff[k];
    r.target_obj = a.t;
    r.triangle = a.fid;
    r.query_behavior = s->meshes[r.query_obj].behavior;
    r.target_behavior = s->meshes[r.target_obj].behavior;
    r.query_shape = s->meshes[r.query_obj].shape;
    r.target_shape = s->meshes[r.target_obj].shape;
  }
  return YSIM_OK;
}

static_assert(YSIM_CUT_CAP == CUT_CAP, "the header's cut capacity is the build's");

int ysim_hip_scene_get_cut(ysim_scene* s, uint32_t mesh, int32_t* h_nodes, uint32_t capacity, uint32_t* h_count) {
  YSIM_ARG(s && h_count && mesh < s->nm, "scene_get_cut: bad argument");
  uint32_t n = 0;
  YSIM_HIP_TRY(hipMemcpyAsync(&n, s->links.cut_cnt + mesh, 4, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  if (n > CUT_CAP) n = 0;  // not listed: walks start from the root
  *h_count = n;
  if (!h_nodes || !n) return YSIM_OK;
  if (n > capacity) return fail(YSIM_ERR_CAPACITY, "scene_get_cut: capacity too small");
  YSIM_HIP_TRY(hipMemcpy(h_nodes, s->links.cut + (size_t)mesh * CUT_CAP, 4 * (size_t)n, hipMemcpyDeviceToHost));
  return YSIM_OK;
}

int ysim_hip_scene_set_analytic(ysim_scene* s, const ysim_analytic_shape* h_shapes, uint32_t n) {
  YSIM_ARG(s, "null scene");
  return s->set_analytic(h_shapes, n);
}

int ysim_hip_scene_set_collision(ysim_scene* s, float margin, float radius, float narrow_thickness,
                                 uint32_t self_collision) {
  YSIM_ARG(s, "null scene");
  return s->set_collision(margin, radius, narrow_thickness, self_collision);
}

int ysim_hip_scene_flags(ysim_scene* s, uint32_t* h_flags) {
  YSIM_ARG(s && h_flags, "scene_flags: null argument");
  YSIM_HIP_TRY(hipMemcpyAsync(h_flags, s->flags, 4, hipMemcpyDeviceToHost, s->st));
  YSIM_HIP_TRY(hipStreamSynchronize(s->st));
  return YSIM_OK;
}

}  // extern "C"

// ====================================================================== adjoint
namespace {
int adj_check(ysim_scene* s) {
  if (!s->ashapes.empty())
    return fail(YSIM_ERR_UNSUPPORTED, "adjoint: analytic colliders are not differentiated");
  if (s->dyn.size() != 1)
    return fail(YSIM_ERR_UNSUPPORTED, "adjoint: the scene must hold exactly one FastGrid cloth");
  if (s->P.self_collision)
    return fail(YSIM_ERR_UNSUPPORTED, "adjoint: contacts must be against static meshes (self_collision = 0)");
  if (s->cap > 64) return fail(YSIM_ERR_UNSUPPORTED, "adjoint: contact_capacity above 64");
  if (s->stage_phase != 0) return fail(YSIM_ERR_STATE, "adjoint: a staged substep is in progress");
  const ysim_mesh_desc& d = s->meshes[s->dyn[0]];
  if (d.behavior != YSIM_BEHAVIOR_FASTGRID_CLOTH)
    return fail(YSIM_ERR_UNSUPPORTED, "adjoint: the cloth must be a FastGrid cloth");
  if (d.cloth.particle_num_1d < 3 || (uint64_t)d.cloth.particle_num_1d * d.cloth.particle_num_1d != d.num_vertices)
    return fail(YSIM_ERR_INVALID, "adjoint: cloth is not an N x N grid with N >= 3");
  return YSIM_OK;
}
}  // namespace

extern "C" {

int ysim_hip_adj_num_springs(ysim_scene* s, uint32_t* h_out) {
  YSIM_ARG(s && h_out, "adj_num_springs: null argument");
  YSIM_TRY(adj_check(s));
  *h_out = SpringIdx(s->meshes[s->dyn[0]].cloth.particle_num_1d).total;
  return YSIM_OK;
}

int ysim_hip_adj_forward(ysim_scene* s, const float* d_log_k, float kd, uint32_t steps, uint32_t checkpoint_every) {
  YSIM_ARG(s && d_log_k, "adj_forward: null argument");
  YSIM_ARG(steps > 0 && checkpoint_every > 0, "adj_forward: steps and checkpoint_every must be positive");
  YSIM_TRY(adj_check(s));
  auto& A = s->adj;
  A.valid = false;
  A.mesh = s->dyn[0];
  A.V = s->meshes[A.mesh].num_vertices;
  A.voff = s->voff[A.mesh];
  A.S = SpringIdx(s->meshes[A.mesh].cloth.particle_num_1d).total;
  if (!s->kspring) YSIM_TRY(s->alloc(&s->kspring, A.S));
  YSIM_TRY(launch_adj_exp(s->st, d_log_k, s->kspring, A.S));
  s->P.kd = kd;
  if (s->graph && (s->graph_kd != kd || s->graph_ks != s->kspring)) s->drop_graph();
  A.steps = steps;
  A.K = checkpoint_every;
  A.nseg = (steps + A.K - 1) / A.K;
  const size_t blk = 9 * (size_t)A.V;
  YSIM_TRY(s->grow(&A.ck, A.ck_floats, blk * (A.nseg + 1)));
  A.step0 = s->step;
  for (uint32_t sg = 0; sg < A.nseg; ++sg) {
    YSIM_TRY(s->save_state(A.ck + blk * sg, s->parity));
    YSIM_TRY(s->steps(std::min(A.K, steps - sg * A.K), true));
  }
  YSIM_TRY(s->save_state(A.ck + blk * A.nseg, s->parity));
  A.par_end = s->parity;
  A.kd = kd;
  A.valid = true;
  return YSIM_OK;
}

int ysim_hip_adj_backward(ysim_scene* s, const float* d_x_bar, const float* d_v_bar, float* d_g_log_k, float* d_g_kd,
                          float* d_g_x0, float* d_g_v0) {
  YSIM_ARG(s && d_x_bar, "adj_backward: null argument");
  YSIM_TRY(adj_check(s));
  auto& A = s->adj;
  if (!A.valid) return fail(YSIM_ERR_STATE, "adj_backward: no adjoint forward to differentiate");
  if (s->P.kd != A.kd) return fail(YSIM_ERR_STATE, "adj_backward: damping changed since adj_forward");
  const size_t n3 = 3 * (size_t)A.V, blk = 9 * (size_t)A.V;
  size_t have = 0;
  if (!A.xb) {
    have = 0; YSIM_TRY(s->grow(&A.xb, have, n3));
    have = 0; YSIM_TRY(s->grow(&A.vb, have, n3));
    have = 0; YSIM_TRY(s->grow(&A.fb, have, n3));
    have = 0; YSIM_TRY(s->grow(&A.kdv, have, A.V));
    have = 0; YSIM_TRY(s->grow(&A.gk, have, A.S));
  }
  YSIM_TRY(s->grow(&A.tape, A.tape_floats, blk * A.K));
  YSIM_HIP_TRY(hipMemcpyAsync(A.xb, d_x_bar, 4 * n3, hipMemcpyDeviceToDevice, s->st));
  if (d_v_bar) YSIM_HIP_TRY(hipMemcpyAsync(A.vb, d_v_bar, 4 * n3, hipMemcpyDeviceToDevice, s->st));
  else YSIM_TRY(s->zero(A.vb, n3));
  YSIM_TRY(s->zero(A.gk, A.S));
  YSIM_TRY(s->zero(A.kdv, A.V));

  AdjArgs a;
  a.P = s->meshes[A.mesh].cloth;
  a.V = A.V;
  a.voff = A.voff;
  a.x = s->x[0];
  a.v = s->v[0];
  a.ext = s->ext;
  a.mass = s->mass;
  a.fixed = s->fixed;
  a.ks = s->kspring;
  a.subh = s->P.subh;
  a.kair = s->P.kair;
  a.kd = s->P.kd;
  a.ccount = s->ccount;
  a.cnd = s->cnd;
  a.cid = s->cid;
  a.cap = s->cap;
  a.xb = A.xb;
  a.vb = A.vb;
  a.fb = A.fb;
  a.gk = A.gk;
  a.kdv = A.kdv;
  s->timing = false;
  for (uint32_t sg = A.nseg; sg-- > 0;) {
    const uint32_t n0 = sg * A.K, n1 = std::min(A.steps, n0 + A.K);
    // recompute the segment from its checkpoint, taping each step's input state
    YSIM_TRY(s->load_state(A.ck + blk * sg, 0));
    s->parity = 0;
    s->step = A.step0 + n0;
    for (uint32_t n = n0; n < n1; ++n) {
      YSIM_TRY(s->save_state(A.tape + blk * (n - n0), s->parity));
      YSIM_TRY(s->eager_steps(1));
    }
    // reverse through the segment: contacts of step n at its input state, then
    // the two adjoint passes
    for (uint32_t n = n1; n-- > n0;) {
      YSIM_TRY(s->load_state(A.tape + blk * (n - n0), 0));
      s->parity = 0;
      s->step = A.step0 + n;
      YSIM_TRY(s->enqueue_collision(0, false));
      YSIM_TRY(s->enqueue_query(0));
      YSIM_TRY(launch_adj_step(s->st, a));
    }
  }
  YSIM_TRY(launch_adj_finish(s->st, A.gk, s->kspring, d_g_log_k, A.S, A.kdv, A.V, d_g_kd));
  if (d_g_x0) YSIM_HIP_TRY(hipMemcpyAsync(d_g_x0, A.xb, 4 * n3, hipMemcpyDeviceToDevice, s->st));
  if (d_g_v0) YSIM_HIP_TRY(hipMemcpyAsync(d_g_v0, A.vb, 4 * n3, hipMemcpyDeviceToDevice, s->st));
  // leave the scene where the forward left it
  YSIM_TRY(s->load_state(A.ck + blk * A.nseg, A.par_end));
  s->parity = A.par_end;
  s->step = A.step0 + A.steps;
  return YSIM_OK;
}

}  // extern "C"

// ====================================================================== tuning
extern "C" {

int ysim_hip_scene_get_tuning(ysim_scene* s, ysim_tuning* h_out) {
  YSIM_ARG(s && h_out, "scene_get_tuning: null argument");
  *h_out = s->tune;
  return YSIM_OK;
}

int ysim_hip_scene_set_tuning(ysim_scene* s, const ysim_tuning* h_in) {
  YSIM_ARG(s && h_in, "scene_set_tuning: null argument");
  if (s->stage_phase != 0) return fail(YSIM_ERR_STATE, "scene_set_tuning: a staged substep is in progress");
  YSIM_TRY(s->apply_tuning(*h_in));
  s->tune.ms_per_substep = 0.f;
  s->tune.ms_per_substep_default = 0.f;
  return YSIM_OK;
}

int ysim_hip_scene_autotune(ysim_scene* s, uint32_t trial_substeps, ysim_tuning* h_out) {
  YSIM_ARG(s, "scene_autotune: null scene");
  return s->autotune(trial_substeps, h_out);
}

}  // extern "C"

// ====================================================================== measurement
namespace ysim {
__global__ __launch_bounds__(256) void k_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace ysim

extern "C" {

int ysim_hip_copy_bandwidth(ysim_ctx* ctx, size_t bytes, uint32_t reps, float* h_ms) {
  YSIM_ARG(ctx && h_ms && bytes >= 16 && reps > 0, "copy_bandwidth: bad argument");
  const size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  YSIM_HIP_TRY(hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(YSIM_ERR_HIP, "copy_bandwidth: hipMalloc failed");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = YSIM_OK;
  float ms = 0.f;
  const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 8);
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(YSIM_ERR_HIP, "event");
  if (rc == YSIM_OK) {
    (void)hipMemsetAsync(a, 0, n * 16, ctx->stream);
    hipLaunchKernelGGL(k_copy_u4, dim3(grid), dim3(256), 0, ctx->stream, (const uint4*)a, (uint4*)b, n);
    (void)hipEventRecord(e0, ctx->stream);
    for (uint32_t r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_copy_u4, dim3(grid), dim3(256), 0, ctx->stream, (const uint4*)a, (uint4*)b, n);
    (void)hipEventRecord(e1, ctx->stream);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = fail(YSIM_ERR_HIP, "copy_bandwidth: timing failed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  YSIM_TRY(rc);
  *h_ms = ms / (float)reps;
  return YSIM_OK;
}

}  // extern "C"
