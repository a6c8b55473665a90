// hash.hip — spatial-hash broad phase with exactly the LBVH candidate set.
//
// The reference's hash broad phases (SpatialHashing<METAL>,
// spatialhashing.metal:22-478, six vertex-triangle rows per triangle pair, and
// CpuSpatialHash, spatial_hash.hpp:1547-2016) do not produce the LBVH's
// candidate set.  Here the hash answers the same predicate as the tree query
// (bvh.metal:1285-1412): vertex p and facet t are candidates iff the closed
// box [x_p - m, x_p + m] overlaps t's (swept) leaf box, facets containing p
// excluded on self queries.  Triangles are inserted into every grid cell their
// box touches (slot = hash(cell, group) mod 2^k, counting sort into a CSR);
// a query visits the cells of its box and reports a triangle only from the
// cell that holds the lower corner of the box intersection, so each pair is
// reported exactly once however many cells the two boxes share.  Each entry
// stores its cell key, so slot collisions never produce false matches.
#include "hash.hpp"

namespace ysim {

namespace {

__device__ __forceinline__ int cell_of(float v, float origin, float inv_h) {
  return (int)floorf((v - origin) * inv_h);
}
__device__ __forceinline__ uint64_t cell_key(int cx, int cy, int cz, uint32_t group) {
  // 20 bits per axis (offset), 4 bits of group: exact identity of a cell
  return ((uint64_t)(uint32_t)(cx + (1 << 19)) & 0xfffffull) |
         (((uint64_t)(uint32_t)(cy + (1 << 19)) & 0xfffffull) << 20) |
         (((uint64_t)(uint32_t)(cz + (1 << 19)) & 0xfffffull) << 40) | ((uint64_t)(group & 0xfu) << 60);
}
__device__ __forceinline__ uint32_t slot_of(int cx, int cy, int cz, uint32_t group, uint32_t mask) {
  uint32_t h = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u) ^
               (group * 2654435761u);
  return h & mask;
}

struct TriRef {
  uint32_t mesh, fid;
};
__device__ __forceinline__ TriRef tri_ref(const HashArgs& a, uint32_t gt) {
  // target triangles are numbered through a.tsegs (one segment per target mesh)
  const uint32_t sg = seg_of(a.tsegs, gt);
  return TriRef{a.tsegs.mesh[sg], gt - a.tsegs.off[sg]};
}
}  // namespace

// Leaf boxes of every target triangle (swept for dynamic meshes) and their
// cell ranges; counts the entries per slot.
__global__ __launch_bounds__(256) void k_hash_boxes_count(HashArgs a) {
  const uint32_t gt = blockIdx.x * 256 + threadIdx.x;
  if (gt >= a.ntri) return;
  const TriRef r = tri_ref(a, gt);
  const uint32_t vb = a.mt.voff[r.mesh];
  const bool swept = a.swept && a.mesh_dynamic[r.mesh];
  f3 lo, hi;
  tri_box(a.x, a.v, a.dt, swept, a.facets + 3 * (size_t)(a.mt.foff[r.mesh] + r.fid), vb, lo, hi);
  float* B = a.boxes + 6 * (size_t)gt;
  B[0] = lo.x; B[1] = lo.y; B[2] = lo.z; B[3] = hi.x; B[4] = hi.y; B[5] = hi.z;
  const uint32_t grp = a.mesh_group[r.mesh];
  const int x0 = cell_of(lo.x, a.origin.x, a.inv_h), x1 = cell_of(hi.x, a.origin.x, a.inv_h);
  const int y0 = cell_of(lo.y, a.origin.y, a.inv_h), y1 = cell_of(hi.y, a.origin.y, a.inv_h);
  const int z0 = cell_of(lo.z, a.origin.z, a.inv_h), z1 = cell_of(hi.z, a.origin.z, a.inv_h);
  const int64_t span = ((int64_t)x1 - x0 + 1) * ((int64_t)y1 - y0 + 1) * ((int64_t)z1 - z0 + 1);
  if (!(span >= 1 && span <= HASH_MAX_CELLS)) {  // direct list (ntri entries: never full)
    a.big[atomicAdd(a.big_cnt, 1u)] = gt;
    a.ncells[gt] = 0;
    return;
  }
  uint32_t n = 0;
  for (int cz = z0; cz <= z1; ++cz)
    for (int cy = y0; cy <= y1; ++cy)
      for (int cx = x0; cx <= x1; ++cx) {
        atomicAdd(a.count + slot_of(cx, cy, cz, grp, a.mask), 1u);
        ++n;
      }
  a.ncells[gt] = n;
}

// Fill: every (triangle, cell) entry at slot offset + running cursor.
__global__ __launch_bounds__(256) void k_hash_fill(HashArgs a) {
  const uint32_t gt = blockIdx.x * 256 + threadIdx.x;
  if (gt >= a.ntri) return;
  const TriRef r = tri_ref(a, gt);
  if (a.ncells[gt] == 0) return;  // on the direct list
  const float* B = a.boxes + 6 * (size_t)gt;
  const uint32_t grp = a.mesh_group[r.mesh];
  const int x0 = cell_of(B[0], a.origin.x, a.inv_h), x1 = cell_of(B[3], a.origin.x, a.inv_h);
  const int y0 = cell_of(B[1], a.origin.y, a.inv_h), y1 = cell_of(B[4], a.origin.y, a.inv_h);
  const int z0 = cell_of(B[2], a.origin.z, a.inv_h), z1 = cell_of(B[5], a.origin.z, a.inv_h);
  for (int cz = z0; cz <= z1; ++cz)
    for (int cy = y0; cy <= y1; ++cy)
      for (int cx = x0; cx <= x1; ++cx) {
        const uint32_t s = slot_of(cx, cy, cz, grp, a.mask);
        const uint32_t pos = a.offsets[s] + atomicAdd(a.cursor + s, 1u);
        if (pos < a.entry_cap) a.entries[pos] = HashEntry{cell_key(cx, cy, cz, grp), gt, 0u};
        else atomicOr(a.flags, 8u);
      }
}

int launch_hash_build(hipStream_t s, const HashArgs& a) {
  const uint32_t nslot = a.mask + 1;
  const uint32_t zb = std::min<uint32_t>(4096, ceil_div(nslot, 256));
  hipLaunchKernelGGL(k_zero_u32, dim3(zb), dim3(256), 0, s, a.count, nslot);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_zero_u32, dim3(zb), dim3(256), 0, s, a.cursor, nslot);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, a.big_cnt, 1u);
  YSIM_CHECK_LAUNCH();
  if (!a.ntri) return YSIM_OK;
  hipLaunchKernelGGL(k_hash_boxes_count, dim3(ceil_div(a.ntri, 256)), dim3(256), 0, s, a);
  YSIM_CHECK_LAUNCH();
  {
    int rc = launch_scan_large(s, a.count, nslot, a.offsets, a.scan_tmp);
    if (rc != YSIM_OK) return rc;
  }
  hipLaunchKernelGGL(k_hash_fill, dim3(ceil_div(a.ntri, 256)), dim3(256), 0, s, a);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// ------------------------------------------------------------------ large exclusive scan
// out[i] = sum(in[0..i)), out[n] = total; n <= 2^26 (1024 tiles of 64 Ki at most).
__global__ __launch_bounds__(256) void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t n,
                                                    uint32_t* __restrict__ out, uint32_t* __restrict__ tsum,
                                                    uint32_t per_thread) {
  __shared__ uint32_t s[256];
  const uint32_t base = blockIdx.x * 256 * per_thread + threadIdx.x * per_thread;
  uint32_t tot = 0;
  for (uint32_t k = 0; k < per_thread; ++k) {
    uint32_t i = base + k;
    tot += i < n ? in[i] : 0u;
  }
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    uint32_t t = threadIdx.x >= (uint32_t)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tot;
  for (uint32_t k = 0; k < per_thread; ++k) {
    uint32_t i = base + k;
    if (i < n) {
      uint32_t v = in[i];
      out[i] = run;
      run += v;
    }
  }
  if (threadIdx.x == 255) tsum[blockIdx.x] = s[255];
}
__global__ __launch_bounds__(1024) void k_scan_tile_sums(uint32_t* __restrict__ tsum, uint32_t nt) {
  __shared__ uint32_t s[1024];
  uint32_t v = threadIdx.x < nt ? tsum[threadIdx.x] : 0u;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t t = threadIdx.x >= (uint32_t)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (threadIdx.x < nt) tsum[threadIdx.x] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) tsum[1024] = s[1023];
}
__global__ void k_scan_add_tiles(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ tsum,
                                 uint32_t tile) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] += tsum[i / tile];
  if (i == 0) out[n] = tsum[1024];
}

int launch_scan_large(hipStream_t s, const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* tmp) {
  uint32_t per_thread = 1;
  while ((uint64_t)256 * per_thread * 1024 < n) per_thread *= 2;
  const uint32_t tile = 256 * per_thread;
  const uint32_t nt = std::max<uint32_t>(1, ceil_div(n, tile));
  if (nt > 1024) return fail(YSIM_ERR_UNSUPPORTED, "scan: too many elements");
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, s, in, n, out, tmp, per_thread);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, s, tmp, nt);
  YSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_add_tiles, dim3(std::max<uint32_t>(1, ceil_div(n, 256))), dim3(256), 0, s, out, n, tmp,
                     tile);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

// ------------------------------------------------------------------ query
template <bool EMIT_PAIRS>
__global__ __launch_bounds__(128) void k_hash_query(HashArgs h, QueryArgs a) {
  const uint32_t j = xcd_block(blockIdx.x, gridDim.x) * 128 + threadIdx.x;
  uint32_t npair = 0;
  if (j < a.nitems) {
    const uint32_t sg = seg_of(a.qsegs, j);
    const uint32_t m = a.qsegs.mesh[sg];
    const uint32_t p = j - a.qsegs.off[sg];
    const uint32_t g = a.mt.voff[m] + p;
    const f3 qcur = ld3(a.x, g);
    const f3 qprev = ld3(a.xprev, g);
    const float margin = a.mesh_margin[m];
    const float band = a.mesh_band[m];
    const f3 qmin = {qcur.x - margin, qcur.y - margin, qcur.z - margin};
    const f3 qmax = {qcur.x + margin, qcur.y + margin, qcur.z + margin};
    const uint32_t grp = h.mesh_group[m];
    const bool collidable = a.tgt_off[sg + 1] > a.tgt_off[sg];
    uint32_t cnt = 0;
    const uint32_t cap = a.cap;
    const int x0 = cell_of(qmin.x, h.origin.x, h.inv_h), x1 = cell_of(qmax.x, h.origin.x, h.inv_h);
    const int y0 = cell_of(qmin.y, h.origin.y, h.inv_h), y1 = cell_of(qmax.y, h.origin.y, h.inv_h);
    const int z0 = cell_of(qmin.z, h.origin.z, h.inv_h), z1 = cell_of(qmax.z, h.origin.z, h.inv_h);
    // a triangle whose box overlaps the query box: the group, self-collision
    // and incident-facet rules, then the narrow phase (or a pair record)
    auto candidate = [&](uint32_t gt) {
      const TriRef r = tri_ref(h, gt);
      if (h.mesh_group[r.mesh] != grp) return;  // the key keeps only 4 group bits
      const bool self = (r.mesh == m);
      if (self && !h.self_collision) return;
      const uint32_t* f = a.facets + 3 * (size_t)(a.mt.foff[r.mesh] + r.fid);
      const uint32_t f0 = f[0], f1 = f[1], f2 = f[2];
      if (self && (p == f0 || p == f1 || p == f2)) return;
      ++npair;
      if (a.audit) audit_lane(a, j, r.mesh, r.fid);
      if (EMIT_PAIRS) {
        uint32_t slot = atomicAdd(a.npairs, 1u);
        if (slot < a.pair_cap) reinterpret_cast<uint4*>(a.pairs)[slot] = make_uint4(m, p, r.mesh, r.fid);
        return;
      }
      const uint32_t tvb = a.mt.voff[r.mesh];
      f3 nrm;
      float d;
      if (narrow_test(qcur, qprev, ld3(a.x, tvb + f0), ld3(a.x, tvb + f1), ld3(a.x, tvb + f2), band, nrm, d)) {
        if (cnt < cap) {
          a.cnd[(size_t)j * cap + cnt] = make_float4(nrm.x, nrm.y, nrm.z, d);
          a.cid[(size_t)j * cap + cnt] = make_uint2(r.mesh, r.fid);
        }
        ++cnt;
      }
    };
    for (int cz = z0; collidable && cz <= z1; ++cz)
      for (int cy = y0; cy <= y1; ++cy)
        for (int cx = x0; cx <= x1; ++cx) {
          const uint64_t key = cell_key(cx, cy, cz, grp);
          const uint32_t sl = slot_of(cx, cy, cz, grp, h.mask);
          const uint32_t e1 = min(h.offsets[sl + 1], h.entry_cap);
          for (uint32_t e = h.offsets[sl]; e < e1; ++e) {
            const HashEntry E = h.entries[e];
            if (E.key != key) continue;
            const float* B = h.boxes + 6 * (size_t)E.tri;
            const f3 bmin = {B[0], B[1], B[2]}, bmax = {B[3], B[4], B[5]};
            if (!overlap(qmin, qmax, bmin, bmax)) continue;
            // report from the cell holding the lower corner of the intersection only
            const f3 imin = vmax(qmin, bmin);
            if (cell_of(imin.x, h.origin.x, h.inv_h) != cx || cell_of(imin.y, h.origin.y, h.inv_h) != cy ||
                cell_of(imin.z, h.origin.z, h.inv_h) != cz)
              continue;
            candidate(E.tri);
          }
        }
    // triangles too large for the table, tested directly (each reported once)
    const uint32_t nbig = collidable ? min(*h.big_cnt, h.ntri) : 0u;
    for (uint32_t k = 0; k < nbig; ++k) {
      const uint32_t gt = h.big[k];
      const float* B = h.boxes + 6 * (size_t)gt;
      if (overlap(qmin, qmax, f3{B[0], B[1], B[2]}, f3{B[3], B[4], B[5]})) candidate(gt);
    }
    if (!EMIT_PAIRS) {
      a.ccount[j] = cnt;
      if (cnt > cap) atomicOr(a.flags, 2u);
    }
  }
  if (!EMIT_PAIRS && a.wstats) {
    // per-wave running count, one word per wave (no same-address atomics)
    const uint32_t s = wave_sum_u32(npair);
    if ((threadIdx.x & 63) == 0 && s) a.wstats[j >> 6] += s;
  }
}

int launch_hash_query(hipStream_t s, const HashArgs& h, const QueryArgs& a, bool emit_pairs) {
  if (!a.nitems) return YSIM_OK;
  if (emit_pairs)
    hipLaunchKernelGGL(k_hash_query<true>, dim3(ceil_div(a.nitems, 128)), dim3(128), 0, s, h, a);
  else
    hipLaunchKernelGGL(k_hash_query<false>, dim3(ceil_div(a.nitems, 128)), dim3(128), 0, s, h, a);
  YSIM_CHECK_LAUNCH();
  return YSIM_OK;
}

}  // namespace ysim
