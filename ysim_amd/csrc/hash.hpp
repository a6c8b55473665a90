// hash.hpp — spatial-hash broad phase (hash.hip).
#pragma once
#include "kernels.hpp"

namespace ysim {

struct HashEntry {
  uint64_t key;  // exact cell identity (cell coordinates + group)
  uint32_t tri;  // target triangle (numbering of HashArgs::tsegs)
  uint32_t pad;
};

struct HashArgs {
  SegTable tsegs;                      // target triangles, one segment per collidable mesh
  MeshTables mt;
  const float* x = nullptr;
  const float* v = nullptr;
  float dt = 0.f;
  int swept = 0;
  const uint32_t* facets = nullptr;
  const uint32_t* mesh_dynamic = nullptr;
  const uint32_t* mesh_group = nullptr;
  uint32_t ntri = 0;
  float* boxes = nullptr;              // 6 floats per target triangle
  uint32_t* ncells = nullptr;          // cells touched per triangle
  uint32_t* count = nullptr;           // per slot
  uint32_t* cursor = nullptr;          // per slot
  uint32_t* offsets = nullptr;         // per slot + 1
  uint32_t* scan_tmp = nullptr;        // 1025 words
  HashEntry* entries = nullptr;
  uint32_t entry_cap = 0;
  uint32_t* flags = nullptr;           // bit3: entry capacity exceeded
  uint32_t mask = 0;                   // slots - 1
  f3 origin = {0.f, 0.f, 0.f};
  float inv_h = 1.f;                   // 1 / cell edge
  uint32_t self_collision = 1;
  // triangles whose box touches more than HASH_MAX_CELLS cells are not put in
  // the table: they are listed here and every query tests them directly
  uint32_t* big = nullptr;             // ntri entries
  uint32_t* big_cnt = nullptr;         // [0] listed this build
};
// Cells a triangle may occupy in the table (a cloth thrown to the world bound
// spans ~10^7 cells per axis; such boxes take the direct list instead).
constexpr int64_t HASH_MAX_CELLS = 4096;

int launch_hash_build(hipStream_t s, const HashArgs& a);
int launch_hash_query(hipStream_t s, const HashArgs& h, const QueryArgs& a, bool emit_pairs);
// out[i] = sum(in[0..i)), out[n] = total; tmp holds 1025 words.
int launch_scan_large(hipStream_t s, const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* tmp);

}  // namespace ysim
