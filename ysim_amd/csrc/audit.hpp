// audit.hpp — candidate-pair audit records of the query kernels, and the host
// filter that turns a substep's records into its candidate set.
//
// While a scene's audit is on, every query kernel (the tile walk, the
// per-vertex walk, the hash query) appends one 16-byte record per candidate
// it reports, {query item j, target mesh t, facet, tag}, where
// tag = kind << 30 | wave.  Kinds:
//   AUDIT_TILE     reported by a tile walk (k_query_tile) of wave `wave`;
//   AUDIT_LANE     reported by a per-lane walk (k_query, the tile kernel's
//                  mixed-segment waves and its redo, the hash query): final;
//   AUDIT_RETRACT  wave `wave` abandoned its tile walk of target t (pending
//                  list overflow) and redid t per lane: its AUDIT_TILE records
//                  of t are void (the redo's AUDIT_LANE records replace them).
// A tile wave walks each target once, so "every AUDIT_TILE record of (wave, t)
// with a matching retract" is exactly the abandoned walk's, whatever order the
// records landed in.  The candidate set of the substep is every remaining
// record; a walk that reports a candidate twice leaves two records, so callers
// can check the multiset, not only the set.
//
// The tag helper is constexpr, so device code uses it too; the filter is host
// code, which the library and the CPU host check build both compile.
#pragma once
#include <algorithm>
#include <cstdint>
#include <unordered_set>
#include <vector>

namespace ysim {

constexpr uint32_t AUDIT_TILE = 0u, AUDIT_LANE = 1u, AUDIT_RETRACT = 2u;
constexpr uint32_t AUDIT_WAVE_MASK = (1u << 30) - 1u;

constexpr uint32_t audit_tag(uint32_t kind, uint32_t wave) { return kind << 30 | (wave & AUDIT_WAVE_MASK); }

struct AuditTriple {
  uint32_t j, t, fid;
};

// rec: n records of 4 words.  Returns the kept (j, t, fid) triples sorted by
// (j, t, fid); *n_retract = number of retract records (tile walks redone per
// lane); *n_bad = records of an unknown kind (none in a valid run).
inline std::vector<AuditTriple> audit_collect(const uint32_t* rec, size_t n, uint32_t* n_retract, uint32_t* n_bad) {
  std::unordered_set<uint64_t> void_walks;  // (wave, target) of retracted tile walks
  uint32_t nr = 0, nb = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t kind = rec[4 * i + 3] >> 30;
    if (kind == AUDIT_RETRACT) {
      void_walks.insert((uint64_t)(rec[4 * i + 3] & AUDIT_WAVE_MASK) << 32 | rec[4 * i + 1]);
      ++nr;
    } else if (kind != AUDIT_TILE && kind != AUDIT_LANE) {
      ++nb;
    }
  }
  std::vector<AuditTriple> out;
  out.reserve(n);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t* r = rec + 4 * i;
    const uint32_t kind = r[3] >> 30;
    if (kind == AUDIT_LANE ||
        (kind == AUDIT_TILE && !void_walks.count((uint64_t)(r[3] & AUDIT_WAVE_MASK) << 32 | r[1])))
      out.push_back({r[0], r[1], r[2]});
  }
  std::sort(out.begin(), out.end(), [](const AuditTriple& a, const AuditTriple& b) {
    return a.j != b.j ? a.j < b.j : a.t != b.t ? a.t < b.t : a.fid < b.fid;
  });
  if (n_retract) *n_retract = nr;
  if (n_bad) *n_bad = nb;
  return out;
}

}  // namespace ysim
