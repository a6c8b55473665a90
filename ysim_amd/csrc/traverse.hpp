// traverse.hpp — the BVH traversal shared by the fused query kernel and its
// host build (the CPU check of the kernel logic in tests/).
//
// Short-stack DFS that tests both children of a node at once: overlapping leaf
// children are reported immediately, overlapping internal children are
// descended (child_a first), and the right child is pushed only when both
// children are descended.  When the stack is full the walk continues without
// it: climbing through parent links, every ancestor reached from its child_a
// has its child_b tested, which is exactly the set of pending stack entries, so
// the visited set does not depend on the stack depth.
#pragma once
#include "common.hpp"

namespace ysim {

// Query order over an N x N grid cloth: the i-th query of the mesh handles
// vertex tile_order(i, N).  Full 8 x 8 tiles first (tile-major, a wave of 64
// lanes covers one tile, so its lanes walk nearby subtrees), then the columns
// right of the last full tile (row-major), then the rows below it (row-major).
// A bijection on [0, N*N).
YSIM_HD uint32_t tile_order(uint32_t i, uint32_t N) {
  const uint32_t nt = N / 8, Nf = nt * 8;
  if (i < Nf * Nf) {
    const uint32_t tile = i >> 6, w = i & 63u;
    const uint32_t r = (tile / nt) * 8 + (w >> 3), c = (tile % nt) * 8 + (w & 7u);
    return r * N + c;
  }
  uint32_t k = i - Nf * Nf;
  const uint32_t wb = N - Nf;
  if (k < Nf * wb) return (k / wb) * N + Nf + k % wb;
  k -= Nf * wb;
  return (Nf + k / N) * N + k % N;
}

struct NodeV {
  f3 lo, hi;
  int a, b;
};

// i >= 0 (a node of T's window); indexed unsigned, so the address needs no
// sign extension
YSIM_HD NodeV load_node(const ysim_bvh_node* T, int i) {
  const ysim_bvh_node& n = T[(uint32_t)i];
  return NodeV{{n.min[0], n.min[1], n.min[2]}, {n.max[0], n.max[1], n.max[2]}, n.child_a, n.child_b};
}

// stk[k * stride] is stack slot k: the child links (a, b) of a deferred node,
// so a pop needs no reload of the node.  on_leaf(facet) per overlapping leaf.
// Returns the number of loop iterations, or -1 if the guard (corrupt tree) fired.
template <int SDEPTH, class F>
YSIM_HD int traverse(const ysim_bvh_node* T, const int32_t* P, uint32_t nT, f3 qmin, f3 qmax, int2* stk,
                     int stride, F&& on_leaf) {
  NodeV r = load_node(T, 0);
  if (!overlap(qmin, qmax, r.lo, r.hi)) return 0;
  if (r.a < 0) {
    on_leaf(r.b);
    return 0;
  }
  int c = 0, ca = r.a, cb = r.b, sp = 0;
  bool stackless = false;
  const uint32_t guard_max = 4 * nT + 16;
  uint32_t guard = 0;
  while (true) {
    if (++guard > guard_max) return -1;
    NodeV A = load_node(T, ca), B = load_node(T, cb);
    const bool hA = overlap(qmin, qmax, A.lo, A.hi), hB = overlap(qmin, qmax, B.lo, B.hi);
    if (hA && A.a < 0) on_leaf(A.b);
    if (hB && B.a < 0) on_leaf(B.b);
    const bool goA = hA && A.a >= 0, goB = hB && B.a >= 0;
    if (goA) {
      if (goB && !stackless) {
        if (sp < SDEPTH) stk[(sp++) * stride] = make_int2(B.a, B.b);
        else stackless = true;
      }
      c = ca; ca = A.a; cb = A.b;
      continue;
    }
    if (goB) {
      c = cb; ca = B.a; cb = B.b;
      continue;
    }
    if (!stackless) {
      // The current node c is not needed after a pop: the walk only turns
      // stackless at a push, which is followed by a descent that sets c.
      if (sp == 0) return (int)guard;
      const int2 e = stk[(--sp) * stride];
      ca = e.x; cb = e.y;
      continue;
    }
    bool found = false;
    while (!found) {
      if (c == 0) return (int)guard;
      if (++guard > guard_max) return -1;
      const int p = P[c];
      if (p < 0 || p >= (int)nT - 1) return -1;
      NodeV PN = load_node(T, p);
      if (PN.a == c) {
        NodeV S = load_node(T, PN.b);
        if (S.a >= 0 && overlap(qmin, qmax, S.lo, S.hi)) {
          c = PN.b; ca = S.a; cb = S.b;
          found = true;
        } else {
          c = PN.b;  // leaf siblings were handled with their parent
        }
      } else {
        c = p;
      }
    }
  }
}

}  // namespace ysim
