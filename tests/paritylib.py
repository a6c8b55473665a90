"""Checkers shared by the parity tests (CPU and GPU).

Candidate sets are compared as multisets of (query_obj, point, target_obj,
triangle) rows: the oracle reports each candidate once, so a kernel that drops
one, reports one twice or invents one differs from it in at least one row.
"""
from __future__ import annotations

import numpy as np


def rows(a) -> np.ndarray:
    """(n, 4) uint32 rows sorted lexicographically (duplicates kept)."""
    a = np.ascontiguousarray(np.asarray(a, np.uint32).reshape(-1, 4))
    if a.shape[0] == 0:
        return a
    order = np.lexsort((a[:, 3], a[:, 2], a[:, 1], a[:, 0]))
    return a[order]


def candidate_diff(got, want) -> dict:
    """Multiset difference of two candidate lists: rows missing from `got`,
    rows extra in `got`, and rows `got` repeats."""
    g, w = rows(got), rows(want)

    def counts(r):
        if r.shape[0] == 0:
            return {}
        u, c = np.unique(r, axis=0, return_counts=True)
        return {tuple(map(int, k)): int(n) for k, n in zip(u, c)}

    cg, cw = counts(g), counts(w)
    missing = [k for k, n in cw.items() if cg.get(k, 0) < n]
    extra = [k for k, n in cg.items() if cw.get(k, 0) < n]
    dup = [k for k, n in cg.items() if n > 1]
    return {"missing": missing, "extra": extra, "duplicated": dup, "n_got": int(g.shape[0]),
            "n_want": int(w.shape[0])}


def assert_same_candidates(got, want, what: str = "candidate set"):
    g, w = rows(got), rows(want)
    if g.shape == w.shape and np.array_equal(g, w):
        return
    d = candidate_diff(g, w)
    raise AssertionError(
        f"{what}: {d['n_got']} rows vs {d['n_want']} expected; {len(d['missing'])} missing "
        f"(first {d['missing'][:4]}), {len(d['extra'])} extra (first {d['extra'][:4]}), "
        f"{len(d['duplicated'])} duplicated (first {d['duplicated'][:4]})")


def cut_cover(nodes, cut) -> np.ndarray:
    """How many cut nodes cover each leaf slot of a tree (leaves at T-1+i):
    a valid cut covers every leaf exactly once."""
    n = nodes.shape[0]
    T = (n + 1) // 2
    cover = np.zeros(T, np.int64)
    ca = nodes["child_a"]
    cb = nodes["child_b"]
    for c in np.asarray(cut, np.int64):
        stack = [int(c)]
        steps = 0
        while stack:
            k = stack.pop()
            steps += 1
            if steps > 2 * n or k < 0 or k >= n:
                raise AssertionError(f"cut node {c}: walk leaves the tree")
            if ca[k] < 0:
                cover[k - (T - 1)] += 1
            else:
                stack.append(int(ca[k]))
                stack.append(int(cb[k]))
    return cover
