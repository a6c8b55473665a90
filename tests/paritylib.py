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


def heavy_tile_leaves(x, v, facets, n: int, margin: float, dt: float, heavy: int = 16) -> int:
    """(tile, leaf) pairs of an n x n grid cloth's self query that the tile
    walk tests with the whole wave: the leaf's box (the facet's corners at x
    and x + v dt) overlaps more than `heavy` crossings of the 8 x 8 tile's row
    and column boxes (each the union of its 8 vertices' [x - m, x + m]).
    Counted on the CPU from a state, so a GPU test can show that the case it
    audits takes that path."""
    x = np.asarray(x, np.float32).reshape(n, n, 3)
    s = (np.asarray(x, np.float32).reshape(-1, 3) + np.asarray(v, np.float32).reshape(-1, 3) * np.float32(dt))
    f = np.asarray(facets).reshape(-1, 3)
    X = x.reshape(-1, 3)
    llo = np.minimum(X[f].min(1), s[f].min(1))
    lhi = np.maximum(X[f].max(1), s[f].max(1))
    # a leaf crossing more than 16 of a tile's 64 row x column crossings
    # spans at least 5 rows or 5 columns, so it is wider than 1.5 grid steps
    ext = (lhi - llo).max(1)
    step = float(np.median(np.linalg.norm(x[:, 1:] - x[:, :-1], axis=-1)))
    big = np.nonzero(ext > 1.5 * step)[0]
    if big.size == 0:
        return 0
    t = x.reshape(n // 8, 8, n // 8, 8, 3)
    m = np.float32(margin)
    # tile rows: (tile row, row, tile column, 3) -> (tiles, 8, 3)
    rlo = (t - m).min(axis=3).transpose(0, 2, 1, 3).reshape(-1, 8, 3)
    rhi = (t + m).max(axis=3).transpose(0, 2, 1, 3).reshape(-1, 8, 3)
    clo = (t - m).min(axis=1).reshape(-1, 8, 3)  # tile columns
    chi = (t + m).max(axis=1).reshape(-1, 8, 3)
    count = 0
    for i in big:
        ro = np.all((rlo <= lhi[i]) & (rhi >= llo[i]), axis=-1).sum(1)
        co = np.all((clo <= lhi[i]) & (chi >= llo[i]), axis=-1).sum(1)
        count += int(np.count_nonzero(ro * co > heavy))
    return count
