"""Pins the tree-cut rule the tile query starts its walk from: the children
of at most S = ceil(leaves / 128) leaves under a parent of more form a
frontier that covers every leaf exactly once, so a walk over boxes that
starts from the cut finds exactly the leaves a walk from the root finds
(checked on the oracle's Karras trees of a fold and of a sphere, for random
query boxes).  The GPU scenes that start their walks from the cut are held
to the oracle's contacts by the GPU suite."""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import scenes

from test_oracle_agglo import _ranges


def _cut(tree, n):
    S = max(1, (n + 127) // 128)
    r = _ranges(tree, n)
    size = {nd: hi - lo + 1 for nd, (lo, hi) in r.items()}
    cut = []
    for p in range(n - 1):
        if size[p] > S:
            for c in (int(tree[p]["child_a"]), int(tree[p]["child_b"])):
                if size[c] <= S:
                    cut.append(c)
    return cut, r, S


def _walk(tree, starts, lo, hi):
    found, stack = [], list(starts)
    while stack:
        nd = stack.pop()
        mn, mx = tree[nd]["min"], tree[nd]["max"]
        if np.any(mx < lo) or np.any(mn > hi):
            continue
        a = int(tree[nd]["child_a"])
        if a < 0:
            found.append(int(tree[nd]["child_b"]))
        else:
            stack += [a, int(tree[nd]["child_b"])]
    return sorted(found)


@pytest.mark.parametrize("which", ["fold", "sphere"])
def test_cut_is_a_frontier_and_walks_agree(which):
    if which == "fold":
        sc = scenes.config_c2(n=64)
        mesh = 0
    else:
        sc = scenes.config_drape(n=16)
        mesh = 1
    S_ = O.OracleScene(sc)
    S_.step(2)
    tree, _, _, _ = S_.tree(mesh)
    n = sc.meshes[mesh].facets.shape[0]
    cut, r, S = _cut(tree, n)
    assert 0 < len(cut) <= 512
    cover = np.zeros(n, np.int32)
    for c in cut:
        lo, hi = r[c]
        assert hi - lo + 1 <= S
        cover[lo:hi + 1] += 1
    assert np.all(cover == 1)
    rng = np.random.default_rng(5)
    root_lo, root_hi = tree[0]["min"], tree[0]["max"]
    for _ in range(40):
        c = rng.uniform(root_lo, root_hi).astype(np.float32)
        h = (rng.uniform(0.0, 0.15) * (root_hi - root_lo)).astype(np.float32)
        lo, hi = c - h, c + h
        assert _walk(tree, cut, lo, hi) == _walk(tree, [0], lo, hi)


def test_planted_wrong_threshold_breaks_the_cover():
    """A cut that also lists the children of nodes at the threshold (a >=
    instead of >) covers some leaves twice: the cover check above would fail."""
    sc = scenes.config_c2(n=32)
    S_ = O.OracleScene(sc)
    tree, _, _, _ = S_.tree(0)
    n = sc.meshes[0].facets.shape[0]
    S = max(1, (n + 127) // 128)
    r = _ranges(tree, n)
    size = {nd: hi - lo + 1 for nd, (lo, hi) in r.items()}
    wrong = [c for p in range(n - 1) if size[p] >= S
             for c in (int(tree[p]["child_a"]), int(tree[p]["child_b"])) if size[c] <= S]
    cover = np.zeros(n, np.int32)
    for c in wrong:
        lo, hi = r[c]
        cover[lo:hi + 1] += 1
    assert np.any(cover > 1)
