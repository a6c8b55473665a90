"""Pins the decomposition the two-kernel refit finishes its spanning nodes
with (k_refit_lds<..., true> + k_refit_span): with the sorted leaves cut into
blocks of RB, the maximal nodes inside a block whose parent spans blocks
("escapes") partition that block's leaves, and the box of every node that
spans blocks equals

    U(escapes of its first block from its range start on)
  U U(whole blocks strictly between its first and last block)
  U U(escapes of its last block up to its range end),

the first union written by the child_a escape that starts the node's range,
the last by the child_b escape that ends it.  Checked bit for bit against the
oracle's bottom-up boxes (bvh.metal:510-586 order) on the Karras trees of a
fold and of a sphere, for every block size the refit takes; a planted slip
(a first-block union that stops one escape short) is caught."""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import scenes

from test_oracle_agglo import _ranges


def _decompose(tree, n, RB, slip=False):
    r = _ranges(tree, n)
    parent = {}
    for p in range(n - 1):
        for c in (int(tree[p]["child_a"]), int(tree[p]["child_b"])):
            parent[c] = (p, 0 if c == int(tree[p]["child_a"]) else 1)
    span = {nd for nd, (lo, hi) in r.items() if nd < n - 1 and lo // RB != hi // RB}
    esc = [nd for nd in r if nd != 0 and nd not in span and parent[nd][0] in span]
    mn = lambda nd: tree[nd]["min"].astype(np.float32)
    mx = lambda nd: tree[nd]["max"].astype(np.float32)
    part, blk = {}, {}
    by_block = {}
    for e in esc:
        by_block.setdefault(r[e][0] // RB, []).append(e)
    for b, E in by_block.items():
        E = sorted(E, key=lambda e: r[e][0])
        # the escapes partition the block's leaves
        assert r[E[0]][0] == b * RB and r[E[-1]][1] == min(n, (b + 1) * RB) - 1
        assert all(r[E[k]][1] + 1 == r[E[k + 1]][0] for k in range(len(E) - 1))
        blk[b] = (np.min([mn(e) for e in E], 0), np.max([mx(e) for e in E], 0))
        for k, e in enumerate(E):
            lo, hi = r[e]
            sel = E[k:] if parent[e][1] == 0 else E[:k + 1]
            if slip and parent[e][1] == 0 and len(sel) > 1:
                sel = sel[:-1]
            pos = lo if parent[e][1] == 0 else hi
            assert pos not in part  # one writer per position
            part[pos] = (np.min([mn(x) for x in sel], 0), np.max([mx(x) for x in sel], 0))
    bad = 0
    for s in span:
        lo, hi = r[s]
        a, b = lo // RB, hi // RB
        pieces = [part[lo], part[hi]] + [blk[c] for c in range(a + 1, b)]
        l = np.min([p[0] for p in pieces], 0)
        h = np.max([p[1] for p in pieces], 0)
        if not (np.array_equal(l.view(np.uint32), mn(s).view(np.uint32))
                and np.array_equal(h.view(np.uint32), mx(s).view(np.uint32))):
            bad += 1
    return len(span), len(esc), bad


def _tree(which):
    if which == "fold":
        sc, mesh = scenes.config_c2(n=96), 0
    else:
        sc, mesh = scenes.config_drape(n=24), 1
    S = O.OracleScene(sc)
    S.step(3)
    tree, _, _, _ = S.tree(mesh)
    return tree, sc.meshes[mesh].facets.shape[0]


@pytest.mark.parametrize("which", ["fold", "sphere"])
@pytest.mark.parametrize("RB", [256, 512, 1024])
def test_spanning_boxes_from_block_partials(which, RB):
    tree, n = _tree(which)
    nspan, nesc, bad = _decompose(tree, n, RB)
    if n > RB:
        assert nspan > 0 and nesc == nspan + 1  # a binary tree over the escapes
    assert bad == 0


def test_planted_partial_slip_is_caught():
    tree, n = _tree("fold")
    assert _decompose(tree, n, 256, slip=True)[2] > 0
