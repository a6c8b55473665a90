"""The segmented broad query (queryPointsSegmented's per-workgroup emission):
the same pair set and count as the per-pair query and as brute force, also
when a workgroup's pairs overflow its LDS staging, and at 512 x 512 size."""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import hip, scenes

pytestmark = pytest.mark.gpu


def _tree(t, v, swept):
    k, i = O.morton(t.x, t.facets, O.scene_box(t.x), 30)
    k, i = O.sort_pairs(k, i, 30)
    tree, _ = O.karras(k, i, 30)
    return O.refit(tree, t.x, v, np.float32(1 / 600), swept, t.facets)


def _pairs(rec):
    p = np.stack([rec["point"], rec["triangle"]], 1)
    s = set(map(tuple, p.tolist()))
    assert len(s) == len(p), "duplicate pairs"
    return s


@pytest.mark.parametrize("margin_scale", [1.0, 12.0])
def test_segmented_equals_brute_force(ctx, margin_scale):
    """margin x 12 gives ~60 pairs per vertex: every workgroup overflows its
    512-record LDS staging and spills through the per-pair path."""
    sc = scenes.config_drape(n=32, height=0.79)
    cl, sp = sc.meshes
    rng = np.random.default_rng(3)
    v = (rng.standard_normal(cl.x.shape) * 0.05).astype(np.float32)
    margin = sc.params["margin"] * margin_scale
    for t, tobj in ((cl, 0), (sp, 1)):
        swept = 1 if tobj == 0 else 0
        tree = _tree(t, v if swept else np.zeros_like(t.x), swept)
        seg = hip.broad_query(ctx, cl.x, 0, tree, t.facets, tobj, margin, segmented=True)
        plain = hip.broad_query(ctx, cl.x, 0, tree, t.facets, tobj, margin)
        T = t.facets.shape[0]
        leaf = tree[T - 1:]
        lb = np.zeros((T, 6), np.float32)
        lb[leaf["child_b"], 0:3] = leaf["min"]
        lb[leaf["child_b"], 3:6] = leaf["max"]
        brute = set(map(tuple, O.brute_candidates(cl.x, 0, lb, t.facets, tobj, margin).tolist()))
        assert _pairs(seg) == brute == _pairs(plain)
        assert np.all(seg["query_obj"] == 0) and np.all(seg["target_obj"] == tobj)
    if margin_scale > 1:
        assert len(seg) > 128 * 512 // 4


def test_segmented_full_size(ctx):
    sc = scenes.config_c2()
    m = sc.meshes[0]
    tree = _tree(m, np.zeros_like(m.x), 0)
    seg = hip.broad_query(ctx, m.x, 0, tree, m.facets, 0, sc.params["margin"], segmented=True)
    plain = hip.broad_query(ctx, m.x, 0, tree, m.facets, 0, sc.params["margin"])
    assert len(seg) == len(plain) > 100000 and _pairs(seg) == _pairs(plain)
