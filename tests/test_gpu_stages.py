"""Stage-level parity of the HIP kernels against the CPU oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import abi, hip, scenes

pytestmark = pytest.mark.gpu


def cloth_and_sphere(seed=0):
    sc = scenes.config_c1(seed=seed)
    return sc.meshes[0], sc.meshes[1]


@pytest.mark.parametrize("n", [1, 7, 1000, 262144])
def test_scene_box(ctx, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((n, 3)) * 3).astype(np.float32)
    x[0, 1] = -0.0
    got = hip.scene_box(ctx, x)
    exp = O.scene_box(x)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_matches_oracle(ctx, bits):
    cl, sp = cloth_and_sphere()
    for m in (cl, sp):
        box = O.scene_box(m.x)
        kg, ig = hip.morton(ctx, m.x, m.facets, box, bits)
        ko, io = O.morton(m.x, m.facets, box, bits)
        assert np.array_equal(kg, ko)
        assert np.array_equal(ig, io)


@pytest.mark.parametrize("n,bits", [(1, 30), (2, 30), (5000, 30), (2049, 63), (100003, 38), (522242, 63)])
def test_sort_matches_oracle(ctx, n, bits):
    rng = np.random.default_rng(n + bits)
    hi = np.uint64((1 << bits) - 1)
    keys = rng.integers(0, 1 << min(bits, 62), n, dtype=np.uint64) & hi
    keys[: n // 3] = keys[0]  # many ties: stability decides the permutation
    vals = np.arange(n, dtype=np.uint32)
    kg, vg = hip.RadixSorter(ctx).sort(keys, vals, bits)
    ko, vo = O.sort_pairs(keys, vals, bits)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(vo, perm)
    assert np.array_equal(vg, vo)
    assert np.array_equal(kg, ko)


SORT_CFGS = [(8, 16, False), (11, 8, False), (11, 16, False), (8, 8, True), (8, 16, True)]


@pytest.mark.parametrize("cfg", SORT_CFGS)
@pytest.mark.parametrize("n,bits", [(1, 63), (100, 7), (5000, 11), (5000, 12), (70000, 19), (70000, 30),
                                    (300000, 63), (522242, 63)])
def test_sort_configs_match_argsort(ctx, cfg, n, bits):
    """Every sort configuration, including the hybrid MSD + per-bucket sort with
    buckets far above LDS capacity (a third of the keys share one key), gives
    numpy's stable permutation."""
    db, items, hybrid = cfg
    print(f"sort cfg={cfg} n={n} bits={bits}", flush=True)
    rng = np.random.default_rng(7 * n + bits)
    hi = np.uint64((1 << bits) - 1)
    keys = rng.integers(0, 1 << min(bits, 62), n, dtype=np.uint64) & hi
    if bits == 63:
        keys |= rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(62)
    keys[: n // 3] = keys[0]
    keys[n // 3: n // 2] = rng.integers(0, 64, n // 2 - n // 3, dtype=np.uint64) & hi  # one crowded low bucket
    vals = np.arange(n, dtype=np.uint32)
    kg, vg = hip.RadixSorter(ctx).sort(keys, vals, bits, digit_bits=db, items=items, hybrid=hybrid)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(vg, perm)
    assert np.array_equal(kg, keys[perm])


@pytest.mark.parametrize("bits", [30, 63])
def test_hierarchy_matches_oracle(ctx, bits):
    cl, _ = cloth_and_sphere(1)
    box = O.scene_box(cl.x)
    k, i = O.morton(cl.x, cl.facets, box, bits)
    k, i = O.sort_pairs(k, i, bits)
    tg, pg = hip.lbvh_hierarchy(ctx, k, i, bits)
    to, po = O.karras(k, i, bits)
    assert np.array_equal(tg["child_a"], to["child_a"])
    assert np.array_equal(tg["child_b"], to["child_b"])
    assert np.array_equal(pg, po)


def test_hierarchy_duplicate_codes(ctx):
    n = 4097
    k = np.repeat(np.arange(0, 64, dtype=np.uint64), n // 64 + 1)[:n]
    i = np.arange(n, dtype=np.uint32)
    tg, pg = hip.lbvh_hierarchy(ctx, k, i, 30)
    to, po = O.karras(k, i, 30)
    assert np.array_equal(tg["child_a"], to["child_a"]) and np.array_equal(tg["child_b"], to["child_b"])
    assert np.array_equal(pg, po)


@pytest.mark.parametrize("swept", [0, 1])
@pytest.mark.parametrize("formal", [False, True])
def test_refit_matches_oracle(ctx, ctx_formal, swept, formal):
    c = ctx_formal if formal else ctx
    cl, _ = cloth_and_sphere(2)
    rng = np.random.default_rng(5)
    v = (rng.standard_normal(cl.x.shape) * 0.1).astype(np.float32)
    box = O.scene_box(cl.x)
    k, i = O.morton(cl.x, cl.facets, box, 30)
    k, i = O.sort_pairs(k, i, 30)
    tree, parent = O.karras(k, i, 30)
    exp = O.refit(tree.copy(), cl.x, v, np.float32(1 / 600), swept, cl.facets)
    for _ in range(3):  # repeated refits reuse the visit counters (parity trick)
        got = hip.lbvh_refit(c, tree, parent, cl.x, v, 1 / 600, swept, cl.facets)
        for f in ("min", "max", "child_a", "child_b"):
            assert np.array_equal(got[f], exp[f]), f


def test_broad_query_matches_brute_force(ctx):
    sc = scenes.config_drape(n=32, height=0.79)
    cl, sp = sc.meshes
    rng = np.random.default_rng(3)
    v = (rng.standard_normal(cl.x.shape) * 0.05).astype(np.float32)
    margin = sc.params["margin"]
    for (t, tobj) in ((cl, 0), (sp, 1)):
        box = O.scene_box(t.x)
        k, i = O.morton(t.x, t.facets, box, 30)
        k, i = O.sort_pairs(k, i, 30)
        tree, parent = O.karras(k, i, 30)
        tree = O.refit(tree, t.x, v if tobj == 0 else np.zeros_like(t.x), np.float32(1 / 600), 1 if tobj == 0 else 0,
                       t.facets)
        got = hip.broad_query(ctx, cl.x, 0, tree, t.facets, tobj, margin)
        gp = np.stack([got["point"], got["triangle"]], 1)
        T = t.facets.shape[0]
        leaf = tree[T - 1:]
        lb = np.zeros((T, 6), np.float32)
        lb[leaf["child_b"], 0:3] = leaf["min"]
        lb[leaf["child_b"], 3:6] = leaf["max"]
        bp = O.brute_candidates(cl.x, 0, lb, t.facets, tobj, margin)
        a = set(map(tuple, gp.tolist()))
        b = set(map(tuple, bp.tolist()))
        assert len(gp) == len(a), "duplicate pairs"
        assert a == b


# ------------------------------------------------------------------ full size (512 x 512 fold)
@pytest.fixture(scope="module")
def c2_full():
    sc = scenes.config_c2()
    m = sc.meshes[0]
    rng = np.random.default_rng(11)
    v = (rng.standard_normal(m.x.shape) * 0.05).astype(np.float32)
    box = O.scene_box(m.x)
    k, i = O.morton(m.x, m.facets, box, 63)
    ks, is_ = O.sort_pairs(k, i, 63)
    tree, parent = O.karras(ks, is_, 63)
    swept = O.refit(tree.copy(), m.x, v, np.float32(1 / 1200), 1, m.facets)
    return dict(sc=sc, m=m, v=v, box=box, k=k, i=i, ks=ks, is_=is_, tree=tree, parent=parent, swept=swept)


def test_full_morton_and_sort(ctx, c2_full):
    d = c2_full
    kg, ig = hip.morton(ctx, d["m"].x, d["m"].facets, d["box"], 63)
    assert np.array_equal(kg, d["k"]) and np.array_equal(ig, d["i"])
    ks, vs = hip.RadixSorter(ctx).sort(d["k"], d["i"], 63)
    assert np.array_equal(ks, d["ks"]) and np.array_equal(vs, d["is_"])


def test_full_hierarchy(ctx, c2_full):
    d = c2_full
    tg, pg = hip.lbvh_hierarchy(ctx, d["ks"], d["is_"], 63)
    assert np.array_equal(tg["child_a"], d["tree"]["child_a"])
    assert np.array_equal(tg["child_b"], d["tree"]["child_b"])
    assert np.array_equal(pg, d["parent"])


@pytest.mark.parametrize("formal", [False, True])
def test_full_refit(ctx, ctx_formal, c2_full, formal):
    d = c2_full
    c = ctx_formal if formal else ctx
    got = hip.lbvh_refit(c, d["tree"], d["parent"], d["m"].x, d["v"], 1 / 1200, 1, d["m"].facets)
    for f in ("min", "max", "child_a", "child_b"):
        assert np.array_equal(got[f], d["swept"][f]), f


def test_full_broad_query(ctx, c2_full):
    d = c2_full
    margin = d["sc"].params["margin"]
    got = hip.broad_query(ctx, d["m"].x, 0, d["swept"], d["m"].facets, 0, margin)
    exp = O.broad_query(d["m"].x, 0, d["swept"], d["m"].facets, 0, margin)
    g = np.stack([got["point"], got["triangle"]], 1).astype(np.uint64)
    e = exp.astype(np.uint64)
    gk = np.sort(g[:, 0] << np.uint64(32) | g[:, 1])
    ek = np.sort(e[:, 0] << np.uint64(32) | e[:, 1])
    assert np.array_equal(gk, ek)
