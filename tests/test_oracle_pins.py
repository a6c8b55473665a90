"""Pins of the CPU oracle against independent closed forms, brute force and the
reference's own test invariants.  CPU only.

The reference holds no numeric golden vectors for this path (its self-test
checks invariants only), so each oracle function is checked here against an
algorithm it does not share code with, chosen so a plausible mistake (a wrong
mask, a flipped comparison, a swapped child) fails.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import abi, scenes


# ------------------------------------------------------------------ Morton
def spread_loop(v: np.ndarray, bits: int) -> np.ndarray:
    """Bit-by-bit spread: bit i of v moves to bit 3i (textbook definition)."""
    out = np.zeros_like(v, dtype=np.uint64)
    for i in range(bits):
        out |= ((v.astype(np.uint64) >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i)
    return out


def test_expand10_exhaustive():
    L = O.lib()
    v = np.arange(1024, dtype=np.uint64)
    got = np.array([L.yo_expand10(int(i)) for i in v], np.uint64)
    assert np.array_equal(got, spread_loop(v, 10))


def test_expand21_exhaustive():
    L = O.lib()
    v = np.arange(1 << 21, dtype=np.uint64)
    ref = spread_loop(v, 21)
    idx = np.random.default_rng(0).integers(0, 1 << 21, 20000)
    idx = np.concatenate([idx, [0, 1, 2, (1 << 21) - 1, (1 << 20), (1 << 20) - 1]])
    got = np.array([L.yo_expand21(int(i)) for i in idx], np.uint64)
    assert np.array_equal(got, ref[idx])


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_special_cases(bits):
    # box corners: all-min -> 0, all-max -> every bit set (clamped to 2^k - 1)
    x = np.array([[0, 0, 0], [1, 2, 4], [0, 0, 0], [1, 2, 4], [1, 2, 4]], np.float32)
    f = np.array([[0, 2, 0], [1, 3, 4]], np.uint32)
    box = O.scene_box(x)
    keys, idx = O.morton(x, f, box, bits)
    assert keys[0] == 0
    assert keys[1] == (1 << bits) - 1
    assert list(idx) == [0, 1]
    # degenerate width: a flat (y = const) box uses max(width, 1e-8) and y maps to 0
    y = np.array([[0, 5, 0], [1, 5, 1], [1, 5, 0]], np.float32)
    k2, _ = O.morton(y, np.array([[0, 1, 2]], np.uint32), O.scene_box(y), bits)
    mask_y = int(spread_loop(np.array([(1 << (bits // 3)) - 1], np.uint64), bits // 3)[0]) << 1
    assert int(k2[0]) & mask_y == 0


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_interleaves_quantized_centroid(bits):
    sc = scenes.config_c1()
    m = sc.meshes[0]
    box = O.scene_box(m.x)
    keys, _ = O.morton(m.x, m.facets, box, bits)
    q = bits // 3
    # quantized centroid per axis, in float32 as the kernel does
    c = ((m.x[m.facets[:, 0]] + m.x[m.facets[:, 1]]) + m.x[m.facets[:, 2]]) / np.float32(3)
    w = np.maximum(box[3:] - box[:3], np.float32(1e-8))
    u = np.clip((c - box[:3]) / w * np.float32(1 << q), 0, (1 << q) - 1).astype(np.uint64)
    ref = (spread_loop(u[:, 0], q) << np.uint64(2)) | (spread_loop(u[:, 1], q) << np.uint64(1)) | spread_loop(u[:, 2], q)
    assert np.array_equal(keys, ref)


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("n,bits", [(1, 30), (1000, 30), (54321, 63), (4096, 38)])
def test_sort_is_stable_argsort(n, bits):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << min(bits, 62), n, dtype=np.uint64) & np.uint64((1 << bits) - 1)
    keys[: n // 2] = keys[n // 2:][: n // 2] if n > 1 else keys[:0]
    vals = np.arange(n, dtype=np.uint32)
    ks, vs = O.sort_pairs(keys, vals, bits)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(vs, perm)
    assert np.array_equal(ks, keys[perm])


# ------------------------------------------------------------------ Karras
def topdown_tree(keys, idx, bits):
    """Radix-tree construction from the augmented keys (code, index): split each
    range where the highest differing bit flips.  Independent of delta/findSplit."""
    n = len(keys)
    aug = [(int(k) << 32) | int(i) for k, i in zip(keys, idx)]
    child = {}
    stack = [(0, n - 1, 0)]
    while stack:
        lo, hi, node = stack.pop()
        a, b = aug[lo], aug[hi]
        top = (a ^ b).bit_length() - 1
        s = lo
        while s + 1 <= hi and not ((aug[s + 1] >> top) & 1):
            s += 1
        ca = s + n - 1 if s == lo else s
        cb = s + 1 + n - 1 if s + 1 == hi else s + 1
        child[node] = (ca, cb)
        if s != lo:
            stack.append((lo, s, s))
        if s + 1 != hi:
            stack.append((s + 1, hi, s + 1))
    return child


def test_karras_long_duplicate_runs():
    """Runs of many equal codes: the split follows the index tie-break."""
    n = 4097
    k = np.repeat(np.arange(0, 64, dtype=np.uint64), n // 64 + 1)[:n]
    i = np.arange(n, dtype=np.uint32)
    tree, parent = O.karras(k, i, 30)
    child = topdown_tree(k, i, 30)
    for node, (ca, cb) in child.items():
        assert (tree[node]["child_a"], tree[node]["child_b"]) == (ca, cb)
        assert parent[ca] == node and parent[cb] == node
    # every slot except the root has exactly one parent
    refs = np.concatenate([tree[: n - 1]["child_a"], tree[: n - 1]["child_b"]])
    assert np.array_equal(np.sort(refs), np.arange(1, 2 * n - 1))


@pytest.mark.parametrize("bits,dups", [(30, False), (63, False), (30, True)])
def test_karras_matches_topdown(bits, dups):
    sc = scenes.config_c1(n=40)
    m = sc.meshes[0]
    box = O.scene_box(m.x)
    k, i = O.morton(m.x, m.facets, box, bits)
    if dups:
        k = k >> np.uint64(12)  # many equal codes: the index tie-break decides
    k, i = O.sort_pairs(k, i, bits)
    tree, parent = O.karras(k, i, bits)
    n = len(k)
    child = topdown_tree(k, i, bits)
    assert len(child) == n - 1
    for node, (ca, cb) in child.items():
        assert (tree[node]["child_a"], tree[node]["child_b"]) == (ca, cb)
        assert parent[ca] == node and parent[cb] == node
    assert parent[0] == -1
    assert np.array_equal(tree[n - 1:]["child_a"], np.full(n, -1))
    assert np.array_equal(tree[n - 1:]["child_b"], i.astype(np.int32))


class MidpointKarras:
    """The reference's own build rule, restated for n small: delta and
    determineRange (src/metal/bvh.metal:226-258) with the index tie-break, and
    findSplit (bvh.metal:197-224), which returns the midpoint when the range's
    end codes are equal and otherwise measures prefixes on the codes alone.
    The child numbering follows buildTree_Tri (bvh.metal:383-392)."""

    def __init__(self, keys, idx):
        self.k = [int(v) for v in keys]
        self.i = [int(v) for v in idx]
        self.n = len(self.k)

    @staticmethod
    def clz(v):
        return 32 - int(v).bit_length()

    def delta(self, a, b):
        if b < 0 or b >= self.n:
            return -1
        if self.k[a] != self.k[b]:
            return self.clz(self.k[a] ^ self.k[b])
        return 32 + self.clz(self.i[a] ^ self.i[b])

    def range(self, x):
        d = 1 if self.delta(x, x + 1) - self.delta(x, x - 1) >= 0 else -1
        dmin = self.delta(x, x - d)
        lmax = 2
        while self.delta(x, x + lmax * d) > dmin:
            lmax <<= 1
        length, t = 0, lmax >> 1
        while t >= 1:
            if self.delta(x, x + (length + t) * d) > dmin:
                length += t
            t >>= 1
        j = x + length * d
        return (j, x) if d < 0 else (x, j)

    def split(self, lo, hi):
        c0, c1 = self.k[lo], self.k[hi]
        if c0 == c1:
            return (lo + hi) >> 1
        common = self.clz(c0 ^ c1)
        s, step = lo, hi - lo
        while True:
            step = (step + 1) >> 1
            ns = s + step
            if ns < hi and self.clz(c0 ^ self.k[ns]) > common:
                s = ns
            if step <= 1:
                return s

    def children(self, node):
        lo, hi = self.range(node)
        s = self.split(lo, hi)
        ca = s + self.n - 1 if s == lo else s
        cb = s + self.n if s + 1 == hi else s + 1
        return (lo, hi), (ca, cb)


def claims(children, n):
    """Parent count of every node slot 0 .. 2n-2."""
    c = np.zeros(2 * n - 1, np.int64)
    for ca, cb in children:
        c[ca] += 1
        c[cb] += 1
    return c


def test_midpoint_findsplit_double_claims_on_equal_runs():
    """Five equal codes: the reference's midpoint split makes internal node 2
    a child of both node 0 and node 3 and leaves leaf 4 without a parent; the
    oracle (prefix by delta, index tie-break) gives every slot but the root
    exactly one parent.  Runs of 2 to 4 equal codes happen to agree."""
    n = 5
    k = np.zeros(n, np.uint64)
    i = np.arange(n, dtype=np.uint32)
    ref = MidpointKarras(k, i)
    kids = [ref.children(x)[1] for x in range(n - 1)]
    assert kids[0] == (2, 3) and kids[3] == (1, 2)
    c = claims(kids, n)
    assert c[2] == 2 and c[n - 1 + 4] == 0
    tree, parent = O.karras(k, i, 30)
    got = claims(zip(tree[: n - 1]["child_a"], tree[: n - 1]["child_b"]), n)
    assert got[0] == 0 and np.all(got[1:] == 1)
    assert (tree[0]["child_a"], tree[0]["child_b"]) == (3, n - 1 + 4)
    for run in range(2, 5):
        kr = np.zeros(run, np.uint64)
        ir = np.arange(run, dtype=np.uint32)
        mk = MidpointKarras(kr, ir)
        tr, _ = O.karras(kr, ir, 30)
        for x in range(run - 1):
            assert mk.children(x)[1] == (tr[x]["child_a"], tr[x]["child_b"])


@pytest.mark.parametrize("seed", range(4))
def test_midpoint_findsplit_agrees_when_end_codes_differ(seed):
    """On every range whose end codes differ the reference's findSplit and the
    oracle's pick the same child pair, duplicates inside the range included;
    with equal end codes and runs of >= 5 some node gets another child."""
    rng = np.random.default_rng(seed)
    n = 300
    k = np.sort(rng.integers(0, 40, n).astype(np.uint64))  # runs of about 7 equal codes
    i = rng.permutation(n).astype(np.uint32)
    order = np.lexsort((i, k))
    k, i = k[order], i[order]
    ref = MidpointKarras(k, i)
    tree, _ = O.karras(k, i, 30)
    differ = same_end = 0
    for x in range(n - 1):
        (lo, hi), kids = ref.children(x)
        mine = (int(tree[x]["child_a"]), int(tree[x]["child_b"]))
        if k[lo] != k[hi]:
            assert kids == mine, x
        else:
            same_end += 1
            differ += kids != mine
    assert same_end > 0 and differ > 0
    assert np.count_nonzero(claims([ref.children(x)[1] for x in range(n - 1)], n)[1:] != 1) > 0


def test_refit_is_exact_union():
    sc = scenes.config_c1(n=24)
    m = sc.meshes[0]
    rng = np.random.default_rng(1)
    v = (rng.standard_normal(m.x.shape) * 0.3).astype(np.float32)
    dt = np.float32(1 / 600)
    box = O.scene_box(m.x)
    k, i = O.morton(m.x, m.facets, box, 30)
    k, i = O.sort_pairs(k, i, 30)
    tree, _ = O.karras(k, i, 30)
    tree = O.refit(tree, m.x, v, dt, 1, m.facets)
    n = len(k)

    def leafbox(fid):
        f = m.facets[fid]
        p = m.x[f]
        s = (p + v[f] * dt).astype(np.float32)
        allp = np.concatenate([p, s])
        return allp.min(0), allp.max(0)

    def rec(node):
        if tree[node]["child_a"] < 0:
            return leafbox(tree[node]["child_b"])
        a = rec(tree[node]["child_a"])
        b = rec(tree[node]["child_b"])
        return np.minimum(a[0], b[0]), np.maximum(a[1], b[1])

    for node in [0, 1, n // 2, n - 2, n - 1, 2 * n - 2]:
        lo, hi = rec(node)
        assert np.array_equal(tree[node]["min"], lo) and np.array_equal(tree[node]["max"], hi)


# ------------------------------------------------------------------ candidates
def test_traversal_equals_brute_force():
    sc = scenes.config_drape(n=32, height=0.79)
    cl, sp = sc.meshes
    margin = sc.params["margin"]
    for t, tobj in ((cl, 0), (sp, 1)):
        box = O.scene_box(t.x)
        k, i = O.morton(t.x, t.facets, box, 30)
        k, i = O.sort_pairs(k, i, 30)
        tree, _ = O.karras(k, i, 30)
        tree = O.refit(tree, t.x, np.zeros_like(t.x), np.float32(0), 0, t.facets)
        got = O.broad_query(cl.x, 0, tree, t.facets, tobj, margin)
        T = t.facets.shape[0]
        leaf = tree[T - 1:]
        lb = np.zeros((T, 6), np.float32)
        lb[leaf["child_b"], :3] = leaf["min"]
        lb[leaf["child_b"], 3:] = leaf["max"]
        bf = O.brute_candidates(cl.x, 0, lb, t.facets, tobj, margin)
        assert set(map(tuple, got.tolist())) == set(map(tuple, bf.tolist()))
        if tobj == 0:
            # self query never returns a facet that contains the vertex
            f = t.facets[got[:, 1]]
            assert not np.any(f == got[:, :1])


# ------------------------------------------------------------------ narrow
def tri_case(p, pprev, tri):
    pairs = np.array([[0, 0]], np.uint32)
    xq = np.array([p], np.float32)
    xqp = np.array([pprev], np.float32)
    xt = np.array(tri, np.float32)
    return pairs, xq, xqp, xt, np.array([[0, 1, 2]], np.uint32)


def test_narrow_hand_cases():
    tri = [[0, 0, 0], [1, 0, 0], [0, 0, 1]]
    band = 0.01
    # above the triangle, inside the band: contact at signed distance 0.005
    fl, nd = O.narrow(*tri_case([0.2, 0.005, 0.2], [0.2, 0.005, 0.2], tri), False, 0.0, band)
    assert fl[0] and abs(nd[0, 3] - 0.005) < 1e-7 and abs(abs(nd[0, 1]) - 1) < 1e-7
    # outside the band: no contact
    fl, _ = O.narrow(*tri_case([0.2, 0.02, 0.2], [0.2, 0.02, 0.2], tri), False, 0.0, band)
    assert not fl[0]
    # crossed the plane this step: contact with negative distance, normal toward xprev
    fl, nd = O.narrow(*tri_case([0.2, -0.03, 0.2], [0.2, 0.05, 0.2], tri), False, 0.0, band)
    assert fl[0] and nd[0, 3] < 0 and nd[0, 1] > 0
    # projects outside the triangle: no contact
    fl, _ = O.narrow(*tri_case([0.9, 0.001, 0.9], [0.9, 0.001, 0.9], tri), False, 0.0, band)
    assert not fl[0]
    # degenerate triangle: no contact
    fl, _ = O.narrow(*tri_case([0.0, 0.0, 0.0], [0, 0, 0], [[0, 0, 0], [1, 0, 0], [2, 0, 0]]), False, 0.0, band)
    assert not fl[0]


def test_narrow_fp32_vs_fp64_ties_reported():
    rng = np.random.default_rng(7)
    n = 20000
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    pts = np.stack([rng.uniform(-0.2, 1.2, n), rng.normal(0, 0.01, n), rng.uniform(-0.2, 1.2, n)], 1).astype(np.float32)
    prev = (pts + rng.normal(0, 0.01, pts.shape)).astype(np.float32)
    pairs = np.stack([np.arange(n), np.zeros(n)], 1).astype(np.uint32)
    f = np.array([[0, 1, 2]], np.uint32)
    f32, nd32 = O.narrow(pairs, pts, prev, tri, f, False, 0.0, 0.01)
    f64, nd64 = O.narrow(pairs, pts, prev, tri, f, False, 0.0, 0.01, f64=True)
    ties = np.nonzero(f32 != f64)[0]
    # decisions differ only at a threshold (|d - band|, barycentric edge) within float rounding
    assert len(ties) <= n * 1e-3
    both = f32 & f64
    assert np.abs(nd32[both, 3] - nd64[both, 3]).max() < 1e-6
    print(f"narrow FP32/FP64 decision ties: {len(ties)} of {n}")


# ------------------------------------------------------------------ springs / integrate
def test_forces_fp32_close_to_fp64():
    sc = scenes.config_c1(n=32)
    m = sc.meshes[0]
    rng = np.random.default_rng(2)
    x = (m.x + rng.normal(0, 1e-3, m.x.shape)).astype(np.float32)
    v = rng.normal(0, 0.1, m.x.shape).astype(np.float32)
    ext = np.tile(np.array([0, -9.81, 0], np.float32), (len(x), 1))
    f32 = O.fastgrid_forces(m.cloth, x, v, ext, -0.001, 0.01).reshape(-1, 3)
    f64 = O.fastgrid_forces(m.cloth, x, v, ext, -0.001, 0.01, f64=True).reshape(-1, 3)
    scale = np.abs(f64).max()
    assert np.abs(f32 - f64).max() / scale < 1e-5


def test_spring_forces_cancel():
    """Newton's third law: with zero ext and air drag, the spring forces sum to ~0."""
    sc = scenes.config_c1(n=20)
    m = sc.meshes[0]
    rng = np.random.default_rng(3)
    x = (m.x + rng.normal(0, 5e-3, m.x.shape)).astype(np.float64)
    v = rng.normal(0, 0.1, m.x.shape)
    f = O.fastgrid_forces(m.cloth, x, v, np.zeros_like(x), 0.0, 0.01, f64=True).reshape(-1, 3)
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max() * len(f)
    assert np.abs(f).max() > 1.0


def test_free_fall_closed_form():
    """No springs stretched (k = 0), no contacts: v_y = n*subh*g (FP32 accumulation)."""
    sc = scenes.config_c1(n=16)
    for m in sc.meshes:
        if m.cloth:
            m.cloth.update(kstretch=0.0, kshear=0.0, kbend=0.0)
    sc.params["kair"] = 0.0
    sc.meshes[1].x = sc.meshes[1].x - np.float32(5.0)  # sphere out of the way
    S = O.OracleScene(sc, twin=True)
    S.step(12)
    x, v, _ = S.state()
    vy = v[:256, 1]
    exp = 12 * (1 / 600) * -9.81
    assert np.abs(vy - exp).max() < 1e-5
    xd, vd = S.twin_state()
    exp64 = 12 * float(np.float32(1 / 600)) * float(np.float32(-9.81))  # the FP32 inputs, exact in FP64
    assert np.abs(vd[:256, 1] - exp64).max() < 1e-14


def test_integrate_contact_projection():
    """integrate_cloth_grid contact response on a hand case (physics.metal:243-311)."""
    x = np.zeros((9, 3), np.float32)
    v = np.array([[0, -1.0, 0]] * 9, np.float32)
    f = np.zeros((9, 3), np.float32)
    mass = np.ones(9, np.float32)
    fixed = np.ones(9, np.float32)
    off = np.zeros(10, np.uint32)
    off[5:] = 1  # vertex 4 has one contact
    rec = np.zeros(1, abi.NARROW)
    rec["target_obj"] = 1
    rec["normal_distance"] = [0, 2.0, 0, 0.001]  # non-unit normal is renormalized
    xo, vo = x.copy().reshape(-1), v.copy().reshape(-1)
    L = O.lib()
    L.yo_integrate_f32(9, np.float32(0.01), np.float32(0.005), O.P(xo, O.f32p), O.P(vo, O.f32p), O.P(f.reshape(-1), O.f32p),
                       O.P(mass, O.f32p), O.P(fixed, O.f32p), O.P(off, O.u32p), rec.ctypes.data)
    xo, vo = xo.reshape(-1, 3), vo.reshape(-1, 3)
    assert vo[4, 1] == 0.0  # inward normal velocity removed
    assert xo[4, 1] == np.float32(0.005) - np.float32(0.001)  # pushed out to thickness, then x += h*v (v = 0)
    assert vo[0, 1] == -1.0 and xo[0, 1] == np.float32(-0.01)


def test_nan_guard():
    """NAN-GUARD (self_test_inline.hpp:7104-7169): non-finite -> fallback, v = 0, flag."""
    x = np.zeros((4, 3), np.float32)
    v = np.zeros((4, 3), np.float32)
    f = np.zeros((4, 3), np.float32)
    f[1, 0] = np.inf
    f[2, 2] = 1e12
    mass = np.ones(4, np.float32)
    fixed = np.ones(4, np.float32)
    xo, vo = x.copy().reshape(-1), v.copy().reshape(-1)
    touched = O.lib().yo_integrate_f32(4, np.float32(0.01), np.float32(0.0), O.P(xo, O.f32p), O.P(vo, O.f32p),
                                       O.P(f.reshape(-1), O.f32p), O.P(mass, O.f32p), O.P(fixed, O.f32p), None, None)
    xo, vo = xo.reshape(-1, 3), vo.reshape(-1, 3)
    assert touched == 2
    assert np.all(np.isfinite(xo)) and np.abs(xo).max() <= 1e4
    assert np.all(vo[1] == 0) and np.all(vo[2] == 0)
    assert xo[2, 2] == np.float32(1e4)


# ------------------------------------------------------------------ scene invariants
def test_primitive_closed_forms():
    # test/primitive_test.cpp:44-94 closed forms, generalised to slices x stacks
    for sl, st in ((32, 16), (8, 8), (24, 12)):
        x, f = scenes.uv_sphere(sl, st, 0.3, (0, 0, 0))
        assert len(x) == sl * (st - 1) + 2 and len(f) == 2 * sl * (st - 1)
        assert np.allclose(np.linalg.norm(x, axis=1), 0.3, atol=1e-6)
    for lv in (0, 1, 3):
        x, f = scenes.icosphere(lv, 1.0, (0, 0, 0))
        assert len(x) == 10 * 4 ** lv + 2 and len(f) == 20 * 4 ** lv
        # closed surface: every edge shared by exactly two faces
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)


def test_grid_facets_cover_every_quad():
    f = scenes.grid_facets(5)
    assert f.shape == (32, 3)
    assert list(f[0]) == [0, 5, 6] and list(f[1]) == [0, 6, 1]  # even quad (0,0)
    assert list(f[2]) == [1, 6, 2] and list(f[3]) == [2, 6, 7]  # odd quad (0,1)


def test_fg_scale_rest_lengths():
    """FG-SCALE (self_test_inline.hpp:5124-5193): rest lengths track the geometry."""
    for size in (1.0, 2.5):
        rng = np.random.default_rng(0)
        m = scenes.cloth_mesh(17, size, (0, 0, 0), rng)
        l = size / 16
        c = m.cloth
        for key, ref in (("stretch_rest_x", l), ("stretch_rest_y", l), ("shear_rest_a", l * math.sqrt(2)),
                         ("shear_rest_b", l * math.sqrt(2)), ("bend_rest_x", 2 * l), ("bend_rest_y", 2 * l)):
            assert abs(c[key] - ref) < 1e-3 * ref, key


def test_determinism_block11():
    """Block 11 (self_test_inline.hpp:1971-2053): two runs are bit-identical."""
    sc = scenes.config_drape(n=16, height=0.8)
    a, b = O.OracleScene(sc), O.OracleScene(sc)
    a.step(30)
    b.step(30)
    assert np.array_equal(a.state()[0], b.state()[0])


def test_drape_invariants_block6():
    """Block 6 (self_test_inline.hpp:1413-1517): mean y falls, contacts fire, no tunnelling."""
    sc = scenes.config_drape(n=24, height=0.81)
    S = O.OracleScene(sc)
    y0 = S.state()[0][: 24 * 24, 1].mean()
    fired = 0
    thick = sc.meshes[0].cloth["thickness"]
    sagitta = 0.3 * (1 - math.cos(math.pi / 12))  # facet chord below the true sphere
    for _ in range(12):
        S.step(5)
        fired += len(S.contacts()[1])
        x = S.state()[0][: 24 * 24]
        r = np.linalg.norm(x - np.array([0.5, 0.5, 0.5], np.float32), axis=1)
        assert r.min() > 0.3 - sagitta - 2 * thick  # no vertex tunnelled into the sphere
    assert x[:, 1].mean() < y0
    assert fired > 0


def test_contact_presence_absence_block13():
    """Block 13 (self_test_inline.hpp:2168-2250): contacts for overlap, none when separated."""
    near = scenes.config_drape(n=16, height=0.8)
    far = scenes.config_drape(n=16, height=1.5)
    a, b = O.OracleScene(near), O.OracleScene(far)
    a.step(40)
    b.step(5)
    assert len(a.contacts()[1]) > 0
    assert len(b.contacts()[1]) == 0


def test_literal_total_mass_explodes():
    """Why the configs use mass 1 per vertex: total mass 1 at k = 500 trips the NaN/box guard."""
    sc = scenes.config_c1(n=48, mass_mode="total", slices=8, stacks=6)
    S = O.OracleScene(sc)
    S.step(10)
    assert S.anomaly() == 1
    ok = O.OracleScene(scenes.config_c1(n=48, slices=8, stacks=6))
    ok.step(10)
    assert ok.anomaly() == 0
