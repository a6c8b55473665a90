"""The agglomerative (Apetrei) build on the GPU: the stage entry point equals
the oracle's agglomerative tree (links, parents, boxes) node by node; in a
scene (build_mode = 1) the trees equal the oracle's node by node after each
rebuild, and states and contacts equal the oracle's bit for bit at rebuild
and mixed cadences, across builder switches and in a batched scene."""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import hip, scenes
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu


def _same_tree(got, exp):
    for f in ("min", "max", "child_a", "child_b"):
        assert np.array_equal(got[f], exp[f]), f


def _sorted_keys(m, bits, shift=0):
    k, i = O.morton(m.x, m.facets, O.scene_box(m.x), bits)
    if shift:
        k = k >> np.uint64(shift)
    return O.sort_pairs(k, i, bits)


@pytest.mark.parametrize("bits,shift", [(30, 0), (63, 0), (30, 12)])
def test_stage_matches_oracle(ctx, bits, shift):
    m = scenes.config_c1(seed=1).meshes[0]
    k, i = _sorted_keys(m, bits, shift)
    tg, pg = hip.lbvh_agglomerative(ctx, k, i, bits, m.x, m.facets)
    to, po = O.agglomerative(k, i, bits)
    to = O.refit(to, m.x, m.x, np.float32(0), 0, m.facets)
    _same_tree(tg, to)
    assert np.array_equal(pg, po)


@pytest.mark.parametrize("formal", [False, True])
def test_stage_full_size(ctx, ctx_formal, formal):
    """512 x 512 fold (522,242 leaves, 63-bit keys), both publication variants."""
    m = scenes.config_c2().meshes[0]
    k, i = _sorted_keys(m, 63)
    tg, pg = hip.lbvh_agglomerative(ctx_formal if formal else ctx, k, i, 63, m.x, m.facets)
    to, po = O.agglomerative(k, i, 63)
    to = O.refit(to, m.x, m.x, np.float32(0), 0, m.facets)
    _same_tree(tg, to)
    assert np.array_equal(pg, po)


def test_stage_duplicate_runs_and_tiny(ctx):
    n = 4097
    rng = np.random.default_rng(0)
    x = rng.random((3 * n, 3)).astype(np.float32)
    f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    k = np.repeat(np.arange(0, 64, dtype=np.uint64), n // 64 + 1)[:n]
    i = np.arange(n, dtype=np.uint32)
    tg, pg = hip.lbvh_agglomerative(ctx, k, i, 30, x, f)
    to, po = O.agglomerative(k, i, 30)
    to = O.refit(to, x, x, np.float32(0), 0, f)
    _same_tree(tg, to)
    assert np.array_equal(pg, po)
    for n in (1, 2, 3):
        kk = np.arange(n, dtype=np.uint64) * np.uint64(5)
        ii = np.arange(n, dtype=np.uint32)
        tg, pg = hip.lbvh_agglomerative(ctx, kk, ii, 30, x, f[:n])
        to, po = O.agglomerative(kk, ii, 30)
        to = O.refit(to, x, x, np.float32(0), 0, f[:n])
        _same_tree(tg, to)
        assert np.array_equal(pg, po)


@pytest.mark.parametrize("rebuild_every", [1, 3])
def test_scene_agglomerative_bit_exact(ctx, oracle_threads, rebuild_every):
    """A self-colliding 48 x 48 cloth draping a sphere (63-bit keys), rebuilt
    every substep (the fused build replaces the refit) or every third (refits
    climb parent links)."""
    print(f"agglomerative scene rebuild_every={rebuild_every}", flush=True)
    sc = scenes.config_drape(n=48, bits=63, rebuild_every=rebuild_every)
    hs, os_ = HipScene(ctx, sc), O.OracleScene(sc)
    hs.set_tuning(build_mode=1)
    os_.set_build_mode(1)
    for k in range(8):
        hs.step(6, graph=(k % 2 == 1))
        os_.step(6)
        for u, w in zip(hs.state(), os_.state()):
            assert np.array_equal(u, w)
    if rebuild_every == 1:
        gn, gp = hs.tree(0)
        on, op, _, _ = os_.tree(0)
        _same_tree(gn, on)
        assert np.array_equal(gp, op)
    og, cg = hs.contacts()
    oo, co = os_.contacts()
    assert np.array_equal(og, oo) and np.array_equal(cg, co)
    assert len(co) > 0 and hs.flags() == 0


def test_builder_switches_and_batched(ctx, oracle_threads):
    """Switching between the builders mid-run keeps results identical (the
    tree topology is the same), also for a batched scene of many segments."""
    sc = scenes.config_drape(n=32, rebuild_every=2)
    hs, os_ = HipScene(ctx, sc), O.OracleScene(sc)
    for k, mode in enumerate([1, 0, 1, 1, 0]):
        hs.set_tuning(build_mode=mode)
        hs.step(7, graph=(k % 2 == 0))
        os_.step(7)
        for u, w in zip(hs.state(), os_.state()):
            assert np.array_equal(u, w)
    bsc = scenes.config_c3(count=12, lo=16, hi=48, substeps=20)
    bsc.params["rebuild_every"] = 1
    a, b = HipScene(ctx, bsc), HipScene(ctx, bsc)
    b.set_tuning(build_mode=1)
    for k in range(4):
        a.step(5, graph=True)
        b.step(5, graph=True)
        for u, w in zip(a.state(), b.state()):
            assert np.array_equal(u, w)
    assert a.flags() == 0 and b.flags() == 0


@pytest.mark.parametrize("rebuild_every", [1, 3])
def test_karras_after_agglomerative_under_capture(ctx, oracle_threads, rebuild_every):
    """Regression: Karras nodes whose range and split hints pass skip rewriting
    their children's parent and `up` links.  After agglomerative builds (which
    write other parent links) the next Karras build must rewrite them all,
    also when the switch back captures a graph before its first (eager)
    substep runs.  Trees and parents after each Karras segment, and states
    after every segment, equal the oracle's (which builds with Karras)."""
    print(f"karras after agglomerative, rebuild_every={rebuild_every}", flush=True)
    sc = scenes.config_c2(n=96)
    sc.params["rebuild_every"] = rebuild_every
    hs, os_ = HipScene(ctx, sc), O.OracleScene(sc)
    for mode in (0, 1, 0, 1, 1, 0):
        hs.set_tuning(build_mode=mode)
        os_.set_build_mode(mode)
        hs.step(2 * rebuild_every + 1, graph=True)
        os_.step(2 * rebuild_every + 1)
        for u, w in zip(hs.state(), os_.state()):
            assert np.array_equal(u, w)
        if mode == 0 and rebuild_every == 1:
            gn, gp = hs.tree(0)
            on, op, _, _ = os_.tree(0)
            _same_tree(gn, on)
            assert np.array_equal(gp, op)
    assert hs.flags() == 0
