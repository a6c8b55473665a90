"""Launch tuning: every sort digit width / tile size and every query and cloth
block size gives bit-identical results; the autotuner leaves the state
untouched and its choice keeps results identical."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from ysim_amd import scenes
from ysim_amd._lib import YsimError
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu

SORT_CONFIGS = [(8, 8, 0, 0), (8, 16, 0, 0), (11, 8, 0, 0), (11, 16, 0, 0), (8, 8, 1, 0), (8, 16, 1, 0),
                (8, 8, 0, 1), (8, 16, 0, 1)]


def _same_state(a, b):
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


@pytest.mark.parametrize("mode", ["hybrid", "incremental"])
def test_sort_modes_full_fold_bit_exact(ctx, oracle_threads, mode):
    """The hybrid and incremental sorts on the full 512x512 fold (63-bit, rebuild
    every substep): state, keys and trees equal the oracle's."""
    sc = scenes.config_c2()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(sort_hybrid=int(mode == "hybrid"), sort_incremental=int(mode == "incremental"), sort_items=16)
    hs.step(3, graph=True)
    os_.step(3)
    _same_state(hs.state(), os_.state())
    ng, pg, kg, ig = hs.tree(0, keys=True)
    no, po, ko, io = os_.tree(0)
    assert np.array_equal(kg, ko) and np.array_equal(ig, io)
    for f in ("min", "max", "child_a", "child_b"):
        assert np.array_equal(ng[f], no[f]), f
    assert hs.flags() == 0


def test_incremental_sort_after_state_jumps(ctx, oracle_threads):
    """The incremental sort starts from the last build's order; after a restore
    (or a new state) that order is far from the new one, the window passes fail
    the check and the full sort runs.  Every state stays equal to the oracle's."""
    sc = scenes.config_c2(n=128)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(sort_incremental=1)
    hs.snapshot()
    x0, v0, p0 = os_.state()
    for graph in (True, False):
        hs.step(6, graph=graph)
        os_.step(6)
        _same_state(hs.state(), os_.state())
        hs.restore()
        os_.set_state(x0, v0, p0)
        _same_state(hs.state(), os_.state())
    # the cloth mirrored in x (an isometry, so the physics stays tame): every
    # key changes and the sorted order with it
    x, v, xp = hs.state()
    V0 = sc.meshes[0].x.shape[0]
    x[:V0, 0] = np.float32(1.0) - x[:V0, 0]
    hs.set_state(x, v, x)
    os_.set_state(x, v, x)
    hs.step(3, graph=True)
    os_.step(3)
    _same_state(hs.state(), os_.state())
    ng, pg, kg, ig = hs.tree(0, keys=True)
    no, po, ko, io = os_.tree(0)
    assert np.array_equal(kg, ko) and np.array_equal(ig, io)


@pytest.mark.parametrize("rb", [256, 1024])
def test_refit_blocks_bit_exact_c2(ctx, oracle_threads, rb):
    """Refit with 256- and 1024-leaf LDS blocks on the 256x256 fold: every node box
    equals the oracle's."""
    sc = scenes.config_c2(n=256)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(refit_block=rb)
    hs.step(3, graph=True)
    os_.step(3)
    _same_state(hs.state(), os_.state())
    ng, pg, kg, ig = hs.tree(0, keys=True)
    no, po, ko, io = os_.tree(0)
    for f in ("min", "max", "child_a", "child_b"):
        assert np.array_equal(ng[f], no[f]), f
    assert hs.flags() == 0


@pytest.mark.parametrize("db,items,hyb,inc", SORT_CONFIGS)
def test_sort_configs_bit_exact_63bit(ctx, oracle_threads, db, items, hyb, inc):
    print(f"sort config digit_bits={db} items={items} hybrid={hyb} incremental={inc} on the 96x96 fold", flush=True)
    sc = scenes.config_c2(n=96)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(sort_digit_bits=db, sort_items=items, sort_hybrid=hyb, sort_incremental=inc)
    assert hs.tuning()["sort_digit_bits"] == db
    for graph in (False, True):
        hs.step(2, graph=graph)
        os_.step(2)
        _same_state(hs.state(), os_.state())
        ng, pg, kg, ig = hs.tree(0, keys=True)
        no, po, ko, io = os_.tree(0)
        assert np.array_equal(kg, ko) and np.array_equal(ig, io)
        assert np.array_equal(pg, po)
        for f in ("min", "max", "child_a", "child_b"):
            assert np.array_equal(ng[f], no[f]), f
    assert hs.flags() == 0


@pytest.mark.parametrize("db,items,hyb,inc", SORT_CONFIGS)
def test_sort_configs_bit_exact_30bit_rebuild(ctx, oracle_threads, db, items, hyb, inc):
    sc = scenes.config_drape(n=40, height=0.8, rebuild_every=1)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(sort_digit_bits=db, sort_items=items, sort_hybrid=hyb, sort_incremental=inc)
    hs.step(12, graph=True)
    os_.step(12)
    _same_state(hs.state(), os_.state())
    ng, pg, kg, ig = hs.tree(0, keys=True)
    no, po, ko, io = os_.tree(0)
    assert np.array_equal(kg, ko) and np.array_equal(ig, io)


@pytest.mark.parametrize("qb,rb,qw", [(64, 256, 8), (128, 512, 5), (256, 1024, 6)])
@pytest.mark.parametrize("cb", [128, 256, 512])
def test_block_sizes_bit_exact(ctx, oracle_threads, qb, rb, qw, cb):
    """The per-vertex query kernel at every block size and occupancy target."""
    sc = scenes.config_drape(n=32, height=0.8)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_block=qb, cloth_block=cb, refit_block=rb, query_waves=qw, query_tile=0)
    hs.step(20, graph=True)
    os_.step(20)
    _same_state(hs.state(), os_.state())
    og, cg = hs.contacts()
    oo, co = os_.contacts()
    assert np.array_equal(og, oo) and len(co) > 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_query_modes_drape_bit_exact(ctx, oracle_threads, mode):
    """Per-vertex walk (0), tile walk (1) and the tile walk pushed onto its
    per-vertex fallback (2) on a 30x30 cloth (its edge rows and columns are not
    whole 8x8 tiles) over a UV sphere: two targets per query vertex."""
    print(f"query_tile={mode} on the 30x30 drape", flush=True)
    sc = scenes.config_drape(n=30, height=0.8)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=mode)
    for graph in (False, True):
        hs.step(12, graph=graph)
        os_.step(12)
        _same_state(hs.state(), os_.state())
        og, cg = hs.contacts()
        oo, co = os_.contacts()
        assert np.array_equal(og, oo) and np.array_equal(cg, co)
    assert len(co) > 0 and hs.flags() == 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_query_modes_fold_bit_exact(ctx, oracle_threads, mode):
    """The three query modes on the 96x96 fold (self queries, 63-bit tree)."""
    print(f"query_tile={mode} on the 96x96 fold", flush=True)
    sc = scenes.config_c2(n=96)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=mode)
    hs.step(30, graph=True)
    os_.step(30)
    _same_state(hs.state(), os_.state())
    og, cg = hs.contacts()
    oo, co = os_.contacts()
    assert np.array_equal(og, oo) and np.array_equal(cg, co)
    assert hs.flags() == 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_query_modes_batched_identical(ctx, mode):
    """Ragged batch (tiles straddle instance boundaries): every query mode gives
    the same states."""
    sc = scenes.config_c3(count=12, hi=48)
    ref, hs = HipScene(ctx, sc), HipScene(ctx, sc)
    ref.set_tuning(query_tile=0)
    hs.set_tuning(query_tile=mode)
    ref.step(15, graph=True)
    hs.step(15, graph=True)
    _same_state(hs.state(), ref.state())
    assert hs.counters()["pairs"] == ref.counters()["pairs"]


def test_autotune_keeps_results(ctx, oracle_threads):
    print("autotune on the 128x128 fold (63-bit, rebuild every substep)", flush=True)
    sc = scenes.config_c2(n=128)
    a, b = HipScene(ctx, sc), HipScene(ctx, sc)
    a.step(3, graph=True)
    b.step(3, graph=True)
    before = a.state()
    c0 = a.counters()
    t = a.autotune(6)
    print("autotune:", t, flush=True)
    _same_state(a.state(), before)
    assert a.counters()["steps"] == c0["steps"]
    assert t["ms_per_substep"] > 0 and t["ms_per_substep"] <= t["ms_per_substep_default"]
    assert a.tuning()["query_block"] == t["query_block"]
    a.step(8, graph=True)
    b.step(8, graph=True)
    _same_state(a.state(), b.state())
    os_ = OracleScene(sc)
    os_.step(11)
    _same_state(a.state(), os_.state())


def test_autotune_batched_keeps_results(ctx):
    sc = scenes.config_c3(count=24, hi=96)
    a, b = HipScene(ctx, sc), HipScene(ctx, sc)
    t = a.autotune(4)
    print("autotune batched:", t, flush=True)
    a.step(10, graph=True)
    b.step(10, graph=True)
    _same_state(a.state(), b.state())


def test_invalid_tuning_rejected(ctx):
    hs = HipScene(ctx, scenes.config_drape(n=16))
    for kw in ({"sort_digit_bits": 9}, {"sort_items": 4}, {"query_block": 96}, {"cloth_block": 1024},
               {"refit_block": 128}, {"sort_hybrid": 2}, {"sort_incremental": 3}, {"query_waves": 7},
               {"query_tile": 3}):
        with pytest.raises(YsimError):
            hs.set_tuning(**kw)
    hs.step(2)


def test_graph_replays_keep_topology_when_order_check_fails(ctx, oracle_threads):
    """A spinning fold (20 rad/s about y): the Morton order changes by many
    places per substep, the incremental check fails in graph replays, which
    then refit the previous topology (no fallback kernels in the graph) and
    the next call runs an eager substep with the full sort.  States and
    contacts stay equal to the oracle's; eager-only runs never defer."""
    sc = scenes.config_c2(n=64)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(sort_incremental=1)
    x, v, xp = os_.state()
    V0 = sc.meshes[0].x.shape[0]
    c = x[:V0].mean(axis=0)
    w = np.float32(20.0)
    v[:V0, 0] = -w * (x[:V0, 2] - c[2])
    v[:V0, 2] = w * (x[:V0, 0] - c[0])
    for s in (hs, os_):
        s.set_state(x, v, xp)
    for k in range(6):
        hs.step(8, graph=True)
        os_.step(8)
        _same_state(hs.state(), os_.state())
    og, cg = hs.contacts()
    oo, co = os_.contacts()
    assert np.array_equal(og, oo) and np.array_equal(cg, co)
    assert hs.deferred_builds() > 0 and hs.flags() == 0
    n0 = hs.deferred_builds()
    hs.step(8, graph=False)
    os_.step(8)
    _same_state(hs.state(), os_.state())
    assert hs.deferred_builds() == n0


def _same_keys_and_tree(hs, os_, mesh, keys=True):
    if keys:
        ng, pg, kg, ig = hs.tree(mesh, keys=True)
    else:
        ng, pg = hs.tree(mesh)
    no, po, ko, io = os_.tree(mesh)
    if keys:
        assert np.array_equal(kg, ko) and np.array_equal(ig, io)
    for f in ("min", "max", "child_a", "child_b"):
        assert np.array_equal(ng[f], no[f]), f
    assert np.array_equal(pg, po)


def test_box_from_cloth_kernel_across_block_sizes(ctx, oracle_threads):
    """One dynamic mesh: the cloth kernel reduces the next build's Morton box.
    Switching its block size (another grid, the same partials buffer) between
    graph runs keeps keys and trees equal to the oracle's after each rebuild."""
    sc = scenes.config_c2(n=96)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(sort_incremental=1)
    for cb in (128, 512, 256, 128, 512):
        hs.set_tuning(cloth_block=cb)
        hs.step(6, graph=True)
        os_.step(6)
        _same_state(hs.state(), os_.state())
        # the last substep rebuilt from the box the previous cloth kernel wrote
        _same_keys_and_tree(hs, os_, 0)
    assert hs.flags() == 0


def test_one_mesh_groups_of_different_sizes_share_box_partials(ctx, oracle_threads):
    """63-bit keys leave no room for a mesh index: every cloth of a batched
    scene is its own build group, and their one-mesh box kernels (grids of
    different sizes) share one partials buffer.  Trees equal the oracle's for
    every cloth (the topology follows from each cloth's box and keys), keys
    for the last group, whose sort buffers stay resident; asking another
    cloth for its keys is an error."""
    sc = scenes.config_c3(count=4, lo=16, hi=48, substeps=20)
    sc.params["morton_bits"] = 63
    sc.params["rebuild_every"] = 1
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    for k in range(3):
        hs.step(4, graph=(k % 2 == 0))
        os_.step(4)
        _same_state(hs.state(), os_.state())
    dyn = [m for m, mesh in enumerate(sc.meshes) if mesh.behavior != scenes.BEHAVIOR_FLOAT]
    for m in dyn:
        _same_keys_and_tree(hs, os_, m, keys=(m == dyn[-1]))
    with pytest.raises(YsimError):
        hs.tree(dyn[0], keys=True)
    assert hs.flags() == 0
