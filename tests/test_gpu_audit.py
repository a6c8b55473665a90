"""Candidate-set parity of the query kernel a substep actually runs.

The scene audit (scene_set_audit / scene_get_audit) records every candidate
pair the substep's own query reports: the tile walk fused with the cloth step
(the timed kernel of the plain 512 x 512 workload), the tile walk alone, the
per-vertex walk, the hash query; eager substeps and graph replays alike.  The
records are compared, as multisets, with the oracle's candidate set at the
state the substep started from (queryPoints, src/metal/bvh.metal:1393-1412;
detectCollisions, include/sim/bvh.hpp:2853-2949).  query_tile = 2 shrinks the
tile walk's buffers so its leaf flushes and its per-lane redo both run; the
redo count shows they did.  The tree cut the tile walk starts from is read
back and must cover every leaf of the GPU tree exactly once.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from paritylib import assert_same_candidates, cut_cover
from ysim_amd import scenes
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu


def same_state(hs, os_):
    for a, b in zip(hs.state(), os_.state()):
        assert np.array_equal(a, b)


def audited_rounds(hs, os_, rounds, graph, what):
    """Each round of k substeps: the audit of its last substep against the
    oracle's candidates at the state that substep started from."""
    redo = 0
    for k in rounds:
        if k > 1:
            os_.step(k - 1)
        want = os_.pairs()
        os_.step(1)
        hs.step(k, graph=graph)
        got, r = hs.audit()
        redo += r
        assert_same_candidates(got, want, f"{what}, round of {k}")
        same_state(hs, os_)
    return redo


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("tile", [1, 2, 0])
def test_audit_c1(ctx, oracle_threads, tile, graph):
    sc = scenes.config_c1()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=tile)
    hs.set_audit(1 << 18)
    audited_rounds(hs, os_, (2, 8, 10), graph, f"c1 tile={tile}")
    assert hs.flags() == 0


@pytest.mark.parametrize("rebuild", [0, 1])
@pytest.mark.parametrize("tile", [1, 2])
def test_audit_drape_with_contacts(ctx, oracle_threads, tile, rebuild):
    sc = scenes.config_drape(n=32, height=0.8, rebuild_every=rebuild)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=tile)
    hs.set_audit(1 << 18)
    contacts = 0
    for _ in range(6):
        audited_rounds(hs, os_, (10,), True, f"drape tile={tile} rebuild={rebuild}")
        contacts += len(os_.contacts()[1])
    assert contacts > 0, "the drape produced no contacts"


@pytest.mark.parametrize("tile", [1, 2])
def test_audit_fold_96(ctx, oracle_threads, tile):
    sc = scenes.config_c2(n=96)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=tile, sort_incremental=1)
    hs.set_audit(1 << 20)
    redo = audited_rounds(hs, os_, (1, 3, 4, 8), True, f"96 fold tile={tile}")
    if tile == 2:
        assert redo > 0, "query_tile=2 ran no per-lane redo"
    cut = hs.cut(0)
    nodes, _ = hs.tree(0)
    # a replay whose Karras pass kept the previous topology lists no cut
    assert cut.size > 0 or hs.deferred_builds() > 0
    if cut.size:
        assert np.all(cut_cover(nodes, cut) == 1)


@pytest.mark.parametrize("tile", [1, 2])
def test_audit_c2_full(ctx, oracle_threads, tile):
    """512 x 512 fold, the bench's kernels (incremental sort, fused tile query
    and cloth step): candidates at substep 0 and after 4 and 8 substeps."""
    sc = scenes.config_c2()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=tile, sort_incremental=1, sort_items=16)
    hs.set_audit(1 << 22)
    redo = audited_rounds(hs, os_, (1, 4, 4), True, f"512 fold tile={tile}")
    if tile == 2:
        assert redo > 0
    nodes, _ = hs.tree(0)
    cut = hs.cut(0)
    T = sc.meshes[0].facets.shape[0]
    assert cut.size <= 512 and (cut.size > 0 or hs.deferred_builds() > 0)
    assert np.all(cut_cover(nodes, cut) == 1)
    # every cut node has at most ceil(T / 128) leaves and a parent with more
    _, parent = hs.tree(0)
    S = -(-T // 128)
    for k in cut[:16]:
        assert int(cut_cover(nodes, [k]).sum()) <= S
        if k != 0:
            assert int(cut_cover(nodes, [parent[k]]).sum()) > S
    assert hs.flags() == 0


def test_audit_batched_subset(ctx, oracle_threads):
    """8 ragged instances in one scene: waves that span two instances take the
    per-lane walk, the others the tile walk."""
    sc = scenes.config_c3(count=8, lo=16, hi=48)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_audit(1 << 20)
    audited_rounds(hs, os_, (1, 5, 10, 10), True, "batched 8")
    assert hs.flags() == 0


@pytest.mark.parametrize("broad_mode", [0, 1])
def test_audit_c5_small(ctx, oracle_threads, broad_mode):
    """256^2 cloth over 4 level-4 icospheres: the LBVH tile walk and the hash query."""
    sc = scenes.config_c5(n=256, level=4, height=0.93)
    if broad_mode:
        sc.params = dict(sc.params, broad_mode=1, hash_slots_log2=22, hash_cell=sc.params["hash_cell"])
    hs, os_ = HipScene(ctx, sc), OracleScene(scenes.config_c5(n=256, level=4, height=0.93))
    hs.set_audit(1 << 21)
    audited_rounds(hs, os_, (1, 4, 4), True, f"c5 small broad_mode={broad_mode}")
    assert hs.flags() == 0


def test_audit_toggles(ctx):
    """Off by default (get_audit refuses), an overflowing buffer reports the
    records it needs, and turning it off leaves the scene stepping."""
    from ysim_amd._lib import ERR_CAPACITY, YsimError
    sc = scenes.config_drape(n=24, height=0.8)
    hs = HipScene(ctx, sc)
    with pytest.raises(YsimError):
        hs.audit()
    hs.set_audit(1)
    hs.step(2, graph=True)
    with pytest.raises(YsimError) as e:
        hs.audit()
    assert e.value.code == ERR_CAPACITY
    hs.set_audit(0)
    hs.step(2, graph=True)
    assert hs.flags() == 0
