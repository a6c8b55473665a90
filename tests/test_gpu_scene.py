"""Scene-level parity: the full substep on the GPU vs the CPU oracle."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from ysim_amd import scenes
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu


def pair_set(a):
    return set(map(tuple, np.asarray(a).tolist()))


def assert_same_state(hs, os_):
    xg, vg, pg = hs.state()
    xo, vo, po = os_.state()
    assert np.array_equal(xg, xo)
    assert np.array_equal(vg, vo)
    assert np.array_equal(pg, po)


def assert_same_contacts(hs, os_):
    og, cg = hs.contacts()
    oo, co = os_.contacts()
    assert np.array_equal(og, oo)
    for f in ("point", "triangle", "query_obj", "target_obj", "normal_distance"):
        assert np.array_equal(cg[f], co[f]), f


@pytest.mark.parametrize("graph", [False, True])
def test_c1_bit_exact(ctx, oracle_threads, graph):
    sc = scenes.config_c1()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    assert pair_set(hs.pairs()) == pair_set(os_.pairs())
    hs.step(20, graph=graph)
    os_.step(20)
    assert_same_state(hs, os_)
    assert_same_contacts(hs, os_)
    assert hs.flags() == 0


@pytest.mark.parametrize("rebuild", [0, 1])
def test_drape_contacts_bit_exact(ctx, oracle_threads, rebuild):
    sc = scenes.config_drape(n=32, height=0.8, rebuild_every=rebuild)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    total = 0
    for _ in range(8):
        hs.step(10, graph=True)
        os_.step(10)
        assert_same_state(hs, os_)
        assert_same_contacts(hs, os_)
        total += len(os_.contacts()[1])
    assert total > 0, "scene produced no contacts"
    assert pair_set(hs.pairs()) == pair_set(os_.pairs())


def test_c2_small_fold_63bit(ctx, oracle_threads):
    sc = scenes.config_c2(n=96)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    for _ in range(3):
        hs.step(2)
        os_.step(2)
        assert_same_state(hs, os_)
        ng, pg, kg, ig = hs.tree(0, keys=True)
        no, po, ko, io = os_.tree(0)
        assert np.array_equal(kg, ko) and np.array_equal(ig, io)
        for f in ("min", "max", "child_a", "child_b"):
            assert np.array_equal(ng[f], no[f]), f
        assert np.array_equal(pg, po)
    assert pair_set(hs.pairs()) == pair_set(os_.pairs())


def test_c2_full_size_bit_exact(ctx, oracle_threads):
    """512 x 512 fold, 63-bit rebuild every substep: state, trees and candidate set."""
    sc = scenes.config_c2()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    assert pair_set(hs.pairs()) == pair_set(os_.pairs())
    hs.step(4, graph=True)
    os_.step(4)
    assert_same_state(hs, os_)
    ng, pg, kg, ig = hs.tree(0, keys=True)
    no, po, ko, io = os_.tree(0)
    assert np.array_equal(kg, ko) and np.array_equal(ig, io)
    assert np.array_equal(pg, po)
    for f in ("min", "max", "child_a", "child_b"):
        assert np.array_equal(ng[f], no[f]), f
    assert hs.flags() == 0


def test_snapshot_restore_replays_identically(ctx):
    sc = scenes.config_drape(n=24, height=0.8, rebuild_every=1)
    hs = HipScene(ctx, sc)
    hs.snapshot()
    hs.step(30, graph=True)
    a = hs.state()
    hs.restore()
    hs.step(30, graph=True)
    b = hs.state()
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


def test_set_collision_between_replays(ctx, oracle_threads):
    """Per-call margin / radius / thickness / self collision
    (detectCollisions, bvh.hpp:2853; narrowAndSortByVertices, bruteforce.hpp:311)
    changed between graph replays: the replayed graph reads the new values, a
    self-collision toggle re-captures it; states equal the oracle's."""
    sc = scenes.config_drape(n=32, height=0.8, rebuild_every=1)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    m0 = sc.params["margin"]
    for margin, radius, thick, self_ in ((m0, 0.0, m0, 1), (1.5 * m0, 0.002, 1.2 * m0, 1), (1.5 * m0, 0.002, m0, 0),
                                         (m0, 0.0, m0, 1)):
        hs.set_collision(margin, radius, thick, self_)
        os_.set_collision(margin, radius, thick, self_)
        hs.step(10, graph=True)
        os_.step(10)
        assert_same_state(hs, os_)
        assert_same_contacts(hs, os_)
    assert hs.flags() == 0
