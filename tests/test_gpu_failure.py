"""Failure handling forced from the host, on the GPU, against the oracle.

* An exploding cloth (the reference's NAN-GUARD block,
  test/self_test_inline.hpp:7104-7169: stiff springs the explicit integrator
  cannot hold) is kept finite and inside the world bound by the integrate
  guard (sanitizeIntegrateOutput, src/metal/physics.metal:40-74) through
  rebuilds, with the anomaly flag (bit 0) raised; states equal the oracle's.
  The spatial-hash broad phase takes the same scene: boxes of thrown
  triangles span millions of cells and go to its direct list.
* A NaN velocity or a NaN position written through set_state is sanitized the
  same way as in the oracle; the NaN query vertex's candidate set (every box,
  under the closed overlap test) equals the oracle's too.
* Capacity: a vertex with more contacts than contact_capacity raises flag
  bit 1 and the contact export refuses to truncate; a pair buffer that is too
  small returns YSIM_ERR_CAPACITY with the size it needs (scene.hpp:1135-1139
  drops pairs silently in the reference).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import OracleScene
from paritylib import assert_same_candidates
from ysim_amd import abi, hip, scenes
from ysim_amd._lib import ERR_CAPACITY, YsimError, load
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu
WORLD_BOUND = 1.0e4


def exploding_scene(broad_mode=0, rebuild=1, self_collision=0):
    """An 8 x 8 cloth of total mass 1 with k_stretch 1e5 above a ground plane:
    dt sqrt(k/m) = 4.2 at dt = 1/600, beyond the explicit integrator's bound."""
    rng = np.random.default_rng(21)
    cl = scenes.cloth_mesh(8, 0.8, (0.0, 0.4, 0.0), rng, k=(1e5, 50.0, 5.0), mass=1.0 / 64)
    gx, gf = scenes.plane_mesh(2.0, (0.4, 0.0, 0.4))
    p = scenes.sim_params(subh=1 / 600, margin=0.5 * 0.8 / 7, morton_bits=30, rebuild_every=rebuild,
                          broad_mode=broad_mode, self_collision=self_collision)
    return scenes.Scene([cl, scenes.obstacle(gx, gf)], p, name=f"exploding-8-mode{broad_mode}")


def same_state(hs, os_, what):
    """Bit patterns equal (xPrev keeps an injected NaN for one substep)."""
    for a, b in zip(hs.state(), os_.state()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), what


@pytest.mark.parametrize("broad_mode,rebuild", [(0, 1), (0, 0), (1, 1)])
def test_exploding_cloth_guarded(ctx, oracle_threads, broad_mode, rebuild):
    sc = exploding_scene(broad_mode, rebuild)
    hs = HipScene(ctx, sc)
    os_ = OracleScene(exploding_scene(0, rebuild))
    hs.step(9, graph=True)
    os_.step(9)
    same_state(hs, os_, "after 9 substeps")
    x = hs.state()[0]
    assert np.all(np.isfinite(x)) and np.abs(x).max() <= WORLD_BOUND + 1
    for _ in range(6):
        hs.step(40, graph=True)
        os_.step(40)
        same_state(hs, os_, "through 249 substeps")
    x, v, _ = hs.state()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v)) and np.abs(x).max() <= WORLD_BOUND + 1
    assert np.abs(x[:64]).max() > 10.0, "the cloth did not explode"
    assert hs.flags() & 1 and os_.anomaly()


@pytest.mark.parametrize("broad_mode", [0, 1])
def test_exploding_self_colliding_cloth_stays_finite(ctx, broad_mode):
    """With self collision the thrown vertices pile up at the world bound's
    corners, each within the band of many degenerate facets: more contacts
    than slots (flag bit 1; the kept subset then differs from the oracle's
    unbounded list, so states are not compared).  The guard still keeps every
    position finite and inside the bound through rebuilds."""
    hs = HipScene(ctx, exploding_scene(broad_mode, 1, self_collision=1))
    for _ in range(6):
        hs.step(40, graph=True)
        x, v, _ = hs.state()
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(v)) and np.abs(x).max() <= WORLD_BOUND + 1
    assert hs.flags() & 1
    assert hs.flags() & 2


def nan_drape():
    return scenes.config_drape(n=24, height=0.8, rebuild_every=1)


def test_nan_velocity_sanitized(ctx, oracle_threads):
    sc = nan_drape()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.step(10, graph=True)
    os_.step(10)
    x, v, xp = os_.state()
    v[300] = np.nan
    hs.set_state(x, v, xp)
    os_.set_state(x, v, xp)
    for k in (1, 4, 10):
        hs.step(k, graph=True)
        os_.step(k)
        same_state(hs, os_, f"NaN velocity, round {k}")
    assert hs.flags() & 1 and os_.anomaly()
    assert all(np.all(np.isfinite(a)) for a in hs.state())


@pytest.mark.parametrize("tile", [1, 0])
def test_nan_position_sanitized(ctx, oracle_threads, tile):
    sc = nan_drape()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=tile)
    hs.step(10, graph=True)
    os_.step(10)
    x, v, xp = os_.state()
    x[300] = np.nan
    hs.set_state(x, v, xp)
    os_.set_state(x, v, xp)
    hs.set_audit(1 << 16)
    want = os_.pairs()
    os_.step(1)
    hs.step(1)
    got, _ = hs.audit()
    assert np.count_nonzero(want[:, 1] == 300) > 1000  # the NaN box meets every leaf
    assert_same_candidates(got, want, f"NaN query vertex, tile={tile}")
    same_state(hs, os_, "NaN position, first substep")
    hs.step(9, graph=True)
    os_.step(9)
    same_state(hs, os_, "NaN position, 10 substeps")
    assert hs.flags() & 1 and os_.anomaly()
    assert all(np.all(np.isfinite(a)) for a in hs.state())


def test_contact_capacity_overflow_flagged(ctx):
    # the disk drape: some vertices meet two facets within the band (the
    # oracle finds two contacts on one vertex by substep 70)
    sc = scenes.config_tri_drape()
    sc.params = dict(sc.params, contact_capacity=1)
    hs = HipScene(ctx, sc)
    for _ in range(20):
        hs.step(10, graph=True)
        if hs.flags() & 2:
            break
    assert hs.flags() & 2, "no vertex exceeded one contact slot"
    with pytest.raises(YsimError) as e:
        hs.contacts()
    assert e.value.code == ERR_CAPACITY


def test_pair_capacity_reports_required_size(ctx, oracle_threads):
    sc = scenes.config_c2(n=48)
    m = sc.meshes[0]
    margin = sc.params["margin"]
    k, i = hip.morton(ctx, m.x, m.facets, hip.scene_box(ctx, m.x), 30)
    k, i = hip.RadixSorter(ctx).sort(k, i, 30)
    tree, parent = hip.lbvh_hierarchy(ctx, k, i, 30)
    tree = hip.lbvh_refit(ctx, tree, parent, m.x, m.x, 0.0, 0, m.facets)
    dx = ctx.to_device(np.asarray(m.x, np.float32).reshape(-1))
    dtree = ctx.to_device(np.ascontiguousarray(tree))
    df = ctx.to_device(np.asarray(m.facets, np.uint32).reshape(-1))
    out = ctx.zeros(1, abi.BROAD)
    cnt = ctx.zeros(1, np.uint32)
    req = C.c_uint32(0)
    rc = load().ysim_hip_broad_query(ctx.h, dx.ptr, m.x.shape[0], 0, 1, 0, dtree.ptr, m.facets.shape[0], df.ptr, 0,
                                     1, 0, float(margin), out.ptr, 1, cnt.ptr, C.byref(req))
    assert rc == ERR_CAPACITY
    want = O.broad_query(m.x, 0, tree, m.facets, 0, margin)
    assert req.value == len(want) > 1
    got = hip.broad_query(ctx, m.x, 0, tree, m.facets, 0, margin, capacity=req.value)
    rows = np.stack([got["query_obj"], got["point"], got["target_obj"], got["triangle"]], 1)
    assert_same_candidates(rows, np.stack([np.zeros(len(want), np.uint32), want[:, 0],
                                           np.zeros(len(want), np.uint32), want[:, 1]], 1))
    # the scene's pair export refuses a short buffer and reports the size
    hs = HipScene(ctx, sc)
    n = C.c_uint32(0)
    small = np.zeros(1, abi.BROAD)
    assert load().ysim_hip_scene_get_pairs(hs.h, small.ctypes.data, 1, C.byref(n)) == ERR_CAPACITY
    assert n.value == len(OracleScene(sc).pairs())
