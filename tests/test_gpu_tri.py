"""TriangularCloth on the GPU: springs from the edge / opposite-vertex lists the
library builds at scene creation, in the same fused integrate kernel as the
FastGrid cloth.  States and contacts equal the oracle's bit for bit."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from ysim_amd import scenes
from ysim_amd._lib import YsimError
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu


def _same_state(a, b):
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


@pytest.mark.parametrize("rebuild_every,mode", [(1, 1), (0, 1), (1, 0)])
def test_tri_drape_bit_exact(ctx, oracle_threads, rebuild_every, mode):
    """A 24-ring disk (1801 vertices, irregular valences) falls onto a UV sphere
    with self collision; compared after every 10 substeps through first contact."""
    print(f"tri drape rebuild_every={rebuild_every} query_tile={mode}", flush=True)
    sc = scenes.config_tri_drape(rebuild_every=rebuild_every)
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_tuning(query_tile=mode)
    for k in range(7):
        hs.step(10, graph=(k % 2 == 0))
        os_.step(10)
        _same_state(hs.state(), os_.state())
    og, cg = hs.contacts()
    oo, co = os_.contacts()
    assert np.array_equal(og, oo) and np.array_equal(cg, co)
    assert len(co) > 50 and hs.flags() == 0


def test_tri_and_grid_cloth_together(ctx, oracle_threads):
    """A FastGrid cloth and a TriangularCloth in one scene (one build group,
    one integrate launch with both force paths) match the oracle."""
    rng = np.random.default_rng(8)
    grid = scenes.cloth_mesh(24, 0.5, (0.1, 0.9, 0.1), rng)
    dx, df = scenes.disk_mesh(10, 0.3, (0.6, 0.85, 0.6), rng)
    tri = scenes.tri_cloth_mesh(dx, df, thickness=0.5 * 0.03)
    sx, sf = scenes.uv_sphere(20, 10, 0.3, (0.5, 0.5, 0.5))
    p = scenes.sim_params(subh=1 / 600, margin=0.01, morton_bits=30, rebuild_every=1)
    sc = scenes.Scene([grid, tri, scenes.obstacle(sx, sf)], p, name="grid+tri")
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.step(40, graph=True)
    os_.step(40)
    _same_state(hs.state(), os_.state())
    assert hs.flags() == 0


def test_tri_cloth_rejections(ctx):
    """A non-manifold triangular cloth is refused at creation; the adjoint
    entry points refuse a TriangularCloth scene."""
    x = np.random.default_rng(0).random((5, 3)).astype(np.float32)
    bad = scenes.tri_cloth_mesh(x, np.array([[0, 1, 2], [0, 1, 3], [0, 1, 4]], np.uint32))
    p = scenes.sim_params(subh=1 / 600, margin=0.01)
    with pytest.raises(YsimError):
        HipScene(ctx, scenes.Scene([bad], p))
    sc = scenes.config_tri_drape(rings=6)
    sc.params["self_collision"] = 0
    hs = HipScene(ctx, sc)
    with pytest.raises(YsimError):
        hs.num_springs()
