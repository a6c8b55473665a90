"""Batched ragged instances: one scene of many cloth+sphere instances advanced
per launch, bit-identical to running each instance as its own plain scene."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from ysim_amd import scenes
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu


def cloth_slices(sc):
    vo, _ = sc.offsets()
    return [(int(vo[2 * i]), int(vo[2 * i + 1])) for i in range(len(sc.meshes) // 2)]


@pytest.mark.parametrize("graph", [False, True])
def test_batched_equals_sequential_plain(ctx, graph):
    sc = scenes.config_c3(count=24, lo=16, hi=64)
    b = HipScene(ctx, sc)
    b.step(30, graph=graph)
    xb, vb, _ = b.state()
    assert b.flags() == 0
    for i, (s0, s1) in enumerate(cloth_slices(sc)):
        p = HipScene(ctx, scenes.instance_scene(sc, i))
        p.step(30, graph=graph)
        xp, vp, _ = p.state()
        n = s1 - s0
        assert np.array_equal(xb[s0:s1], xp[:n]), i
        assert np.array_equal(vb[s0:s1], vp[:n]), i


def test_batched_matches_oracle(ctx, oracle_threads):
    sc = scenes.config_c3(count=8, lo=16, hi=48)
    b = HipScene(ctx, sc)
    o = OracleScene(sc)
    for _ in range(4):
        b.step(10, graph=True)
        o.step(10)
        assert np.array_equal(b.state()[0], o.state()[0])
    og, cg = b.contacts()
    oo, co = o.contacts()
    assert np.array_equal(og, oo) and len(cg) == len(co)


def test_batched_full_runs_clean(ctx):
    sc = scenes.config_c3()  # 256 instances, ~3.0 M vertices, ~6.0 M triangles
    b = HipScene(ctx, sc)
    b.step(50, graph=True)
    assert b.flags() == 0
    x, v, _ = b.state()
    assert np.all(np.isfinite(x))
