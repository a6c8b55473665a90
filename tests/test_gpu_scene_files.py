"""The scene files the reference ships, loaded by the scene-file reader and
run on the GPU: states and contacts equal the oracle's bit for bit (a
TriangularCloth falling past an analytic box onto a ground grid; a
TriangularCloth draped over a Rigid sphere answered as an analytic
collider; a 128 x 128 TriangularCloth in free fall)."""
import os

import numpy as np
import pytest

from oracle.oracle import OracleScene
from ysim_amd import scene_io
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def _same_state(a, b):
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


@pytest.mark.parametrize("name,frames,every", [("synthetic.json", 60, 6), ("cloth_drape.json", 30, 3),
                                               ("subobject_cloth.json", 3, 1)])
def test_scene_file_bit_exact(ctx, oracle_threads, name, frames, every):
    print(f"scene file {name}: {frames} frames", flush=True)
    sc = scene_io.load_scene(os.path.join(GOLD, name))
    n = sc.info["substeps"]
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    for k in range(0, frames, every):
        hs.step(every * n, graph=(k % 2 == 0))
        os_.step(every * n)
        _same_state(hs.state(), os_.state())
    og, cg = hs.contacts()
    oo, co = os_.contacts()
    assert np.array_equal(og, oo) and np.array_equal(cg, co)
    assert hs.flags() == 0
    if name != "subobject_cloth.json":
        assert len(co) > 0
