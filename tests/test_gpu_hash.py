"""Spatial-hash broad phase: identical candidate set to the LBVH (and the oracle),
identical dynamics."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from ysim_amd import scenes
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu


def keyset(p):
    p = np.asarray(p, np.uint64)
    k = (p[:, 0] << np.uint64(52)) | (p[:, 1] << np.uint64(31)) | (p[:, 2] << np.uint64(24)) | p[:, 3]
    return np.sort(k)


def hash_variant(sc, slots=22, cell=0.0):
    sc.params = dict(sc.params, broad_mode=1, hash_slots_log2=slots, hash_cell=cell)
    return sc


@pytest.mark.parametrize("slots", [12, 22])
def test_drape_hash_equals_lbvh_and_oracle(ctx, oracle_threads, slots):
    base = scenes.config_drape(n=32, height=0.8, rebuild_every=1)
    hs_h = HipScene(ctx, hash_variant(scenes.config_drape(n=32, height=0.8, rebuild_every=1), slots))
    hs_l = HipScene(ctx, base)
    os_ = OracleScene(base)
    for _ in range(6):
        ph, pl, po = keyset(hs_h.pairs()), keyset(hs_l.pairs()), keyset(os_.pairs())
        assert np.array_equal(ph, pl) and np.array_equal(pl, po)
        hs_h.step(10, graph=True)
        hs_l.step(10)
        os_.step(10)
        xh, vh, _ = hs_h.state()
        xo, vo, _ = os_.state()
        assert np.array_equal(xh, xo) and np.array_equal(vh, vo)
    assert hs_h.flags() == 0 and len(os_.contacts()[1]) > 0


def test_c5_small_hash_equals_lbvh(ctx, oracle_threads):
    lb = scenes.config_c5(n=256, level=4, height=0.93)
    hs_l = HipScene(ctx, lb)
    hs_h = HipScene(ctx, hash_variant(scenes.config_c5(n=256, level=4, height=0.93), cell=lb.params["hash_cell"]))
    os_ = OracleScene(lb)
    for _ in range(3):
        pl, ph, po = keyset(hs_l.pairs()), keyset(hs_h.pairs()), keyset(os_.pairs())
        assert np.array_equal(pl, ph) and np.array_equal(pl, po)
        hs_l.step(4)
        hs_h.step(4)
        os_.step(4)
    assert np.array_equal(hs_l.state()[0], hs_h.state()[0])
    assert np.array_equal(hs_h.state()[0], os_.state()[0])
    assert hs_h.flags() == 0


def test_c5_full_hash_equals_lbvh(ctx):
    """1024^2 cloth + 4 level-6 icospheres, 2^22 slots, cell = 2 * thickness."""
    lb = scenes.config_c5()
    hs_l = HipScene(ctx, lb)
    pl = keyset(hs_l.pairs())
    del hs_l
    hs_h = HipScene(ctx, hash_variant(scenes.config_c5(), slots=22, cell=lb.params["hash_cell"]))
    ph = keyset(hs_h.pairs())
    assert len(pl) > 2_000_000
    assert np.array_equal(pl, ph)
    hs_h.step(4, graph=True)
    assert hs_h.flags() == 0
