"""CPU check of the query kernel's traversal logic (ysim_amd/csrc/traverse.hpp
compiled for the host): the short-stack walk, including its stackless
continuation (forced with a 2-entry stack), returns exactly the oracle's
candidate set on a 128 x 128 fold and on a drape over a sphere."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck")


@pytest.fixture(scope="module")
def L():
    subprocess.run(["make", "-s"], cwd=HERE, check=True)
    return C.CDLL(os.path.join(HERE, "libhosttraverse.so"))


def host_pairs(L, xq, tree, parent, facets, same, margin, small):
    T = facets.shape[0]
    cnt, it, mx = C.c_uint64(), C.c_uint64(), C.c_uint32()
    args = [xq.ctypes.data_as(C.c_void_p), C.c_uint32(len(xq)), 0, tree.ctypes.data_as(C.c_void_p),
            parent.ctypes.data_as(C.c_void_p), C.c_uint32(T), facets.ctypes.data_as(C.c_void_p),
            0 if same else 1, C.c_float(margin), small]
    assert L.yt_query(*args, None, C.c_uint64(0), C.byref(cnt), C.byref(it), C.byref(mx)) == 0
    out = np.zeros(2 * cnt.value, np.uint32)
    assert L.yt_query(*args, out.ctypes.data_as(C.c_void_p), C.c_uint64(cnt.value), C.byref(cnt), C.byref(it),
                      C.byref(mx)) == 0
    return out.reshape(-1, 2), it.value / len(xq)


def test_tile_order_is_a_bijection(L):
    """The fused query visits grid vertices in 8x8 tiles; every vertex exactly once."""
    for N in (3, 7, 8, 9, 17, 31, 33, 100, 128, 512, 1024):
        out = np.zeros(N * N, np.uint32)
        L.yt_tile_order(C.c_uint32(N), out.ctypes.data_as(C.c_void_p))
        assert np.array_equal(np.sort(out), np.arange(N * N, dtype=np.uint32)), N
    out = np.zeros(256, np.uint32)
    L.yt_tile_order(C.c_uint32(16), out.ctypes.data_as(C.c_void_p))
    assert out[:8].tolist() == list(range(8)) and out[8] == 16  # first tile, second row


@pytest.mark.parametrize("small", [0, 1])
def test_fold_self_query(L, small):
    sc = scenes.config_c2(n=128)
    S = O.OracleScene(sc)
    S.step(2)
    nodes, parent, _, _ = S.tree(0)
    x, _, _ = S.state()
    m = sc.meshes[0]
    got, iters = host_pairs(L, x, nodes, parent, m.facets, True, sc.params["margin"], small)
    exp = O.broad_query(x, 0, nodes, m.facets, 0, sc.params["margin"])
    assert set(map(tuple, got.tolist())) == set(map(tuple, exp.tolist()))
    assert len(got) == len(exp)
    assert iters < 200


@pytest.mark.parametrize("small", [0, 1])
def test_drape_sphere_query(L, small):
    sc = scenes.config_drape(n=32, height=0.79)
    S = O.OracleScene(sc)
    S.step(20)
    x, _, _ = S.state()
    cl, sp = sc.meshes
    nodes, parent, _, _ = S.tree(1)
    xc = np.ascontiguousarray(x[: cl.x.shape[0]])
    got, _ = host_pairs(L, xc, nodes, parent, sp.facets, False, sc.params["margin"], small)
    exp = O.broad_query(xc, 0, nodes, sp.facets, 1, sc.params["margin"])
    assert set(map(tuple, got.tolist())) == set(map(tuple, exp.tolist()))


@pytest.mark.parametrize("corrupt", ["self_loop", "out_of_range", "two_cycle"])
def test_corrupt_parent_links_fire_the_guard(L, corrupt):
    """The walk's stackless continuation climbs through parent links (a
    2-entry stack forces it).  Corrupted links must end the walk through its
    guard (return -1) rather than loop or read out of range; checked on the
    host build only."""
    sc = scenes.config_c2(n=32)
    S = O.OracleScene(sc)
    S.step(1)
    nodes, parent, _, _ = S.tree(0)
    x, _, _ = S.state()
    m = sc.meshes[0]
    T = m.facets.shape[0]
    _, iters = host_pairs(L, x, nodes, parent, m.facets, True, 4 * sc.params["margin"], 1)
    bad = parent.copy()
    internal = np.arange(1, T - 1)
    if corrupt == "self_loop":
        bad[internal] = internal
    elif corrupt == "out_of_range":
        bad[internal] = 10 ** 6
    else:  # every internal node's parent is its own child_a
        bad[internal] = nodes["child_a"][internal]
    args = [x.ctypes.data_as(C.c_void_p), C.c_uint32(len(x)), 0, nodes.ctypes.data_as(C.c_void_p),
            bad.ctypes.data_as(C.c_void_p), C.c_uint32(T), m.facets.ctypes.data_as(C.c_void_p), 0,
            C.c_float(4 * sc.params["margin"]), 1]
    cnt, it, mx = C.c_uint64(), C.c_uint64(), C.c_uint32()
    assert L.yt_query(*args, None, C.c_uint64(0), C.byref(cnt), C.byref(it), C.byref(mx)) == -1
