"""CPU checks of the candidate audit: the library's record filter
(ysim_amd/csrc/audit.hpp, built for the host in tests/hostcheck) and the
multiset checker the GPU audit tests use.

Records are synthesised from the oracle's candidate set the way the query
kernels write them: tile walks (kind 0, per wave), per-lane walks (kind 1), and
tile walks abandoned part way (some kind-0 records, a retract, then the
per-lane redo of the same target), in a shuffled order as atomics leave them.
The filter must give back the oracle's set exactly; planted mistakes (a lost
record, a retract naming the wrong wave, a repeated record) must fail the
checker."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleScene
from paritylib import assert_same_candidates, cut_cover, rows
from ysim_amd import scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck")
TILE, LANE, RETRACT = 0, 1, 2


@pytest.fixture(scope="module")
def L():
    subprocess.run(["make", "-s"], cwd=HERE, check=True)
    lib = C.CDLL(os.path.join(HERE, "libhosttraverse.so"))
    lib.yt_audit_collect.restype = C.c_uint64
    return lib


@pytest.fixture(scope="module")
def fold_pairs():
    """Candidate set of a 40 x 40 fold (self collision) from the oracle."""
    sc = scenes.config_c2(n=40)
    return OracleScene(sc).pairs()


def collect(L, rec):
    rec = np.ascontiguousarray(rec, np.uint32).reshape(-1, 4)
    nr, nb = C.c_uint32(), C.c_uint32()
    n = L.yt_audit_collect(rec.ctypes.data_as(C.c_void_p), C.c_uint64(rec.shape[0]), None, C.c_uint64(0),
                           C.byref(nr), C.byref(nb))
    out = np.zeros(3 * n, np.uint32)
    L.yt_audit_collect(rec.ctypes.data_as(C.c_void_p), C.c_uint64(rec.shape[0]), out.ctypes.data_as(C.c_void_p),
                       C.c_uint64(n), C.byref(nr), C.byref(nb))
    t = out.reshape(-1, 3)
    got = np.stack([np.zeros(len(t), np.uint32), t[:, 0], t[:, 1], t[:, 2]], 1)
    return got, nr.value, nb.value


def synth_records(pairs, rng, p_redo=0.3):
    """Kernel-style records of a candidate set (one query mesh, item j = point,
    wave = j // 64).  Returns the records and the (wave, target) walks redone."""
    recs = []
    redone = []
    waves = pairs[:, 1] // 64
    for w in np.unique(waves):
        sel = pairs[waves == w]
        for t in np.unique(sel[:, 2]):
            grp = sel[sel[:, 2] == t]
            tag = lambda kind: (kind << 30) | int(w)  # noqa: E731
            if rng.random() < p_redo:
                part = grp[rng.random(len(grp)) < 0.5]
                recs += [(int(r[1]), int(t), int(r[3]), tag(TILE)) for r in part]
                recs.append((0, int(t), 0, tag(RETRACT)))
                recs += [(int(r[1]), int(t), int(r[3]), tag(LANE)) for r in grp]
                redone.append((int(w), int(t)))
            else:
                kind = LANE if rng.random() < 0.1 else TILE
                recs += [(int(r[1]), int(t), int(r[3]), tag(kind)) for r in grp]
    rec = np.asarray(recs, np.uint32)
    return rec[rng.permutation(len(rec))], redone


def test_checker_sees_planted_mistakes(fold_pairs):
    want = fold_pairs
    assert len(want) > 1000
    assert_same_candidates(want[::-1], want)  # order does not matter
    with pytest.raises(AssertionError, match="1 missing"):
        assert_same_candidates(np.delete(want, 17, axis=0), want)
    with pytest.raises(AssertionError, match="1 duplicated"):
        assert_same_candidates(np.concatenate([want, want[5:6]]), want)
    extra = np.array([[0, 0, 0, 1 << 30]], np.uint32)  # no such facet
    with pytest.raises(AssertionError, match="1 extra"):
        assert_same_candidates(np.concatenate([want, extra]), want)


def test_filter_returns_the_candidate_set(L, fold_pairs):
    rng = np.random.default_rng(11)
    rec, redone = synth_records(fold_pairs, rng)
    assert redone, "the synthetic stream has no redone walk"
    got, nr, nb = collect(L, rec)
    assert nr == len(redone) and nb == 0
    assert_same_candidates(got, fold_pairs, "filtered records")
    assert np.array_equal(rows(got), got), "the filter returns rows sorted"


def test_filter_planted_mistakes_fail(L, fold_pairs):
    rng = np.random.default_rng(12)
    rec, redone = synth_records(fold_pairs, rng)
    # a lost record of a walk that was not redone
    kinds = rec[:, 3] >> 30
    keep_walk = np.array([((int(r[3]) & ((1 << 30) - 1)), int(r[1])) not in set(redone) for r in rec])
    i = int(np.flatnonzero((kinds == TILE) & keep_walk)[0])
    got, _, _ = collect(L, np.delete(rec, i, axis=0))
    with pytest.raises(AssertionError, match="1 missing"):
        assert_same_candidates(got, fold_pairs)
    # a retract that names another wave: the abandoned walk's records stay
    M = (1 << 30) - 1

    def partial(j):
        return int(np.count_nonzero((kinds == TILE) & (rec[:, 1] == rec[j, 1]) & ((rec[:, 3] & M) == (rec[j, 3] & M))))

    j = next(int(k) for k in np.flatnonzero(kinds == RETRACT) if partial(int(k)) > 0)
    bad = rec.copy()
    bad[j, 3] = (RETRACT << 30) | (((int(bad[j, 3]) & M) + 1) & M)
    got, _, _ = collect(L, bad)
    with pytest.raises(AssertionError, match=f"{partial(j)} duplicated"):
        assert_same_candidates(got, fold_pairs)
    # a record reported twice
    got, _, _ = collect(L, np.concatenate([rec, rec[i:i + 1]]))
    with pytest.raises(AssertionError):
        assert_same_candidates(got, fold_pairs)
    # an unknown kind is counted
    odd = rec.copy()
    odd[0, 3] = (3 << 30) | 5
    _, _, nb = collect(L, odd)
    assert nb == 1


def test_cut_cover_on_oracle_tree():
    """cut_cover (the GPU test's cut check) on the oracle's own tree: the
    children of the root cover every leaf once; a node and its child together
    cover some leaves twice; a missing subtree leaves zeros."""
    sc = scenes.config_c2(n=24)
    nodes, _, _, _ = OracleScene(sc).tree(0)
    a, b = int(nodes["child_a"][0]), int(nodes["child_b"][0])
    assert np.all(cut_cover(nodes, [a, b]) == 1)
    assert np.all(cut_cover(nodes, [0]) == 1)
    assert np.any(cut_cover(nodes, [0, a]) == 2)
    assert np.any(cut_cover(nodes, [a]) == 0)


def test_heavy_tile_leaves_known_case():
    """A flat 16 x 16 grid (unit steps, margin 0.5) with its own cell
    triangles and one planted triangle over rows and columns 0-10: the
    planted leaf crosses 64, 24, 24 and 9 row x column crossings of the four
    tiles, so exactly three (tile, leaf) pairs pass the 16 threshold; the
    cell triangles cross at most 3 x 3."""
    from paritylib import heavy_tile_leaves
    n = 16
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = np.stack([c, r, np.zeros_like(c)], -1).astype(np.float32).reshape(-1, 3)
    cells = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, d, e = i * n + j, i * n + j + 1, (i + 1) * n + j, (i + 1) * n + j + 1
            cells += [[a, b, e], [a, e, d]]
    cells = np.array(cells, np.uint32)
    v = np.zeros_like(x)
    assert heavy_tile_leaves(x, v, cells, n, 0.5, 0.0) == 0
    planted = np.vstack([cells, [[0, 10, 10 * n + 10]]]).astype(np.uint32)
    assert heavy_tile_leaves(x, v, planted, n, 0.5, 0.0) == 3
    # a sweep along +x by 2 steps reaches column 12: 8 x 5 in the right tiles
    v[[0, 10, 10 * n + 10], 0] = 2.0
    assert heavy_tile_leaves(x, v, planted, n, 0.5, 1.0) == 3
