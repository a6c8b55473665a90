"""Pins of the oracle's TriangularCloth restatement (MeshAdjacencyInitializer,
include/sim/mesh_state.hpp:236-454; compute_tri_spring_forces,
src/metal/physics.metal:319-378) against independent constructions: the edge
and opposite-vertex lists against a dictionary build from the facets, rest
lengths against numpy, forces against a per-spring FP64 sum written from the
spring law, and the pairwise cancellation of spring forces."""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import scenes


def _meshes():
    rng = np.random.default_rng(11)
    dx, df = scenes.disk_mesh(4, 0.5, (0.1, 0.7, -0.2), rng)
    ix, i_f = scenes.icosphere(1, 0.3, (0.0, 0.0, 0.0))
    ux, uf = scenes.uv_sphere(8, 5, 0.4, (0.5, 0.5, 0.5))
    return [("disk", dx, df), ("icosphere", ix, i_f), ("uv_sphere", ux, uf)]


def _norm32(a, b):
    d = (b - a).astype(np.float32)
    return np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32)


def _brute_adjacency(F, x):
    """Edges sorted by (min, max); per vertex its edges in edge order; an edge
    of two facets makes its opposite vertices list each other, in edge order."""
    opp = {}
    for f in F:
        for k in range(3):
            a, b, c = int(f[k]), int(f[(k + 1) % 3]), int(f[(k + 2) % 3])
            opp.setdefault((min(a, b), max(a, b)), []).append(c)
    keys = sorted(opp)
    nV = x.shape[0]
    adj = [[] for _ in range(nV)]
    oppv = [[] for _ in range(nV)]
    for e, (a, b) in enumerate(keys):
        adj[a].append(e)
        adj[b].append(e)
        if len(opp[(a, b)]) == 2:
            o0, o1 = opp[(a, b)]
            oppv[o0].append(o1)
            oppv[o1].append(o0)
    return np.array(keys, np.uint32), adj, oppv


@pytest.mark.parametrize("name,x,F", _meshes())
def test_adjacency_matches_brute_force(name, x, F):
    A = O.tri_adjacency(F, x)
    edges, adj, oppv = _brute_adjacency(F, x)
    assert np.array_equal(A["edges"], edges)
    assert np.array_equal(A["rest_edge"], _norm32(x[edges[:, 0]], x[edges[:, 1]]))
    for i in range(x.shape[0]):
        assert list(A["adj"][A["adj_off"][i]:A["adj_off"][i + 1]]) == adj[i]
        got = A["opp"][A["opp_off"][i]:A["opp_off"][i + 1]]
        assert list(got) == oppv[i]
        rest = A["rest_opp"][A["opp_off"][i]:A["opp_off"][i + 1]]
        if len(got):
            assert np.array_equal(rest, _norm32(np.repeat(x[i:i + 1], len(got), 0), x[got]))
    # closed surfaces: every edge interior, E = 3F/2; the disk has 6R boundary edges
    nE = edges.shape[0]
    if name != "disk":
        assert nE * 2 == 3 * F.shape[0] and A["opp"].size == 2 * nE
    else:
        assert A["opp"].size == 2 * (nE - 6 * 4)


def test_rejects_degenerate_and_nonmanifold():
    x = np.random.default_rng(0).random((5, 3)).astype(np.float32)
    assert O.tri_adjacency(np.array([[0, 1, 1]], np.uint32), x) is None
    fan3 = np.array([[0, 1, 2], [0, 1, 3], [0, 1, 4]], np.uint32)  # one edge, three facets
    assert O.tri_adjacency(fan3, x) is None
    assert O.tri_adjacency(fan3[:2], x) is not None


def _spring_sum64(F, x0, x, v, ext, ks, kb, kair, kd):
    """FP64 forces from the spring law, one spring at a time (independent of
    the oracle's loops): f = ext + kair v + sum_k (k (|d| - r) + kd dv.n) n."""
    edges, adj, oppv = _brute_adjacency(F, x0)
    x0d, xd, vd = (np.asarray(a, np.float64) for a in (x0, x, v))
    f = np.asarray(ext, np.float64) + kair * vd
    def add(i, j, rest, k):
        d = xd[j] - xd[i]
        L = np.linalg.norm(d)
        n = d / L
        f[i] += (k * (L - rest) + kd * np.dot(vd[j] - vd[i], n)) * n
    for i in range(x.shape[0]):
        for e in adj[i]:
            a, b = edges[e]
            add(i, int(b if a == i else a), float(np.float32(_norm32(x0[a:a + 1], x0[b:b + 1])[0])), ks)
        for o in oppv[i]:
            add(i, o, float(_norm32(x0[i:i + 1], x0[o:o + 1])[0]), kb)
    return f


@pytest.mark.parametrize("name,x0,F", _meshes())
def test_forces_match_spring_law(name, x0, F):
    rng = np.random.default_rng(5)
    x = (x0 + rng.normal(0, 0.01, x0.shape)).astype(np.float32)
    v = rng.normal(0, 0.1, x0.shape).astype(np.float32)
    ext = np.tile(np.array([0, -9.81, 0], np.float32), (x0.shape[0], 1))
    ks, kb, kair, kd = 500.0, 5.0, -0.001, 0.5
    ref = _spring_sum64(F, x0, x, v, ext, ks, kb, kair, kd)
    f64 = O.tri_forces(F, x0, x, v, ext, ks, kb, kair, kd, f64=True)
    f32 = O.tri_forces(F, x0, x, v, ext, np.float32(ks), np.float32(kb), np.float32(kair), np.float32(kd))
    scale = np.abs(ref).max()
    assert np.abs(f64 - ref).max() <= 1e-12 * scale
    assert np.abs(f32 - ref).max() <= 2e-5 * scale


def test_spring_forces_cancel_in_pairs():
    """Each spring pushes its two ends with exactly opposite forces, so the
    internal forces sum to zero (FP64, kd > 0, no external force)."""
    x0, F = scenes.disk_mesh(5, 0.5, (0.0, 0.0, 0.0), np.random.default_rng(2))
    rng = np.random.default_rng(3)
    x = x0 + rng.normal(0, 0.02, x0.shape)
    v = rng.normal(0, 0.3, x0.shape)
    f = O.tri_forces(F, x0, x, v, np.zeros_like(x), 800.0, 20.0, 0.0, 2.0, f64=True)
    assert np.abs(f.sum(axis=0)).max() <= 1e-10 * np.abs(f).max()


def test_planted_bend_stiffness_is_detected():
    """Swapping k_stretch and k_bend (a plausible wiring mistake) changes the
    forces far beyond rounding: the pin above is sensitive to which family
    gets which constant."""
    x0, F = scenes.disk_mesh(3, 0.5, (0.0, 0.0, 0.0), np.random.default_rng(4))
    x = (x0 * np.float32(1.01)).astype(np.float32)
    v = np.zeros_like(x0)
    ext = np.zeros_like(x0)
    ref = _spring_sum64(F, x0, x, v, ext, 500.0, 5.0, 0.0, 0.0)
    swapped = O.tri_forces(F, x0, x, v, ext, 5.0, 500.0, 0.0, 0.0, f64=True)
    assert np.abs(swapped - ref).max() > 1e-3 * np.abs(ref).max()


def test_scene_twin_and_rejection():
    """A TriangularCloth scene steps in the oracle with its FP64 twin close to
    the FP32 path; a non-manifold cloth is rejected at creation."""
    sc = scenes.config_tri_drape(rings=8)
    os_ = O.OracleScene(sc, twin=True)
    os_.step(10)
    x, v, _ = os_.state()
    xd, vd = os_.twin_state()
    assert np.isfinite(x).all() and np.abs(x - xd).max() < 1e-5
    bad = scenes.tri_cloth_mesh(np.random.default_rng(0).random((5, 3)).astype(np.float32),
                                np.array([[0, 1, 2], [0, 1, 3], [0, 1, 4]], np.uint32))
    with pytest.raises(ValueError):
        O.OracleScene(scenes.Scene([bad], sc.params))
