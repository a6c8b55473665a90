"""Pins of the oracle's analytic-collider restatement (narrow_pt_analytic,
src/metal/bruteforce.metal:198-499; the rotation bands of
include/sim/bruteforce.hpp:201-242 and include/sim/bvh.hpp:2431-2518; the
marker carry-over of include/sim/scene.hpp:511-529) against closed forms:
distances to planes, spheres, boxes and cylinders, the first crossing of a
segment with an inflated surface found by bisection, rotated boxes' bounding
half-extents, and the geodesic rotation angle of a quaternion pair."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import scenes
from ysim_amd.scenes import (COLLIDER_BOX, COLLIDER_CYLINDER, COLLIDER_PLANE, COLLIDER_SPHERE, analytic_shape,
                             quat_apply, quat_axis_angle)

TOL = 2e-6


def _narrow(sh, p, pprev=None, band=0.05):
    return O.narrow_analytic(sh, p, p if pprev is None else pprev, band)


def _rot(q, v):
    return quat_apply(q, np.asarray(v, np.float64)[None, :])[0]


def test_plane_distance_and_normal():
    q = quat_axis_angle((0.3, 0.0, 1.0), 0.4)
    c = np.array([0.1, 0.2, -0.3])
    sh = analytic_shape(COLLIDER_PLANE, 0, c, rot=q)
    n_w = _rot(q, (0, 1, 0))
    rng = np.random.default_rng(0)
    for _ in range(50):
        p = c + rng.normal(0, 0.5, 3)
        dist = float(np.dot(n_w, p - c))
        r = _narrow(sh, p, band=0.05)
        if dist < 0.05 - 1e-5:
            assert r is not None
            assert np.abs(r[0] - n_w).max() < TOL and abs(r[1] - dist) < 1e-5
        elif dist > 0.05 + 1e-5:
            assert r is None
    # a half-space has no far side: deep below still pushes along +Y
    r = _narrow(sh, c - 3.0 * n_w, band=0.05)
    assert r is not None and np.abs(r[0] - n_w).max() < TOL and abs(r[1] + 3.0) < 1e-5


def test_sphere_static_distance():
    c, rad = np.array([0.5, 0.5, 0.5]), 0.3
    sh = analytic_shape(COLLIDER_SPHERE, 0, c, half=(rad, rad, rad))
    rng = np.random.default_rng(1)
    for _ in range(100):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        delta = rng.uniform(-0.1, 0.1)
        p = c + (rad + delta) * u
        r = _narrow(sh, p, band=0.04)
        if delta < 0.04 - 1e-5:
            assert r is not None, delta
            assert np.abs(r[0] - u).max() < 1e-5 and abs(r[1] - delta) < 1e-5
        elif delta > 0.04 + 1e-5:
            assert r is None


def _first_crossing(a, b, inside):
    """Smallest t in [0, 1] with inside(a + t (b - a)), by dense scan + bisection."""
    ts = np.linspace(0, 1, 4001)
    hit = [t for t in ts if inside(a + t * (b - a))]
    if not hit:
        return None
    hi = hit[0]
    lo = max(0.0, hi - 1 / 4000)
    for _ in range(60):
        m = 0.5 * (lo + hi)
        if inside(a + m * (b - a)):
            hi = m
        else:
            lo = m
    return hi


def test_sphere_entry_side_rule():
    """A vertex that tunnels from above to just inside the far (bottom) side
    is pushed back out through the top (the entry), not out the bottom."""
    c, rad, band = np.zeros(3), 0.2, 0.02
    sh = analytic_shape(COLLIDER_SPHERE, 0, c, half=(rad, rad, rad))
    a = np.array([0.03, 0.4, 0.01])
    b = np.array([0.02, -0.17, 0.0])
    t = _first_crossing(a, b, lambda p: np.linalg.norm(p) <= rad + band)
    u = a + t * (b - a)
    n_exp = u / np.linalg.norm(u)
    d_exp = float(np.dot(n_exp, b - rad * n_exp))
    n, d = _narrow(sh, b, a, band)
    assert n[1] > 0.9 and np.abs(n - n_exp).max() < 1e-5 and abs(d - d_exp) < 1e-5
    assert d < -0.3
    # started inside: the direction of the start point decides, not the end point
    a_in = np.array([0.0, 0.1, 0.0])
    n2, d2 = _narrow(sh, np.array([0.0, -0.15, 0.0]), a_in, band)
    assert n2[1] > 0.999 and abs(d2 - (-0.15 - 0.2)) < 1e-5


def test_ellipsoid_normal_is_gradient():
    c, e = np.array([0.1, 0.0, -0.2]), np.array([0.3, 0.15, 0.2])
    q = quat_axis_angle((1.0, 1.0, 0.0), 0.7)
    sh = analytic_shape(COLLIDER_SPHERE, 0, c, half=e, rot=q)
    rng = np.random.default_rng(2)
    for _ in range(40):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        loc = e * u * (1.0 + rng.uniform(0.0, 0.05))   # just outside, local frame
        p = c + _rot(q, loc)
        r = _narrow(sh, p, band=0.05)
        assert r is not None
        uh = (loc / e) / np.linalg.norm(loc / e)
        s = e * uh
        n_loc = (uh / e) / np.linalg.norm(uh / e)
        assert np.abs(r[0] - _rot(q, n_loc)).max() < 1e-5
        assert abs(r[1] - float(np.dot(n_loc, loc - s))) < 1e-5
        assert r[1] >= -1e-6


def test_box_faces_entry_and_inside():
    c, h = np.array([0.4, 0.3, 0.2]), np.array([0.2, 0.1, 0.15])
    q = quat_axis_angle((0.2, 1.0, 0.4), 0.5)
    sh = analytic_shape(COLLIDER_BOX, 0, c, half=h, rot=q)
    band = 0.03
    W = lambda v: c + _rot(q, v)
    # resting 0.01 off the +X face
    n, d = _narrow(sh, W((0.21, 0.02, -0.03)), band=band)
    assert np.abs(n - _rot(q, (1, 0, 0))).max() < TOL and abs(d - 0.01) < 1e-5
    # beyond the band: no contact
    assert _narrow(sh, W((0.0, 0.14, 0.0)), band=band) is None
    # falling through the top face in one step: entry face +Y, deep negative d
    n, d = _narrow(sh, W((0.05, -0.08, 0.0)), W((0.05, 0.2, 0.0)), band)
    assert np.abs(n - _rot(q, (0, 1, 0))).max() < TOL and abs(d - (-0.08 - 0.1)) < 1e-5
    # started inside near the -Z face: minimum-penetration axis of the start
    n, d = _narrow(sh, W((0.0, 0.0, 0.05)), W((0.0, 0.0, -0.14)), band)
    assert np.abs(n - _rot(q, (0, 0, -1))).max() < TOL and abs(d - (-0.05 - 0.15)) < 1e-5


def test_cylinder_wall_cap_and_entry():
    c, rr, hh = np.array([0.0, 0.5, 0.0]), 0.1, 0.2
    q = quat_axis_angle((1.0, 0.0, 0.0), 0.3)
    sh = analytic_shape(COLLIDER_CYLINDER, 0, c, radius=rr, half_height=hh, rot=q)
    W = lambda v: c + _rot(q, v)
    band = 0.02
    ang = 0.7
    radial = np.array([math.cos(ang), 0.0, math.sin(ang)])
    n, d = _narrow(sh, W(radial * 0.11 + [0, 0.05, 0]), band=band)
    assert np.abs(n - _rot(q, radial)).max() < 1e-5 and abs(d - 0.01) < 1e-5
    n, d = _narrow(sh, W((0.02, 0.21, -0.01)), band=band)
    assert np.abs(n - _rot(q, (0, 1, 0))).max() < TOL and abs(d - 0.01) < 1e-5
    n, d = _narrow(sh, W((0.0, -0.215, 0.03)), band=band)
    assert np.abs(n - _rot(q, (0, -1, 0))).max() < TOL and abs(d - 0.015) < 1e-5
    # sideways through the wall: entry on the wall, normal toward the start
    a, b = np.array([0.3, 0.0, 0.05]), np.array([-0.05, 0.0, 0.05])
    t = _first_crossing(a, b, lambda p: math.hypot(p[0], p[2]) <= rr + band and abs(p[1]) <= hh + band)
    hit = a + t * (b - a)
    nxz = np.array([hit[0], 0.0, hit[2]]) / math.hypot(hit[0], hit[2])
    n, d = _narrow(sh, W(b), W(a), band)
    assert np.abs(n - _rot(q, nxz)).max() < 1e-5 and abs(d - (float(np.dot(nxz, b)) - rr)) < 1e-5


def test_motion_inversion_moving_box():
    """A box rising into a resting vertex: in the box frame the vertex crossed
    the top face, so the contact normal is +Y and d is the penetration."""
    h = np.array([0.1, 0.05, 0.1])
    sh = analytic_shape(COLLIDER_BOX, 0, (0.0, 0.02, 0.0), half=h, prev_center=(0.0, -0.1, 0.0))
    p = np.array([0.01, 0.06, 0.0])
    n, d = _narrow(sh, p, p, 0.01)
    assert np.abs(n - (0, 1, 0)).max() < TOL and abs(d - (0.04 - 0.05)) < 1e-6
    # a static box sees the same vertex as resting inside near the top: same answer
    n2, d2 = _narrow(analytic_shape(COLLIDER_BOX, 0, (0.0, 0.02, 0.0), half=h), p, p, 0.01)
    assert np.array_equal(n, n2) and d == d2


@pytest.mark.parametrize("kind,half,radius,hh,rmax", [
    (COLLIDER_SPHERE, (0.3, 0.1, 0.2), 0.0, 0.0, 0.3),
    (COLLIDER_BOX, (0.3, 0.1, 0.2), 0.0, 0.0, math.sqrt(0.09 + 0.01 + 0.04)),
    (COLLIDER_CYLINDER, (0.0, 0.0, 0.0), 0.15, 0.25, math.sqrt(0.15 ** 2 + 0.25 ** 2)),
    (COLLIDER_PLANE, (0.0, 0.0, 0.0), 0.0, 0.0, 0.0),
])
def test_rotation_margins(kind, half, radius, hh, rmax):
    phi = 0.3   # float32 quaternions: acos near 1 loses digits for small angles
    sh = analytic_shape(kind, 0, (0, 0, 0), half=half, radius=radius, half_height=hh,
                        rot=quat_axis_angle((0, 1, 0), phi), prev_rot=(1.0, 0.0, 0.0, 0.0))
    lo, hi, rot_band, rot_margin, (theta, rm) = O.analytic_bounds(sh, 0.001, 0.01)
    h = np.array(half, np.float64)
    # broad: theta (|h| + radius) for every kind
    assert abs(rot_band - phi * (np.linalg.norm(h) + radius)) < 1e-6
    if kind == COLLIDER_PLANE:
        assert rot_margin == 0.0 and theta == 0.0
    else:
        assert abs(theta - phi) < 1e-6 and abs(rm - rmax) < 1e-7
        assert abs(rot_margin - phi * (rmax + 0.011)) < 1e-6
    # no rotation: both vanish
    sh0 = analytic_shape(kind, 0, (0, 0, 0), half=half, radius=radius, half_height=hh)
    assert O.analytic_bounds(sh0, 0.001, 0.01)[2:4] == (0.0, 0.0)


def test_bounding_half_extents():
    h = np.array([0.3, 0.1, 0.2])
    s45 = analytic_shape(COLLIDER_BOX, 0, (1.0, 2.0, 3.0), half=h, rot=quat_axis_angle((0, 1, 0), math.pi / 4))
    lo, hi = O.analytic_bounds(s45)[:2]
    w = (h[0] + h[2]) / math.sqrt(2)
    assert np.abs(hi - lo - 2 * np.array([w, h[1], w])).max() < 1e-5
    cyl = analytic_shape(COLLIDER_CYLINDER, 0, (0, 0, 0), radius=0.1, half_height=0.3,
                         rot=quat_axis_angle((0, 0, 1), math.pi / 2))
    lo, hi = O.analytic_bounds(cyl)[:2]
    assert np.abs(hi - np.array([0.3, 0.1, 0.1])).max() < 1e-6 and np.abs(lo + hi).max() < 1e-6
    ell = analytic_shape(COLLIDER_SPHERE, 0, (0, 0, 0), half=(0.1, 0.4, 0.2), rot=quat_axis_angle((1, 2, 3), 1.0))
    lo, hi = O.analytic_bounds(ell)[:2]
    assert np.array_equal(hi, np.full(3, 0.4, np.float32))
    # swept: the union of the previous and current boxes
    mv = analytic_shape(COLLIDER_BOX, 0, (0.5, 0.0, 0.0), half=h, prev_center=(-0.5, 0.1, 0.0))
    lo, hi = O.analytic_bounds(mv)[:2]
    assert np.allclose(lo, [-0.8, -0.1, -0.2], atol=1e-6) and np.allclose(hi, [0.8, 0.2, 0.2], atol=1e-6)


def test_f32_against_f64():
    rng = np.random.default_rng(9)
    shapes = [analytic_shape(COLLIDER_SPHERE, 0, (0, 0, 0), half=(0.2, 0.3, 0.25), rot=quat_axis_angle((1, 0, 1), 0.3)),
              analytic_shape(COLLIDER_BOX, 0, (0, 0, 0), half=(0.2, 0.3, 0.25), rot=quat_axis_angle((0, 1, 1), 0.8)),
              analytic_shape(COLLIDER_CYLINDER, 0, (0, 0, 0), radius=0.2, half_height=0.3),
              analytic_shape(COLLIDER_PLANE, 0, (0, 0, 0), rot=quat_axis_angle((1, 0, 0), 0.2))]
    agree = 0
    for sh in shapes:
        for _ in range(300):
            p = rng.normal(0, 0.25, 3).astype(np.float32)
            pp = (p + rng.normal(0, 0.02, 3)).astype(np.float32)
            r32 = O.narrow_analytic(sh, p, pp, 0.02)
            r64 = O.narrow_analytic(sh, p.astype(np.float64), pp.astype(np.float64), 0.02, f64=True)
            if (r32 is None) != (r64 is None):
                continue   # a gate tie
            if r32 is None:
                continue
            agree += 1
            assert np.abs(r32[0] - r64[0]).max() < 1e-5 and abs(r32[1] - r64[1]) < 1e-5
    assert agree > 300


def _box_scene(shape_flags=1):
    rng = np.random.default_rng(3)
    cl = scenes.cloth_mesh(8, 0.1, (-0.05, 0.3, -0.05), rng)
    h = (0.1, 0.1, 0.1)
    bx = scenes.collider_mesh(*scenes.box_mesh(h, (0.0, 0.0, 0.0)), COLLIDER_BOX)
    p = scenes.sim_params(subh=1 / 600, margin=0.01, rebuild_every=1)
    sh = analytic_shape(COLLIDER_BOX, 1, (0.0, 0.0, 0.0), half=h, flags=shape_flags)
    return scenes.Scene([cl, bx], p, analytic=sh)


def test_markers_and_carry_over():
    """The pair is marked while the cloth's root box meets the padded box, and
    runs one substep longer: a cloth moved far away along x whose previous
    positions are inside the box (nearest the +Y face) still gets contacts
    from that face on the substep after the move, and none on the next."""
    sc = _box_scene()
    o = O.OracleScene(sc)
    o.step(1)
    pairs, marked = o.analytic_markers()
    assert pairs.tolist() == [[0, 1]] and not marked[0]   # 0.3 above, margin 0.01
    x, v, _ = o.state()
    inside = x.copy()
    inside[:64, 1] = 0.07
    o.set_state(x=inside, xprev=inside)
    o.step(1)
    assert o.analytic_markers()[1][0]
    far = inside.copy()
    far[:64, 0] += 5.0
    o.set_state(x=far, v=np.zeros_like(far), xprev=inside)
    o.step(1)
    off, rec = o.contacts()
    assert not o.analytic_markers()[1][0]
    assert len(rec) == 64 and set(rec["target_obj"]) == {1} and set(rec["triangle"]) == {0}
    assert np.all(np.abs(rec["normal_distance"][:, 1] - 1.0) < 1e-6)   # +Y face from the start points
    assert np.all(np.abs(rec["normal_distance"][:, 3] + 0.03) < 1e-6)
    o.set_state(x=far, v=np.zeros_like(far), xprev=far)
    o.step(1)
    assert len(o.contacts()[1]) == 0


def test_analytic_mesh_leaves_the_triangle_path():
    sc = scenes.config_analytic(n=16, mesh_obstacle=True)
    o = O.OracleScene(sc)
    for k in range(12):
        o.step(10)
    pairs = o.pairs()
    analytic_meshes = set(sc.analytic["obj_index"].tolist())
    assert len(pairs) and not (set(pairs[:, 2].tolist()) & analytic_meshes)
    off, rec = o.contacts()
    a = np.isin(rec["target_obj"], list(analytic_meshes))
    assert a.any() and np.all(rec["triangle"][a] == 0)
    assert np.array_equal(rec["target_shape"][a], rec["target_obj"][a])   # shape lane = kind (1..4 here)
    # a non-collidable shape keeps its mesh out of the triangle path without contacts
    sc2 = _box_scene(shape_flags=0)
    o2 = O.OracleScene(sc2)
    x, _, _ = o2.state()
    x[:64, 1] = 0.07
    o2.set_state(x=x, xprev=x)
    o2.step(1)
    assert len(o2.contacts()[1]) == 0 and 1 not in o2.pairs()[:, 2]


def test_set_analytic_rejections():
    sc = _box_scene()
    o = O.OracleScene(sc)
    bad = [analytic_shape(COLLIDER_BOX, 0, (0, 0, 0), half=(0.1, 0.1, 0.1)),       # dynamic mesh
           analytic_shape(COLLIDER_SPHERE, 1, (0, 0, 0), half=(0.1, 0.1, 0.1)),    # shape lane mismatch
           analytic_shape(0, 1, (0, 0, 0)),                                         # kind Mesh
           analytic_shape(COLLIDER_BOX, 2, (0, 0, 0), half=(0.1, 0.1, 0.1)),       # no such mesh
           np.concatenate([sc.analytic, sc.analytic])]                               # two shapes, one mesh
    for b in bad:
        with pytest.raises(ValueError):
            o.set_analytic(b)
    o.set_analytic(sc.analytic[:0])
    assert len(o.analytic_markers()[0]) == 0
