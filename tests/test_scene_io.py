"""Scene files (scene_format JSON + RunConfig) into packed scenes.

Parsing mirrors the reference's own scene_io / sim_config tests
(test/scene_io_test.cpp: BDD-014/015/016, clamping, import paths, collider
defaults, environment defaults); the primitives are pinned by the closed-form
counts of test/primitive_test.cpp and by their geometry (closed, consistently
wound surfaces; vertices on the sphere / cube / cylinder); the jiggle RNG by
the C++ standard's mt19937 known answer; the scene build by its placements
and forces; the three scene files the reference ships (tests/golden/scenes)
by a drop to rest on the ground at ground + thickness in the oracle."""
import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import OracleScene
from ysim_amd import scene_io as S
from ysim_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "scenes")
FILES = ["synthetic.json", "cloth_drape.json", "subobject_cloth.json"]

MAT = {"base_color": [1, 1, 1], "metallic": 0.0, "roughness": 0.5, "specular_weight": 1.0,
       "emission_color": [0, 0, 0]}


def _obj(idx=0, source=None, behavior="Float", params=None, **extra):
    o = {"id": idx, "name": f"o{idx}",
         "source": source or {"type": "primitive", "shape": "grid", "size": 1.0, "tessellation": 2},
         "transform": {"position": [0, 0, 0], "rotation": [1, 0, 0, 0]}, "material": dict(MAT),
         "behavior": {"type": behavior, "params": params or {}}}
    o.update(extra)
    return o


def _scene(*objs, **top):
    d = {"format_version": 1, "objects": list(objs), "environment": {"gravity": [0, 0, 0], "wind": [0, 0, 0]}}
    d.update(top)
    return d


# ------------------------------------------------------------------ parsing
def test_rejections_match_the_reference_texts():
    with pytest.raises(S.SceneFormatError, match="999.*1"):
        S.parse_scene(json.dumps({"format_version": 999, "objects": []}))
    with pytest.raises(S.SceneFormatError, match="missing format_version"):
        S.parse_scene(json.dumps({"objects": []}))
    with pytest.raises(S.SceneFormatError, match="Elastic.*not available"):
        S.parse_scene(_scene(_obj(behavior="Elastic")))
    with pytest.raises(S.SceneFormatError, match="unknown behavior.type 'Soap'"):
        S.parse_scene(_scene(_obj(behavior="Soap")))
    with pytest.raises(S.SceneFormatError, match="foo.fbx"):
        S.parse_scene(_scene(_obj(source={"type": "import", "path": "foo.fbx"})))
    with pytest.raises(S.SceneFormatError, match="unknown source.shape 'torus'"):
        S.parse_scene(_scene(_obj(source={"type": "primitive", "shape": "torus", "size": 1, "tessellation": 2})))
    with pytest.raises(S.SceneFormatError, match="missing 'material'"):
        o = _obj()
        del o["material"]
        S.parse_scene(_scene(o))
    with pytest.raises(S.SceneFormatError, match="4-element"):
        o = _obj()
        o["transform"]["rotation"] = [1, 0, 0]
        S.parse_scene(_scene(o))
    with pytest.raises(S.SceneFormatError, match="substeps must be >= 1"):
        S.parse_scene(_scene(simulation={"substeps": 0}))
    with pytest.raises(S.SceneFormatError, match="JSON parse error"):
        S.parse_scene("{not json")


def test_defaults_clamps_and_optional_keys():
    snap = S.parse_scene(json.dumps({"format_version": 1, "objects": []}))
    assert snap["environment"]["gravity"] == [0.0, -9.81, 0.0] and snap["environment"]["wind"] == [0.0, 0.0, 0.0]
    assert snap["simulation"] == {"time_per_frame": 1.0 / 60.0, "substeps": 60}
    o = _obj(source={"type": "primitive", "shape": "sphere", "size": 1.0, "tessellation": 8})
    o["material"] = {"base_color": [1.5, -0.2, 0.5], "metallic": 2.0, "roughness": -0.1, "specular_weight": 3.0,
                     "emission_color": [-1.0, 0.0, 0.0]}
    o["transform"]["rotation"] = [2, 0, 0, 0]   # renormalised on read
    snap = S.parse_scene(_scene(o))
    m = snap["objects"][0]["material"]
    assert snap["warnings"] and m["base_color"] == [1.0, 0.0, 0.5] and m["metallic"] == 1.0
    assert m["roughness"] == 0.0 and m["specular_weight"] == 1.0 and m["emission_color"] == [0.0, 0.0, 0.0]
    ob = snap["objects"][0]
    assert ob["transform"]["rotation"] == [1.0, 0.0, 0.0, 0.0] and ob["transform"]["scale"] == [1.0, 1.0, 1.0]
    assert ob["collider_kind"] == "" and ob["collidable"] and not ob["self_collide"]
    assert ob["apply_gravity"] and ob["apply_wind"] and ob["cloth_stiffness_scale"] == 1.0


def test_save_load_save_is_byte_identical(tmp_path):
    o0 = _obj(0, behavior="TriangularCloth", params={"stretch": 1234.5, "shear": 567.0, "bend": 78.9,
                                                     "thickness": 0.01},
              collider_kind="plane", collidable=False, self_collide=True,
              fixed_particles=[{"vid": 3, "pos": [0.1, 0.2, 0.3]}, {"vid": "bad"}])
    o0["transform"]["scale"] = [2.0, 1.0, 1.0]
    o1 = _obj(1, source={"type": "import", "path": "assets/teapot.obj", "scale": 1.5, "mass": 0.1},
              collider_kind="box")
    d = _scene(o0, o1, reference_constraints=[{"query": {"object": 0, "vertex": 1},
                                                "target": {"object": 1, "vertex": 2}}])
    snap = S.parse_scene(d)
    p1, p2 = str(tmp_path / "a.json"), str(tmp_path / "b.json")
    S.write_scene(snap, p1)
    back = S.read_scene(p1)
    S.write_scene(back, p2)
    assert open(p1).read() == open(p2).read()
    ob = back["objects"]
    assert ob[0]["collider_kind"] == "plane" and not ob[0]["collidable"] and ob[0]["self_collide"]
    assert ob[1]["collider_kind"] == "box" and ob[1]["collidable"] and not ob[1]["self_collide"]
    assert ob[0]["fixed_particles"] == [{"vid": 3, "pos": [0.1, 0.2, 0.3]}]
    assert back["reference_constraints"] == snap["reference_constraints"]
    j = S.to_json(S.parse_scene(_scene(_obj())))
    assert "collider_kind" not in j["objects"][0] and "scale" not in j["objects"][0]["transform"]


@pytest.mark.parametrize("name", FILES)
def test_reference_scene_files_round_trip(name):
    snap = S.read_scene(os.path.join(GOLD, name))
    again = S.parse_scene(json.dumps(S.to_json(snap)))
    snap.pop("warnings"), again.pop("warnings")
    assert again == snap


def test_run_config_engine_and_profile():
    cfg = S.read_run_config(os.path.join(GOLD, "cloth_drape.json"))
    assert cfg["engine"] == {"backend": "METAL", "system": "Explicit"}
    assert cfg["profile"]["enabled"] and cfg["profile"]["frames"] == 20 and cfg["profile"]["level"] == "in_frame"
    with pytest.raises(S.SceneFormatError, match="unknown engine.backend 'TPU'"):
        S.parse_run_config(dict(_scene(), engine={"backend": "TPU"}))
    # the HIP tag is accepted where the reference's isKnownBackend is extended
    assert S.parse_run_config(dict(_scene(), engine={"backend": "HIP"}))["engine"]["backend"] == "HIP"
    with pytest.raises(S.SceneFormatError, match="profile.frames must be >= 1"):
        S.parse_run_config(dict(_scene(), profile={"frames": 0}))


def test_import_path_helpers():
    assert S.resolve_import_path("/scenes/2026", "assets/teapot.obj") == "/scenes/2026/assets/teapot.obj"
    assert S.resolve_import_path("scenes", "assets/teapot.obj") == "scenes/assets/teapot.obj"
    assert S.resolve_import_path("", "assets/teapot.obj") == "assets/teapot.obj"
    assert S.resolve_import_path("/scenes/2026", "/library/teapot.obj") == "/library/teapot.obj"
    assert S.resolve_import_path("/scenes/2026/", "assets/teapot.obj") == "/scenes/2026/assets/teapot.obj"
    assert S.scene_dir("/scenes/2026/foo.ysim.json") == "/scenes/2026"
    assert S.scene_dir("foo.ysim.json") == "" and S.scene_dir("./foo.ysim.json") == "."


# ------------------------------------------------------------------ geometry
def test_mt19937_known_answers():
    """[rand.predef]: the 10000th output of a default-constructed mt19937 (seed
    5489) is 4123659995."""
    r = S.mt19937(5489, 10000)
    assert int(r[-1]) == 4123659995 and int(r[0]) == 3499211612
    j = S.jiggle_values(7, 1000)
    assert j.dtype == np.float32 and j.min() >= 0 and j.max() < np.float32(1e-4)
    assert np.array_equal(j, S.jiggle_values(7, 1000)) and not np.array_equal(j, S.jiggle_values(8, 1000))


def _closed(f):
    """Every undirected edge in exactly two facets."""
    und = {}
    for a, b, c in f.tolist():
        for u, v in ((a, b), (b, c), (c, a)):
            k = (min(u, v), max(u, v))
            und[k] = und.get(k, 0) + 1
    return all(n == 2 for n in und.values())


def _closed_and_wound(f):
    """Every undirected edge in exactly two facets, every directed edge once."""
    directed = {}
    for a, b, c in f.tolist():
        for u, v in ((a, b), (b, c), (c, a)):
            directed[(u, v)] = directed.get((u, v), 0) + 1
    if any(n != 1 for n in directed.values()):
        return False
    return all((v, u) in directed for (u, v) in directed)


def _volume(x, f):
    a, b, c = (x[f[:, k]].astype(np.float64) for k in range(3))
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


@pytest.mark.parametrize("t", [3, 4, 8, 16])
def test_sphere_primitive(t):
    c = np.array([0.25, -0.5, 1.0])
    x, f = S.sphere_geometry(1.5, t, c)
    assert x.shape[0] == (t - 1) * t + 2 and f.shape[0] == 2 * t * (t - 1)
    assert np.abs(np.linalg.norm(x - c, axis=1) - 0.75).max() < 1e-5
    assert _closed_and_wound(f) and f.max() < x.shape[0]
    assert abs(_volume(x - c, f)) <= 4 / 3 * math.pi * 0.75 ** 3 + 1e-6
    area = np.linalg.norm(np.cross(x[f[:, 1]] - x[f[:, 0]], x[f[:, 2]] - x[f[:, 0]]), axis=1)
    assert area.min() > 0


@pytest.mark.parametrize("t", [1, 2, 4, 8])
def test_cube_primitive(t):
    x, f = S.cube_geometry(2.0, t, (2.0, 3.0, 4.0))
    assert x.shape[0] == 6 * t * t + 2 and f.shape[0] == 12 * t * t
    d = np.abs(x - np.array([2.0, 3.0, 4.0])).max(axis=1)
    assert np.abs(d - 1.0).max() < 1e-6                       # every vertex on a face plane
    assert _closed_and_wound(f)                                 # welded seams: one closed surface
    assert abs(abs(_volume(x - np.array([2.0, 3.0, 4.0]), f)) - 8.0) < 1e-4


@pytest.mark.parametrize("t", [3, 6, 12])
def test_cylinder_primitive(t):
    x, f = S.cylinder_geometry(1.0, t, (0.0, 0.0, 0.0))
    assert x.shape[0] == (t + 1) * t + 2 and f.shape[0] == 2 * t * t + 2 * t
    ring = x[:-2]
    assert np.abs(np.hypot(ring[:, 0], ring[:, 2]) - 0.5).max() < 1e-6
    assert np.allclose(x[-2:], [[0, -0.5, 0], [0, 0.5, 0]])
    # closed; the wall is wound one way and both caps the other, as
    # primitive::cylinder writes them (the narrow phase orients normals by
    # the vertex's side, so winding does not enter the contacts)
    assert _closed(f)
    wall, caps = f[: 2 * t * t], f[2 * t * t:]
    n_gon = 0.5 * t * 0.25 * math.sin(2 * math.pi / t)          # polygon area x height 1
    assert abs(abs(_volume(x, wall) - _volume(x, caps)) - n_gon) < 1e-5


def test_grid_matches_the_cloth_grid_layout():
    x, f = S.grid_geometry(5, 1.0, "XZPlane", (0.5, 0.25, -0.5), False, 0)
    ref = scenes.grid_positions(5, 1.0, (0.0, 0.25, -1.0), np.random.default_rng(0), jitter_rel=0.0)
    assert np.allclose(x, ref, atol=1e-7) and np.array_equal(f, scenes.grid_facets(5))
    y, _ = S.grid_geometry(5, 1.0, "XZPlane", (0.5, 0.25, -0.5), True, 3)
    assert np.array_equal(y[:, [0, 2]], x[:, [0, 2]])
    assert np.array_equal(y[:, 1], (np.float32(0.25) + S.jiggle_values(3, 25)).astype(np.float32))


def test_transform_bake_rotates_about_the_position():
    q = scenes.quat_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
    o = _obj(0, transform=None)
    o["transform"] = {"position": [0.0, 1.0, 0.0], "rotation": list(q), "scale": [2.0, 1.0, 1.0]}
    sc = S.build_scene(S.parse_scene(_scene(o)))
    x = sc.meshes[0].x.astype(np.float64)
    flat, _ = S.grid_geometry(2, 1.0, "XZPlane", (0.0, 1.0, 0.0), False, 0)
    d = (flat - [0.0, 1.0, 0.0]) * [2.0, 1.0, 1.0]
    expect = scenes.quat_apply(q, d) + [0.0, 1.0, 0.0]
    assert np.abs(x - expect).max() < 1e-6 and np.abs(x[:, 2]).max() < 1e-6   # the XZ grid now lies in XY


# ------------------------------------------------------------------ scene build
def test_synthetic_scene_build():
    sc = S.load_scene(os.path.join(GOLD, "synthetic.json"))
    cl, cube, ground = sc.meshes
    assert (cl.x.shape[0], cl.facets.shape[0], cl.behavior) == (16, 18, scenes.BEHAVIOR_TRIANGULAR_CLOTH)
    assert (cube.x.shape[0], cube.facets.shape[0], cube.shape) == (26, 48, scenes.COLLIDER_BOX)
    assert (ground.x.shape[0], ground.behavior, ground.shape) == (4, scenes.BEHAVIOR_FLOAT, 0)
    assert sc.params["subh"] == float(np.float32(np.float32(1 / 60) / np.float32(8)))
    assert sc.params["margin"] == 0.015 and sc.params["radius"] == 0.012 and sc.params["rebuild_every"] == 80
    a = sc.analytic
    assert a.shape == (1,) and a["kind"][0] == 2 and a["obj_index"][0] == 1
    assert np.allclose(a["half_ext_height"][0, :3], 0.1) and np.allclose(a["center_radius"][0, :3], [0.5, -0.2, 0])
    # grid cloth: construction jiggle (seed = id 0) then the cloth jiggle (seed id ^ index = 0)
    j = S.jiggle_values(0, 16)
    assert np.array_equal(cl.x[:, 1], ((np.float32(0.25) + j) + j).astype(np.float32))
    assert cl.cloth["kstretch"] == 1000.0 and cl.cloth["kbend"] == 1000.0
    pk = sc.packed()
    ext = pk["ext"].reshape(-1, 3)
    assert np.array_equal(ext[:16, 1], np.full(16, np.float32(-9.8100004196167) * np.float32(0.1), np.float32))
    assert not ext[16:].any()


def test_forces_pins_stiffness_scale_fastgrid_and_obj(tmp_path):
    obj = tmp_path / "quad.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\nf 1/1 2/2 3/3 4/4\n")
    grid = _obj(0, source={"type": "primitive", "shape": "grid", "size": 1.0, "tessellation": 4, "mass": 0.5},
                behavior="FastGridCloth", params={"particle_num_1d": 4, "k_stretch": 100.0, "k_shear": 10.0,
                                                 "k_bend": 1.0, "thickness": 0.02},
                cloth_stiffness_scale=2.0, apply_wind=True, apply_gravity=False,
                fixed_particles=[{"vid": 0, "pos": [9.0, 9.0, 9.0]}])
    imp = _obj(1, source={"type": "import", "path": "quad.obj", "scale": 2.0})
    imp["transform"]["position"] = [0.0, -1.0, 0.0]
    d = _scene(grid, imp, environment={"gravity": [0, -9.81, 0], "wind": [0.5, 0, 0]})
    sc = S.build_scene(S.parse_scene(d), str(tmp_path))
    g, q = sc.meshes
    assert g.cloth["kstretch"] == 200.0 and g.cloth["kshear"] == 20.0 and g.cloth["kbend"] == 2.0
    assert abs(g.cloth["stretch_rest_x"] - 1 / 3) < 1e-6 and abs(g.cloth["shear_rest_a"] - math.sqrt(2) / 3) < 1e-6
    assert g.fixed[0] == 0 and np.array_equal(g.x[0], [9.0, 9.0, 9.0]) and g.fixed[1:].all()
    ext = sc.packed()["ext"].reshape(-1, 3)
    assert np.array_equal(ext[:16], np.tile(np.float32([0.5, 0, 0]), (16, 1)))   # wind only, no gravity
    assert q.facets.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert np.array_equal(q.x, np.float32([[0, -1, 0], [2, -1, 0], [2, -1, 2], [0, -1, 2]]))
    with pytest.raises(S.SceneFormatError, match="particle_num_1d"):
        bad = _obj(0, behavior="FastGridCloth", params={"particle_num_1d": 5})
        S.build_scene(S.parse_scene(_scene(bad)))


def test_collider_kinds_and_fits():
    cyl = _obj(0, source={"type": "primitive", "shape": "cylinder", "size": 0.4, "tessellation": 12})
    cyl["transform"]["scale"] = [1.0, 2.0, 1.0]
    plane = _obj(1, collider_kind="plane")
    ball = _obj(2, source={"type": "primitive", "shape": "grid", "size": 1.0, "tessellation": 3},
                collider_kind="sphere")
    mesh_cube = _obj(3, source={"type": "primitive", "shape": "cube", "size": 1.0, "tessellation": 1},
                     collider_kind="mesh")
    sc = S.build_scene(S.parse_scene(_scene(cyl, plane, ball, mesh_cube)))
    a = sc.analytic
    assert a["kind"].tolist() == [3, 4, 1] and a["obj_index"].tolist() == [0, 1, 2]
    assert np.allclose(a["center_radius"][0, 3], 0.2) and np.allclose(a["half_ext_height"][0, 3], 0.4)
    # a grid fitted as a sphere: the origin-centred box of its vertices (0.5, 0, 0.5)
    assert np.allclose(a["half_ext_height"][2, :3], [0.5, 0.0, 0.5])
    assert [m.shape for m in sc.meshes] == [3, 4, 1, 0]
    with pytest.raises(S.SceneFormatError, match="analytic collider on a cloth"):
        S.build_scene(S.parse_scene(_scene(_obj(0, behavior="TriangularCloth", collider_kind="box"))))


# ------------------------------------------------------------------ stepping (oracle)
def test_synthetic_scene_comes_to_rest_on_the_ground(oracle_threads):
    """The 4 x 4 cloth misses the cube, falls onto the ground grid at y = -1
    and rests at ground + thickness (the contact push-out), unmoved in x/z."""
    sc = S.load_scene(os.path.join(GOLD, "synthetic.json"))
    o = OracleScene(sc)
    o.step(60 * sc.info["substeps"])
    x, v, _ = o.state()
    assert np.abs(x[:16, 1] - np.float32(-0.99)).max() < 1e-5
    assert np.abs(x[:16, [0, 2]] - sc.meshes[0].x[:, [0, 2]]).max() < 1e-3
    off, rec = o.contacts()
    assert set(rec["target_obj"].tolist()) == {2}


def test_cloth_drape_touches_the_analytic_sphere(oracle_threads):
    sc = S.load_scene(os.path.join(GOLD, "cloth_drape.json"))
    o = OracleScene(sc)
    o.step(30 * sc.info["substeps"])
    off, rec = o.contacts()
    assert len(rec) > 500 and set(rec["target_obj"].tolist()) == {1} and set(rec["triangle"].tolist()) == {0}
    x, _, _ = o.state()
    r = np.linalg.norm(x[:2500] - sc.analytic["center_radius"][0, :3], axis=1)
    assert r.min() > 0.5 - 0.02     # held outside the sphere (radius 0.5) within the band
