"""Scene files: the reference's versioned JSON scene format and run config,
turned into a packed scene for the HIP backend.

Parsing and validation follow scene_format::fromJson and its helpers
(include/scene_format.hpp:175-737) and sim_config (include/sim_config.hpp:
43-293): the same required keys, defaults, reserved and unknown names, error
texts and clamping warnings.  `to_json` is the encoder (scene_format.hpp:
282-418), so a parsed scene saves and reloads field for field.

`build_scene` is Simulator::applySnapshot + Scene::pack for the meshes this
backend runs (include/sim/simulator.hpp:4573-4840, include/sim/scene.hpp:
783-1110):

* primitive sources: the grid of MeshGridInitializer (initializers.hpp:
  47-128) with its optional construction jiggle, and the sphere, cube and
  cylinder of primitive_geometry.hpp; import sources: Wavefront .obj vertices
  and faces scaled and offset by the object position (MeshFileInitializer,
  initializers.hpp:228-295), polygons fanned into triangles;
* the object transform baked into the positions: scale, then rotation about
  the object position (scene.hpp:958-990);
* the cloth jiggle of a grid whose behavior is a cloth, seeded with
  object id ^ load index (applyClothJiggle, initializers.hpp:191-224);
* behaviors: TriangularCloth (stretch and bend are the spring constants,
  times cloth_stiffness_scale, collision_pod.hpp:479-487), FastGridCloth
  (its params, rest lengths re-measured from the placed grid,
  recomputeFastGridRest collision_pod.hpp:419-442, constants times the
  scale), Float; Rigid objects are static here (this path has no rigid-body
  integrator), as they are when the reference runs without its rigid backend;
* per-vertex mass = source mass; external force = gravity * mass when
  apply_gravity, plus wind for cloths when apply_wind, zero for static
  objects (Simulator::applyEnvironmentForces, simulator.hpp:3122-3166);
* collider kind: the initializer's default (sphere, box, cylinder; grid and
  import: mesh) unless `collider_kind` names one (scene.hpp:326-331,
  collision_pod.hpp:247-254); a non-mesh kind becomes an analytic collider
  with the fitted half-extents of fitAnalyticLocalHalf (scene.hpp:603-641);
* pinned vertices (fixed_particles): held at their saved position, mask 0;
* solver timing: substep dt = time_per_frame / substeps in float
  (simulator.hpp:4632-4637); Simulator defaults margin 0.015, radius 0.012,
  kair -0.001, kd 0.5 (simulator.hpp:180-181, symplectic_system.hpp:197-198),
  the narrow band radius + margin, the trees rebuilt every 10 frames and
  refit every substep.

Self collision: the scene-wide switch is on when any object sets
`self_collide` (the reference's per-object toggle feeds a CPU spatial-hash
self pass; here the one tree query serves all self contacts).
"""
from __future__ import annotations

import copy
import json
import math
import os
from typing import Optional

import numpy as np

from . import scenes

F32 = np.float32
FORMAT_VERSION = 1

KNOWN_BEHAVIORS = ("Float", "TriangularCloth", "FastGridCloth", "Rigid")
RESERVED_BEHAVIORS = ("Elastic", "Fluid", "Generator")
KNOWN_SHAPES = ("grid", "sphere", "cube", "cylinder")
# the reference's isKnownBackend (include/sim_config.hpp:100-102) plus the tag
# this backend adds beside them
KNOWN_BACKENDS = ("METAL", "CPU", "CUDA", "HIP")
KNOWN_SYSTEMS = ("Explicit", "PBD", "PD", "LargeSteps")
PROFILE_LEVELS = ("none", "per_frame", "in_frame")
COLLIDER_NAMES = {"mesh": 0, "sphere": 1, "box": 2, "cylinder": 3, "plane": 4}

# Simulator / SymplecticSystem defaults (simulator.hpp:157, 180-181; symplectic_system.hpp:197-198)
DEFAULT_MARGIN = 0.015
DEFAULT_RADIUS = 0.012
DEFAULT_KAIR = -0.001
DEFAULT_KD = 0.5
REBUILD_EVERY_FRAMES = 10


class SceneFormatError(ValueError):
    """A scene or run-config file the loader refuses (the reference's LoadError)."""


# ------------------------------------------------------------------ parsing
def _is_num(v) -> bool:
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _is_int(v) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


def _vec(j, n: int, path: str) -> list:
    what = "3-element array" if n == 3 else "4-element [w,x,y,z] array"
    if not isinstance(j, list) or len(j) != n:
        raise SceneFormatError(f"expected {what} at {path}")
    for i, v in enumerate(j):
        if not _is_num(v):
            raise SceneFormatError(f"expected number at {path}[{i}]")
    return [float(v) for v in j]


def _field(j: dict, key: str, path: str, check, kind: str):
    if key not in j:
        raise SceneFormatError(f"missing '{key}' at {path}")
    v = j[key]
    if not check(v):
        raise SceneFormatError(f"wrong type for '{key}' at {path}: expected {kind}")
    return v


def _normalize_quat(q):
    n = math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
    if n == 0.0:
        return [1.0, 0.0, 0.0, 0.0]
    return [q[0] / n, q[1] / n, q[2] / n, q[3] / n]


def _clamp01(x: float) -> float:
    return 0.0 if x < 0.0 else (1.0 if x > 1.0 else x)


def _source(j, idx: int) -> dict:
    path = f"objects[{idx}].source"
    typ = _field(j, "type", path, lambda v: isinstance(v, str), "string")
    if typ == "primitive":
        shape = _field(j, "shape", path, lambda v: isinstance(v, str), "string")
        if shape not in KNOWN_SHAPES:
            raise SceneFormatError(f"unknown source.shape '{shape}' at objects[{idx}]")
        s = dict(type="primitive", shape=shape,
                 size=float(_field(j, "size", path, _is_num, "number")),
                 tessellation=int(_field(j, "tessellation", path, _is_int, "integer")),
                 direction="XZPlane", mass=0.1, jiggle=False)
        if "direction" in j:
            s["direction"] = str(j["direction"])
        if "mass" in j:
            s["mass"] = float(j["mass"])
        if "jiggle" in j:
            s["jiggle"] = bool(j["jiggle"])
        return s
    if typ == "import":
        p = _field(j, "path", path, lambda v: isinstance(v, str), "string")
        ext = os.path.splitext(p)[1].lower()
        if ext != ".obj":
            raise SceneFormatError(f"unsupported source.path extension '{p}' at objects[{idx}]")
        s = dict(type="import", path=p, scale=1.0, mass=0.1)
        if "scale" in j:
            s["scale"] = float(j["scale"])
        if "mass" in j:
            s["mass"] = float(j["mass"])
        return s
    raise SceneFormatError(f"unknown source.type '{typ}' at objects[{idx}]")


def _object(j, idx: int, warnings: list) -> dict:
    path = f"objects[{idx}]"
    if not isinstance(j, dict):
        raise SceneFormatError(f"{path} must be a JSON object")
    o = dict(id=int(_field(j, "id", path, _is_int, "integer")),
             name=str(_field(j, "name", path, lambda v: isinstance(v, str), "string")))
    if "source" not in j:
        raise SceneFormatError(f"missing 'source' at {path}")
    o["source"] = _source(j["source"], idx)
    if "transform" not in j:
        raise SceneFormatError(f"missing 'transform' at {path}")
    t, tp = j["transform"], f"{path}.transform"
    if "position" not in t:
        raise SceneFormatError(f"missing 'position' at {tp}")
    pos = _vec(t["position"], 3, tp + ".position")
    if "rotation" not in t:
        raise SceneFormatError(f"missing 'rotation' at {tp}")
    rot = _normalize_quat(_vec(t["rotation"], 4, tp + ".rotation"))
    scale = _vec(t["scale"], 3, tp + ".scale") if "scale" in t else [1.0, 1.0, 1.0]
    o["transform"] = dict(position=pos, rotation=rot, scale=scale)
    if "material" not in j:
        raise SceneFormatError(f"missing 'material' at {path}")
    m, mp = j["material"], f"{path}.material"
    if "base_color" not in m:
        raise SceneFormatError(f"missing 'base_color' at {mp}")
    mat = dict(base_color=_vec(m["base_color"], 3, mp + ".base_color"),
               metallic=float(_field(m, "metallic", mp, _is_num, "number")),
               roughness=float(_field(m, "roughness", mp, _is_num, "number")),
               specular_weight=float(_field(m, "specular_weight", mp, _is_num, "number")))
    if "emission_color" not in m:
        raise SceneFormatError(f"missing 'emission_color' at {mp}")
    mat["emission_color"] = _vec(m["emission_color"], 3, mp + ".emission_color")
    clamped = dict(base_color=[_clamp01(c) for c in mat["base_color"]],
                   metallic=_clamp01(mat["metallic"]), roughness=_clamp01(mat["roughness"]),
                   specular_weight=_clamp01(mat["specular_weight"]),
                   emission_color=[max(c, 0.0) for c in mat["emission_color"]])
    if clamped != mat:
        warnings.append(f"clamped out-of-range material values at {mp}")
    o["material"] = clamped
    if "behavior" not in j:
        raise SceneFormatError(f"missing 'behavior' at {path}")
    b, bp = j["behavior"], f"{path}.behavior"
    btype = _field(b, "type", bp, lambda v: isinstance(v, str), "string")
    if btype in RESERVED_BEHAVIORS:
        raise SceneFormatError(f"behavior '{btype}' not available in this build (objects[{idx}])")
    if btype not in KNOWN_BEHAVIORS:
        raise SceneFormatError(f"unknown behavior.type '{btype}' at {bp}")
    o["behavior"] = dict(type=btype, params=copy.deepcopy(b.get("params", {})))
    o["apply_gravity"] = j["apply_gravity"] if isinstance(j.get("apply_gravity"), bool) else True
    o["apply_wind"] = j["apply_wind"] if isinstance(j.get("apply_wind"), bool) else True
    o["is_static"] = j["is_static"] if isinstance(j.get("is_static"), bool) else False
    o["cloth_stiffness_scale"] = float(j["cloth_stiffness_scale"]) \
        if _is_num(j.get("cloth_stiffness_scale")) else 1.0
    o["collider_kind"] = j["collider_kind"] if isinstance(j.get("collider_kind"), str) else ""
    o["collidable"] = j["collidable"] if isinstance(j.get("collidable"), bool) else True
    o["self_collide"] = j["self_collide"] if isinstance(j.get("self_collide"), bool) else False
    fixed = []
    for e in j.get("fixed_particles", []) if isinstance(j.get("fixed_particles"), list) else []:
        if not isinstance(e, dict) or not _is_int(e.get("vid")) or "pos" not in e:
            continue
        try:
            fixed.append(dict(vid=e["vid"], pos=_vec(e["pos"], 3, path)))
        except SceneFormatError:
            continue
    o["fixed_particles"] = fixed
    splines = []
    for e in j.get("spline_constraints", []) if isinstance(j.get("spline_constraints"), list) else []:
        if not isinstance(e, dict) or not _is_int(e.get("vid")) or not isinstance(e.get("points"), list):
            continue
        try:
            pts = [_vec(p, 3, path) for p in e["points"]]
        except SceneFormatError:
            continue
        if len(pts) < 2:
            continue
        splines.append(dict(vid=e["vid"], closed=bool(e.get("closed", False)) if isinstance(e.get("closed"), bool)
                            else False, duration=float(e["duration"]) if _is_num(e.get("duration")) else 4.0,
                            points=pts))
    o["spline_constraints"] = splines
    return o


def parse_scene(data) -> dict:
    """scene_format::fromJson: a JSON text or an already-decoded object."""
    if isinstance(data, (str, bytes)):
        try:
            j = json.loads(data)
        except json.JSONDecodeError as e:
            raise SceneFormatError(f"JSON parse error: {e}") from None
    else:
        j = data
    if not isinstance(j, dict):
        raise SceneFormatError("scene root must be a JSON object")
    if "format_version" not in j:
        raise SceneFormatError("missing format_version")
    fv = j["format_version"]
    if not _is_int(fv):
        raise SceneFormatError(f"format_version must be an integer (got {type(fv).__name__})")
    if fv != FORMAT_VERSION:
        raise SceneFormatError(f"format_version mismatch: found {fv}, expected {FORMAT_VERSION}")
    objs = j.get("objects")
    if not isinstance(objs, list):
        raise SceneFormatError("missing or invalid 'objects' array")
    warnings: list = []
    snap = dict(format_version=fv, objects=[_object(o, i, warnings) for i, o in enumerate(objs)])
    env = dict(gravity=[0.0, -9.81, 0.0], wind=[0.0, 0.0, 0.0], background_color=[0.68, 0.85, 0.95])
    if "environment" in j:
        e = j["environment"]
        for k in ("gravity", "wind", "background_color"):
            if k in e:
                env[k] = _vec(e[k], 3, f"environment.{k}")
    snap["environment"] = env
    sim = dict(time_per_frame=1.0 / 60.0, substeps=60)
    if "simulation" in j:
        s = j["simulation"]
        if not isinstance(s, dict):
            raise SceneFormatError("'simulation' must be a JSON object")
        if "time_per_frame" in s:
            if not _is_num(s["time_per_frame"]):
                raise SceneFormatError("simulation.time_per_frame must be a number")
            if not s["time_per_frame"] > 0.0:
                raise SceneFormatError("simulation.time_per_frame must be > 0")
            sim["time_per_frame"] = float(s["time_per_frame"])
        if "substeps" in s:
            if not _is_int(s["substeps"]):
                raise SceneFormatError("simulation.substeps must be an integer")
            if s["substeps"] < 1:
                raise SceneFormatError("simulation.substeps must be >= 1")
            sim["substeps"] = s["substeps"]
    snap["simulation"] = sim
    refs = []
    for e in j.get("reference_constraints", []) if isinstance(j.get("reference_constraints"), list) else []:
        q, t = (e.get("query"), e.get("target")) if isinstance(e, dict) else (None, None)
        if not isinstance(q, dict) or not isinstance(t, dict):
            continue
        vals = [q.get("object"), q.get("vertex"), t.get("object"), t.get("vertex")]
        if not all(_is_int(v) for v in vals):
            continue
        refs.append(dict(query_object=vals[0], query_vertex=vals[1], target_object=vals[2], target_vertex=vals[3]))
    snap["reference_constraints"] = refs
    snap["warnings"] = warnings
    return snap


def read_scene(path: str) -> dict:
    try:
        with open(path, "r") as f:
            text = f.read()
    except OSError:
        raise SceneFormatError(f"cannot open '{path}' for reading") from None
    return parse_scene(text)


def scene_dir(scene_path: str) -> str:
    """sceneDir (scene_format.hpp:794-798): the path up to its last '/'."""
    pos = scene_path.rfind("/")
    return "" if pos < 0 else scene_path[:pos]


def resolve_import_path(directory: str, import_path: str) -> str:
    """resolveImportPath (scene_format.hpp:809-814): absolute paths pass,
    relative ones join the scene file's directory."""
    if import_path.startswith("/") or not directory:
        return import_path
    return directory + import_path if directory.endswith("/") else directory + "/" + import_path


def to_json(snap: dict) -> dict:
    """scene_format::toJson: optional keys omitted when at their defaults."""
    objs = []
    for o in snap["objects"]:
        s = o["source"]
        src = (dict(type="primitive", shape=s["shape"], size=s["size"], tessellation=s["tessellation"],
                    direction=s["direction"], mass=s["mass"], jiggle=s["jiggle"])
               if s["type"] == "primitive" else dict(type="import", path=s["path"], scale=s["scale"], mass=s["mass"]))
        t = o["transform"]
        tr = dict(position=list(t["position"]), rotation=_normalize_quat(t["rotation"]))
        if list(t["scale"]) != [1.0, 1.0, 1.0]:
            tr["scale"] = list(t["scale"])
        j = dict(id=o["id"], name=o["name"], source=src, transform=tr, material=copy.deepcopy(o["material"]),
                 behavior=dict(type=o["behavior"]["type"], params=copy.deepcopy(o["behavior"]["params"])),
                 apply_gravity=o["apply_gravity"], apply_wind=o["apply_wind"], is_static=o["is_static"],
                 cloth_stiffness_scale=o["cloth_stiffness_scale"], collidable=o["collidable"],
                 self_collide=o["self_collide"])
        if o["collider_kind"]:
            j["collider_kind"] = o["collider_kind"]
        if o["fixed_particles"]:
            j["fixed_particles"] = [dict(vid=f["vid"], pos=list(f["pos"])) for f in o["fixed_particles"]]
        if o["spline_constraints"]:
            j["spline_constraints"] = [dict(vid=c["vid"], closed=c["closed"], duration=c["duration"],
                                            points=[list(p) for p in c["points"]]) for c in o["spline_constraints"]]
        objs.append(j)
    out = dict(format_version=snap["format_version"], objects=objs, environment=copy.deepcopy(snap["environment"]),
               simulation=copy.deepcopy(snap["simulation"]))
    if snap.get("reference_constraints"):
        out["reference_constraints"] = [
            dict(query=dict(object=c["query_object"], vertex=c["query_vertex"]),
                 target=dict(object=c["target_object"], vertex=c["target_vertex"]))
            for c in snap["reference_constraints"]]
    return out


def write_scene(snap: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(to_json(snap), f, indent=2, sort_keys=True)


def parse_run_config(data) -> dict:
    """sim_config::fromJson: the scene plus optional engine and profile blocks."""
    j = json.loads(data) if isinstance(data, (str, bytes)) else data
    cfg = dict(scene=parse_scene(j), engine=dict(backend="METAL", system="Explicit"),
               profile=dict(enabled=False, frames=30, realtime_sync=False, output_path="", level="in_frame"))
    if "engine" in j:
        e = j["engine"]
        if not isinstance(e, dict):
            raise SceneFormatError("'engine' must be a JSON object")
        if "backend" in e:
            if not isinstance(e["backend"], str):
                raise SceneFormatError("engine.backend must be a string")
            if e["backend"] not in KNOWN_BACKENDS:
                raise SceneFormatError(f"unknown engine.backend '{e['backend']}' (expected METAL|CPU|CUDA|HIP)")
            cfg["engine"]["backend"] = e["backend"]
        if "system" in e:
            if not isinstance(e["system"], str):
                raise SceneFormatError("engine.system must be a string")
            if e["system"] not in KNOWN_SYSTEMS:
                raise SceneFormatError(f"unknown engine.system '{e['system']}' (expected Explicit|PBD|PD|LargeSteps)")
            cfg["engine"]["system"] = e["system"]
    if "profile" in j:
        p = j["profile"]
        if not isinstance(p, dict):
            raise SceneFormatError("'profile' must be a JSON object")
        if "enabled" in p:
            if not isinstance(p["enabled"], bool):
                raise SceneFormatError("profile.enabled must be a boolean")
            cfg["profile"]["enabled"] = p["enabled"]
        if "frames" in p:
            if not _is_int(p["frames"]):
                raise SceneFormatError("profile.frames must be an integer")
            if p["frames"] < 1:
                raise SceneFormatError("profile.frames must be >= 1")
            cfg["profile"]["frames"] = p["frames"]
        if "realtime_sync" in p:
            if not isinstance(p["realtime_sync"], bool):
                raise SceneFormatError("profile.realtime_sync must be a boolean")
            cfg["profile"]["realtime_sync"] = p["realtime_sync"]
        if "output_path" in p:
            if not isinstance(p["output_path"], str):
                raise SceneFormatError("profile.output_path must be a string")
            cfg["profile"]["output_path"] = p["output_path"]
        if "level" in p:
            if not isinstance(p["level"], str):
                raise SceneFormatError("profile.level must be a string")
            if p["level"] not in PROFILE_LEVELS:
                raise SceneFormatError(f"unknown profile.level '{p['level']}' (expected none|per_frame|in_frame)")
            cfg["profile"]["level"] = p["level"]
    return cfg


def read_run_config(path: str) -> dict:
    try:
        with open(path, "r") as f:
            text = f.read()
    except OSError:
        raise SceneFormatError(f"cannot open '{path}' for reading") from None
    return parse_run_config(text)


# ------------------------------------------------------------------ geometry
def mt19937(seed: int, n: int) -> np.ndarray:
    """The first n outputs of std::mt19937 seeded with `seed` (init_genrand)."""
    mt = np.zeros(624, np.uint64)
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        prev = int(mt[i - 1])
        mt[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & 0xFFFFFFFF
    bg = np.random.MT19937(0)
    st = bg.state
    st["state"] = dict(key=mt.astype(np.uint32), pos=624)
    bg.state = st
    return bg.random_raw(n).astype(np.uint64)


def jiggle_values(seed: int, n: int) -> np.ndarray:
    """n draws of std::uniform_real_distribution<float>(0, 1e-4) on mt19937(seed):
    generate_canonical<float> = float(u) / 2^32 (held below 1), times
    float(1e-4), plus 0, in float32."""
    u = mt19937(seed, n).astype(np.float32)            # float(u32), round to nearest
    c = (u / np.float32(4294967296.0)).astype(F32)
    c = np.where(c >= F32(1.0), np.nextafter(F32(1.0), F32(0.0)), c).astype(F32)
    return (c * F32(1.0 / 10000.0)).astype(F32)


def grid_geometry(n: int, size: float, direction: str, center, jiggle: bool, seed: int):
    """MeshGridInitializer::initialize (initializers.hpp:47-128), float32."""
    n = int(n)
    half = F32(size) / F32(2.0)
    length = F32(size) / F32(n - 1)
    rows, cols = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    px = (cols.astype(F32) * length - half).astype(F32).reshape(-1)
    py = ((-rows).astype(F32) * length + half).astype(F32).reshape(-1)
    pz = jiggle_values(seed, n * n) if jiggle else np.zeros(n * n, F32)
    c = np.asarray(center, F32)
    x = np.empty((n * n, 3), F32)
    if direction == "XYPlane":
        x[:, 0], x[:, 1], x[:, 2] = px + c[0], py + c[1], pz + c[2]
    elif direction == "YZPlane":
        x[:, 0], x[:, 1], x[:, 2] = pz + c[0], px + c[1], -py + c[2]
    else:
        x[:, 0], x[:, 1], x[:, 2] = px + c[0], pz + c[1], -py + c[2]
    return x, scenes.grid_facets(n)


def _normal_axis(direction: str) -> int:
    return {"XYPlane": 2, "YZPlane": 0}.get(direction, 1)


def sphere_geometry(size: float, tessellation: int, center):
    """primitive::sphere (include/primitive_geometry.hpp): north pole, t - 1
    rings of t vertices, south pole; V = (t-1) t + 2, F = 2 t (t-1)."""
    t = max(3, int(tessellation))
    r = F32(size) * F32(0.5)
    pi = F32(math.pi)
    c = np.asarray(center, F32)
    pts = [(F32(0), r, F32(0))]
    for i in range(1, t):
        th = F32(pi * F32(i) / F32(t))
        sy, sr = F32(np.cos(th) * r), F32(np.sin(th) * r)
        for j in range(t):
            ph = F32(F32(F32(2.0) * pi) * F32(j) / F32(t))
            pts.append((F32(sr * np.cos(ph)), sy, F32(sr * np.sin(ph))))
    pts.append((F32(0), -r, F32(0)))
    x = (np.array(pts, F32) + c[None, :]).astype(F32)
    south = (t - 1) * t + 1
    inner = lambda i, j: 1 + (i - 1) * t + (j % t)
    f = [(0, inner(1, j), inner(1, j + 1)) for j in range(t)]
    for i in range(1, t - 1):
        for j in range(t):
            a, b, cc, d = inner(i, j), inner(i, j + 1), inner(i + 1, j), inner(i + 1, j + 1)
            f += [(a, cc, d), (a, d, b)]
    f += [(south, inner(t - 1, j + 1), inner(t - 1, j)) for j in range(t)]
    return x, np.array(f, np.uint32)


def cube_geometry(size: float, tessellation: int, center):
    """primitive::cube: six t x t faces, then the coincident seam vertices
    welded (quantised to size * 1e-5, first occurrence kept); V = 6 t^2 + 2."""
    t = max(1, int(tessellation))
    s = F32(size)
    h = s * F32(0.5)
    z = F32(0)
    faces = [((h, -h, -h), (z, s, z), (z, z, s)), ((-h, -h, h), (z, s, z), (z, z, -s)),
             ((-h, h, h), (s, z, z), (z, z, -s)), ((-h, -h, -h), (s, z, z), (z, z, s)),
             ((-h, -h, h), (s, z, z), (z, s, z)), ((h, -h, -h), (-s, z, z), (z, s, z))]
    c = np.asarray(center, F32)
    pos, fac = [], []
    for o, u, v in faces:
        base = len(pos)
        for i in range(t + 1):
            vi = F32(F32(i) / F32(t))
            for j in range(t + 1):
                uj = F32(F32(j) / F32(t))
                pos.append(tuple(F32(F32(F32(o[k] + F32(uj * u[k])) + F32(vi * v[k])) + c[k]) for k in range(3)))
        for i in range(t):
            for j in range(t):
                p00 = base + i * (t + 1) + j
                p10, p01 = p00 + 1, p00 + t + 1
                fac += [(p00, p01, p01 + 1), (p00, p01 + 1, p10)]
    q = F32(1.0) / (s * F32(1e-5) if s > 0 else F32(1e-5))

    def llround(v: float) -> int:  # halves away from zero
        r = math.floor(abs(v) + 0.5)
        return int(r) if v >= 0 else -int(r)

    seen, out, remap = {}, [], []
    for p in pos:
        key = tuple(llround(float(k) * float(q)) for k in p)
        if key not in seen:
            seen[key] = len(out)
            out.append(p)
        remap.append(seen[key])
    f = np.array([[remap[a] for a in tri] for tri in fac], np.uint32)
    return np.array(out, F32), f


def cylinder_geometry(size: float, tessellation: int, center):
    """primitive::cylinder: t + 1 rings of t vertices from y = -size/2 to
    +size/2 (radius size/2), then the bottom and top cap centres."""
    t = max(3, int(tessellation))
    s = F32(size)
    r = s * F32(0.5)
    hh = s * F32(0.5)
    pi = F32(math.pi)
    c = np.asarray(center, F32)
    pts = []
    for i in range(t + 1):
        y = F32(-hh + F32(s * F32(i) / F32(t)))
        for j in range(t):
            ph = F32(F32(F32(2.0) * pi) * F32(j) / F32(t))
            pts.append((F32(r * np.cos(ph)), y, F32(r * np.sin(ph))))
    pts += [(F32(0), -hh, F32(0)), (F32(0), hh, F32(0))]
    x = (np.array(pts, F32) + c[None, :]).astype(F32)
    bottom, top = (t + 1) * t, (t + 1) * t + 1
    side = lambda i, j: i * t + (j % t)
    f = []
    for i in range(t):
        for j in range(t):
            a, b, cc, d = side(i, j), side(i, j + 1), side(i + 1, j), side(i + 1, j + 1)
            f += [(a, cc, d), (a, d, b)]
    f += [(bottom, side(0, j + 1), side(0, j)) for j in range(t)]
    f += [(top, side(t, j), side(t, j + 1)) for j in range(t)]
    return x, np.array(f, np.uint32)


def obj_geometry(path: str, scale: float, offset):
    """Vertices and faces of a Wavefront .obj: x = v * scale + offset;
    1-based (or negative, relative) indices; polygons fanned from their first
    corner."""
    verts, faces = [], []
    try:
        lines = open(path, "r").read().splitlines()
    except OSError:
        raise SceneFormatError(f"cannot open '{path}' for reading") from None
    for ln in lines:
        p = ln.split()
        if not p:
            continue
        if p[0] == "v" and len(p) >= 4:
            verts.append([float(p[1]), float(p[2]), float(p[3])])
        elif p[0] == "f" and len(p) >= 4:
            ids = []
            for tok in p[1:]:
                k = int(tok.split("/")[0])
                ids.append(k - 1 if k > 0 else len(verts) + k)
            for q in range(1, len(ids) - 1):
                faces.append((ids[0], ids[q], ids[q + 1]))
    v = np.array(verts, F32).reshape(-1, 3)
    x = (v * F32(scale) + np.asarray(offset, F32)[None, :]).astype(F32)
    f = np.array(faces, np.uint32).reshape(-1, 3)
    if f.size and int(f.max()) >= len(verts):
        raise SceneFormatError(f"face index out of range in '{path}'")
    return x, f


def _rotate(q, d):
    """rotateVector(q, d) for points (n, 3), float32: q * (0, d) * conj(q)."""
    w, qx, qy, qz = (F32(c) for c in q)
    dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
    z = F32(0)
    # r = q * (0, d): Hamilton product, terms summed left to right
    rw = ((w * z - qx * dx) - qy * dy) - qz * dz
    rx = ((w * dx + qx * z) + qy * dz) - qz * dy
    ry = ((w * dy - qx * dz) + qy * z) + qz * dx
    rz = ((w * dz + qx * dy) - qy * dx) + qz * z
    # s = r * conj(q) = r * (w, -qx, -qy, -qz)
    cx, cy, cz = -qx, -qy, -qz
    sx = ((rw * cx + rx * w) + ry * cz) - rz * cy
    sy = ((rw * cy - rx * cz) + ry * w) + rz * cx
    sz = ((rw * cz + rx * cy) - ry * cx) + rz * w
    return np.stack([sx, sy, sz], 1).astype(F32)


def bake_transform(x: np.ndarray, pivot, rot, scale) -> np.ndarray:
    """Scene::pack's transform bake: d = (x - pivot) * scale, rotated, + pivot
    (skipped for unit scale and identity rotation, as the reference does)."""
    s = np.asarray(scale, F32)
    q = np.asarray(rot, F32)
    unit_s = bool(np.all(np.abs(s - F32(1.0)) < F32(1e-7)))
    ident = bool(abs(q[0] - F32(1.0)) < F32(1e-7) and np.all(np.abs(q[1:]) < F32(1e-7)))
    if unit_s and ident:
        return x
    p = np.asarray(pivot, F32)
    d = (x - p[None, :]).astype(F32) * s[None, :]
    if not ident:
        d = _rotate(q, d.astype(F32))
    return (p[None, :] + d).astype(F32)


def fastgrid_rest(x: np.ndarray, n: int, p: dict) -> dict:
    """recomputeFastGridRest (collision_pod.hpp:419-442): the six rest lengths
    measured at the grid's first cell, kept where the distance is <= 1e-9."""
    out = dict(p)

    def dist(a, b):
        d = (x[a] - x[b]).astype(F32)
        return F32(np.sqrt(F32(F32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))

    for key, a, b in (("stretch_rest_x", 0, 1), ("stretch_rest_y", 0, n), ("shear_rest_a", 0, n + 1),
                      ("shear_rest_b", 1, n), ("bend_rest_x", 0, 2), ("bend_rest_y", 0, 2 * n)):
        if a < x.shape[0] and b < x.shape[0]:
            d = dist(a, b)
            if d > F32(1e-9):
                out[key] = float(d)
    return out


def _local_half(o: dict, kind: int, x: np.ndarray) -> np.ndarray:
    """fitAnalyticLocalHalf (scene.hpp:603-641)."""
    s = o["source"]
    match = {1: "sphere", 2: "cube", 3: "cylinder"}.get(kind)
    if s["type"] == "primitive" and s["shape"] == match and s["size"] > 0:
        h = F32(s["size"]) * F32(0.5)
        return (h * np.asarray(o["transform"]["scale"], F32)).astype(F32)
    c = np.asarray(o["transform"]["position"], F32)
    q = np.asarray(o["transform"]["rotation"], F32)
    l = _rotate((q[0], -q[1], -q[2], -q[3]), (x - c[None, :]).astype(F32))
    return np.abs(l).max(axis=0).astype(F32) if len(l) else np.zeros(3, F32)


def build_scene(snap: dict, directory: str = "", margin: float = DEFAULT_MARGIN, radius: float = DEFAULT_RADIUS,
                kair: float = DEFAULT_KAIR, kd: float = DEFAULT_KD, morton_bits: int = 30,
                rebuild_every: Optional[int] = None, self_collision: Optional[int] = None,
                contact_capacity: int = 16) -> scenes.Scene:
    """A packed scene (meshes, analytic colliders, stepping parameters) from
    a parsed scene file; `info` carries the substeps per frame."""
    sim = snap["simulation"]
    h = F32(sim["time_per_frame"])
    subh = float(F32(h / F32(sim["substeps"])))
    env = snap["environment"]
    meshes, shapes = [], []
    for idx, o in enumerate(snap["objects"]):
        s, tr, btype = o["source"], o["transform"], o["behavior"]["type"]
        pos = np.asarray(tr["position"], F32)
        grid_dir = None
        if s["type"] == "primitive":
            if s["shape"] == "sphere":
                x, f = sphere_geometry(s["size"], s["tessellation"], pos)
            elif s["shape"] == "cube":
                x, f = cube_geometry(s["size"], s["tessellation"], pos)
            elif s["shape"] == "cylinder":
                x, f = cylinder_geometry(s["size"], s["tessellation"], pos)
            else:
                grid_dir = s["direction"]
                x, f = grid_geometry(s["tessellation"], s["size"], grid_dir, pos, s["jiggle"], o["id"] & 0xFFFFFFFF)
            mass = s["mass"]
        else:
            x, f = obj_geometry(resolve_import_path(directory, s["path"]), s["scale"], pos)
            mass = s["mass"]
        x = bake_transform(x, pos, tr["rotation"], tr["scale"])
        cloth_b = btype in ("TriangularCloth", "FastGridCloth")
        if cloth_b and grid_dir is not None:
            ax = _normal_axis(grid_dir)
            x = x.copy()
            x[:, ax] = (x[:, ax] + jiggle_values((o["id"] ^ idx) & 0xFFFFFFFF, x.shape[0])).astype(F32)
        prm = o["behavior"]["params"]
        ks = F32(o["cloth_stiffness_scale"])
        if btype == "TriangularCloth":
            cloth = dict(kstretch=float(F32(prm.get("stretch", 0.0)) * ks),
                         kshear=float(F32(prm.get("shear", 0.0)) * ks),
                         kbend=float(F32(prm.get("bend", 0.0)) * ks), thickness=float(prm.get("thickness", 0.0)))
            behavior = scenes.BEHAVIOR_TRIANGULAR_CLOTH
        elif btype == "FastGridCloth":
            n1 = int(prm.get("particle_num_1d", 0))
            if n1 * n1 != x.shape[0]:
                raise SceneFormatError(f"objects[{idx}]: FastGridCloth needs particle_num_1d^2 = {x.shape[0]} vertices")
            legacy = {k: prm.get(k, 0.0) for k in ("stretch_rest", "shear_rest", "bend_rest")}
            cloth = dict(particle_num_1d=n1,
                         stretch_rest_x=prm.get("stretch_rest_x", legacy["stretch_rest"]),
                         stretch_rest_y=prm.get("stretch_rest_y", legacy["stretch_rest"]),
                         shear_rest_a=prm.get("shear_rest_a", legacy["shear_rest"]),
                         shear_rest_b=prm.get("shear_rest_b", legacy["shear_rest"]),
                         bend_rest_x=prm.get("bend_rest_x", legacy["bend_rest"]),
                         bend_rest_y=prm.get("bend_rest_y", legacy["bend_rest"]),
                         kstretch=float(F32(prm.get("k_stretch", 0.0)) * ks),
                         kshear=float(F32(prm.get("k_shear", 0.0)) * ks),
                         kbend=float(F32(prm.get("k_bend", 0.0)) * ks), thickness=float(prm.get("thickness", 0.0)))
            cloth = fastgrid_rest(x, n1, cloth)
            behavior = scenes.BEHAVIOR_FASTGRID_CLOTH
        else:
            cloth = None
            behavior = scenes.BEHAVIOR_FLOAT
        # pins last: the held positions win over jiggle and rest lengths
        # (Scene::pack re-applies them after recomputeFastGridRest)
        fixed = np.ones(x.shape[0], F32)
        for fp in o["fixed_particles"]:
            if 0 <= fp["vid"] < x.shape[0]:
                x = x.copy() if not x.flags.writeable else x
                x[fp["vid"]] = np.asarray(fp["pos"], F32)
                fixed[fp["vid"]] = 0.0
        kind = {"sphere": 1, "cube": 2, "cylinder": 3}.get(s.get("shape"), 0) if s["type"] == "primitive" else 0
        if o["collider_kind"] in COLLIDER_NAMES:
            kind = COLLIDER_NAMES[o["collider_kind"]]
        if kind and behavior != scenes.BEHAVIOR_FLOAT:
            raise SceneFormatError(f"objects[{idx}]: an analytic collider on a cloth is not supported by this backend")
        m = scenes.Mesh(x=x, facets=f, behavior=behavior, cloth=cloth, mass=float(F32(mass)),
                        collidable=o["collidable"], shape=kind, fixed=fixed,
                        apply_gravity=o["apply_gravity"], apply_wind=o["apply_wind"])
        meshes.append(m)
        if kind:
            hl = _local_half(o, kind, x)
            if btype == "Rigid":
                # a Rigid collider follows its vertex centroid (double sums,
                # refreshAnalyticShapes, scene.hpp:675-687)
                pos = (x.astype(np.float64).sum(axis=0) / x.shape[0]).astype(F32)
            q = tuple(float(F32(c)) for c in tr["rotation"])
            half = (float(hl[0]), float(hl[1]), float(hl[2]))
            radius_c = float(max(hl[0], hl[2])) if kind == 3 else (float(hl[0]) if kind == 1 else 0.0)
            shapes.append(scenes.analytic_shape(kind, idx, tuple(float(c) for c in pos),
                                                half=half if kind in (1, 2, 3) else (0.0, 0.0, 0.0),
                                                radius=radius_c, half_height=float(hl[1]) if kind in (1, 3) else 0.0,
                                                rot=q, flags=1 if o["collidable"] else 0))
    if self_collision is None:
        self_collision = int(any(o["self_collide"] for o in snap["objects"]))
    if rebuild_every is None:
        rebuild_every = REBUILD_EVERY_FRAMES * sim["substeps"]
    p = scenes.sim_params(subh=subh, margin=margin, morton_bits=morton_bits, rebuild_every=rebuild_every,
                          self_collision=self_collision, radius=radius, kd=kd, kair=kair,
                          contact_capacity=contact_capacity)
    sc = scenes.Scene(meshes, p, name="scene-file",
                      info=dict(substeps=sim["substeps"]),
                      analytic=np.concatenate(shapes) if shapes else None,
                      gravity=tuple(env["gravity"]), wind=tuple(env["wind"]))
    return sc


def load_scene(path: str, **kw) -> scenes.Scene:
    """read_scene + build_scene, imports resolved against the file's directory."""
    return build_scene(read_scene(path), scene_dir(path), **kw)
