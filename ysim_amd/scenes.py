"""Synthetic scene generators and the packed-scene layout.

The packed layout follows the reference's Scene::pack contract
(include/sim/scene.hpp:783-1182): per-mesh vertex and facet ranges concatenated
into packed float3 arrays (x, v, ext), per-vertex mass and pin mask, and facets
as uint3 with mesh-local vertex ids.

Generators (numpy, float32 arithmetic so the inputs are bit-identical for the
HIP path and the CPU oracle):
  * grid cloth — MeshGridInitializer on the XZ plane
    (include/sim/initializers.hpp:40-127): vertex id = row*N + col, facet order
    per quad alternating the diagonal with (row+col) parity;
  * optional fold of the grid along x = 0.5 into two layers `gap` apart;
  * UV sphere with independent slice and stack counts (the reference's
    primitive is t x t, include/primitive_geometry.hpp:57-140);
  * icosphere of a given subdivision level;
  * triangulated disk (rings of 6r vertices, zipped by angle): an irregular
    manifold mesh for TriangularCloth.
Rest lengths follow recomputeFastGridRest (include/sim/collision_pod.hpp:418-442).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

F32 = np.float32
GRAVITY = (0.0, -9.81, 0.0)  # SceneEnvironment.gravity, include/sim/scene.hpp:187

BEHAVIOR_FLOAT = 0
BEHAVIOR_FASTGRID_CLOTH = 1
BEHAVIOR_TRIANGULAR_CLOTH = 2


@dataclasses.dataclass
class Mesh:
    x: np.ndarray            # (V, 3) float32
    facets: np.ndarray       # (T, 3) uint32, mesh-local ids
    behavior: int
    cloth: Optional[dict] = None   # FastGrid params when behavior == cloth
    mass: float = 1.0        # per-vertex mass
    collidable: bool = True
    group: int = 0
    shape: int = 0
    fixed: Optional[np.ndarray] = None   # (V,) float32 pin mask, 1 = free
    apply_gravity: bool = True           # per-object environment-force gates
    apply_wind: bool = True


@dataclasses.dataclass
class Scene:
    meshes: List[Mesh]
    params: dict
    name: str = ""
    info: dict = dataclasses.field(default_factory=dict)
    analytic: Optional[np.ndarray] = None   # abi.ANALYTIC records (analytic colliders)
    gravity: tuple = GRAVITY
    wind: tuple = (0.0, 0.0, 0.0)

    # -------------------------------------------------------------- packing
    def packed(self):
        xs, vs, ms, es, fx, fs = [], [], [], [], [], []
        g = np.array(self.gravity, dtype=F32)
        w = np.array(self.wind, dtype=F32)
        for m in self.meshes:
            V = m.x.shape[0]
            xs.append(np.ascontiguousarray(m.x, dtype=F32))
            vs.append(np.zeros((V, 3), F32))
            mass = np.full(V, m.mass, F32)
            ms.append(mass)
            if m.behavior != BEHAVIOR_FLOAT:
                # applyEnvironmentForces: ext = gravity * m when the object
                # takes gravity, + wind for cloths that take wind, in float
                # (simulator.hpp:3122-3166)
                e = (g[None, :] * mass[:, None]).astype(F32) if m.apply_gravity else np.zeros((V, 3), F32)
                if m.apply_wind:
                    e = (e + w[None, :]).astype(F32)
                es.append(e)
            else:
                es.append(np.zeros((V, 3), F32))
            fx.append(np.ones(V, F32) if m.fixed is None else m.fixed.astype(F32))
            fs.append(np.ascontiguousarray(m.facets, dtype=np.uint32))
        return dict(
            x=np.concatenate(xs).reshape(-1), v=np.concatenate(vs).reshape(-1),
            mass=np.concatenate(ms), ext=np.concatenate(es).reshape(-1),
            fixed=np.concatenate(fx), facets=np.concatenate(fs).reshape(-1))

    def offsets(self):
        vo = np.zeros(len(self.meshes) + 1, np.uint32)
        fo = np.zeros(len(self.meshes) + 1, np.uint32)
        for i, m in enumerate(self.meshes):
            vo[i + 1] = vo[i] + m.x.shape[0]
            fo[i + 1] = fo[i] + m.facets.shape[0]
        return vo, fo

    @property
    def num_vertices(self):
        return int(sum(m.x.shape[0] for m in self.meshes))

    @property
    def num_facets(self):
        return int(sum(m.facets.shape[0] for m in self.meshes))


# ------------------------------------------------------------------ geometry
def grid_facets(n: int) -> np.ndarray:
    """Facets of an n x n grid, MeshGridInitializer order (initializers.hpp:95-122)."""
    r, c = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    r = r.reshape(-1).astype(np.int64)
    c = c.reshape(-1).astype(np.int64)
    p00 = r * n + c
    p10 = r * n + c + 1
    p01 = (r + 1) * n + c
    p11 = (r + 1) * n + c + 1
    even = ((r + c) & 1) == 0
    f = np.empty((r.size, 2, 3), np.int64)
    f[even, 0] = np.stack([p00, p01, p11], 1)[even]
    f[even, 1] = np.stack([p00, p11, p10], 1)[even]
    f[~even, 0] = np.stack([p00, p01, p10], 1)[~even]
    f[~even, 1] = np.stack([p10, p01, p11], 1)[~even]
    return f.reshape(-1, 3).astype(np.uint32)


def grid_positions(n: int, size: float, origin, rng: np.random.Generator,
                   jitter_rel: float = 1e-3, fold_gap: float = 0.0) -> np.ndarray:
    """XZ-plane grid: x = origin.x + col*l, z = origin.z + row*l, y = origin.y.

    Jitter ~ N(0, (jitter_rel*l)^2) is added along the plane normal (y), in the
    slot the reference's applyClothJiggle uses (initializers.hpp:209-222).
    With fold_gap > 0 the sheet is folded along x = origin.x + size/2 into two
    layers fold_gap apart, joined by a half-cylinder of radius fold_gap/2 whose
    arc length equals the flat length it replaces (springs start near rest).
    """
    l = size / (n - 1)
    col = np.arange(n, dtype=np.float64)
    xf = col * l
    y = np.zeros(n)
    if fold_gap > 0:
        r = fold_gap / 2.0
        L = math.pi * r
        a = size / 2 - L / 2
        b = size / 2 + L / 2
        xs = np.empty(n)
        for i, u in enumerate(xf):
            if u <= a:
                xs[i], y[i] = u, 0.0
            elif u >= b:
                xs[i], y[i] = a - (u - b), 2 * r
            else:
                th = (u - a) / r
                xs[i], y[i] = a + r * math.sin(th), r * (1 - math.cos(th))
        xf = xs
    X = np.empty((n, n, 3), np.float64)
    X[:, :, 0] = origin[0] + xf[None, :]
    X[:, :, 1] = origin[1] + y[None, :]
    X[:, :, 2] = origin[2] + (np.arange(n) * l)[:, None]
    X = X.reshape(-1, 3).astype(F32)
    if jitter_rel > 0:
        jit = (rng.standard_normal(n * n) * (jitter_rel * l)).astype(F32)
        X[:, 1] = (X[:, 1] + jit).astype(F32)
    return X


def fastgrid_rest(x: np.ndarray, n: int, defaults: dict) -> dict:
    """recomputeFastGridRest (include/sim/collision_pod.hpp:418-442), float32."""
    out = dict(defaults)
    V = x.shape[0]

    def dist(a, b):
        d = (x[a] - x[b]).astype(F32)
        return F32(np.sqrt(F32(F32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))

    def setv(key, a, b):
        if a < V and b < V:
            d = dist(a, b)
            if d > F32(1e-9):
                out[key] = float(d)

    setv("stretch_rest_x", 0, 1)
    setv("stretch_rest_y", 0, n)
    setv("shear_rest_a", 0, n + 1)
    setv("shear_rest_b", 1, n)
    setv("bend_rest_x", 0, 2)
    setv("bend_rest_y", 0, 2 * n)
    return out


def cloth_mesh(n: int, size: float, origin, rng, k=(500.0, 50.0, 5.0), mass=1.0,
               thickness=None, fold_gap=0.0, group=0, jitter_rel=1e-3) -> Mesh:
    x = grid_positions(n, size, origin, rng, jitter_rel=jitter_rel, fold_gap=fold_gap)
    l = size / (n - 1)
    defaults = dict(stretch_rest_x=l, stretch_rest_y=l, shear_rest_a=l * math.sqrt(2),
                    shear_rest_b=l * math.sqrt(2), bend_rest_x=2 * l, bend_rest_y=2 * l)
    rest = fastgrid_rest(x, n, defaults)
    cloth = dict(particle_num_1d=n, **rest, kstretch=float(k[0]), kshear=float(k[1]),
                 kbend=float(k[2]), thickness=float(0.5 * l if thickness is None else thickness))
    return Mesh(x=x, facets=grid_facets(n), behavior=BEHAVIOR_FASTGRID_CLOTH, cloth=cloth,
                mass=mass, group=group)


def disk_mesh(rings: int, radius: float, center, rng=None, jitter_rel: float = 1e-3) -> tuple:
    """XZ-plane disk: a centre vertex and rings r = 1..rings of 6r vertices at
    radius r/rings * radius, consecutive rings zipped into triangles in angle
    order (valences vary, unlike a grid).  V = 1 + 3 R (R + 1), F = 6 R^2."""
    cx, cy, cz = (F32(c) for c in center)
    pts = [(cx, cy, cz)]
    rings_idx = [[0]]
    ang = [[0.0]]
    h = F32(radius) / F32(rings)
    for r in range(1, rings + 1):
        n = 6 * r
        a = [2.0 * math.pi * k / n for k in range(n)]
        rr = F32(r) * h
        idx = []
        for k in range(n):
            idx.append(len(pts))
            pts.append((cx + rr * F32(math.cos(a[k])), cy, cz + rr * F32(math.sin(a[k]))))
        rings_idx.append(idx)
        ang.append(a)
    x = np.array(pts, dtype=F32)
    if rng is not None and jitter_rel > 0:
        x[1:, [0, 2]] += (rng.uniform(-1, 1, (x.shape[0] - 1, 2)) * (jitter_rel * float(h))).astype(F32)
    fac = []
    for k in range(6):  # centre fan
        fac.append((0, rings_idx[1][k], rings_idx[1][(k + 1) % 6]))
    for r in range(2, rings + 1):
        inner, outer, ai, ao = rings_idx[r - 1], rings_idx[r], ang[r - 1], ang[r]
        i = j = 0
        ni, no = len(inner), len(outer)
        while i < ni or j < no:
            a_next_i = ai[i + 1] if i + 1 < ni else 2.0 * math.pi
            a_next_o = ao[j + 1] if j + 1 < no else 2.0 * math.pi
            if j < no and (i >= ni or a_next_o <= a_next_i):
                fac.append((inner[i % ni], outer[j], outer[(j + 1) % no]))
                j += 1
            else:
                fac.append((inner[i % ni], outer[j % no], inner[(i + 1) % ni]))
                i += 1
    return x, np.array(fac, dtype=np.uint32)


def tri_cloth_mesh(x, facets, k=(500.0, 50.0, 5.0), mass=1.0, thickness=0.01, group=0) -> Mesh:
    """TriangularCloth: springs along the edges (k[0]) and across interior
    edges (k[2]); k[1] (shear) is carried but unused, as in
    compute_tri_spring_forces."""
    cloth = dict(kstretch=float(k[0]), kshear=float(k[1]), kbend=float(k[2]), thickness=float(thickness))
    return Mesh(x=np.ascontiguousarray(x, dtype=F32), facets=np.ascontiguousarray(facets, dtype=np.uint32),
                behavior=BEHAVIOR_TRIANGULAR_CLOTH, cloth=cloth, mass=mass, group=group)


def uv_sphere(slices: int, stacks: int, radius: float, center) -> tuple:
    """UV sphere: poles + (stacks-1) rings of `slices` vertices.

    V = slices*(stacks-1) + 2, F = 2*slices*(stacks-1) (the closed forms the
    reference's primitive test checks for slices == stacks,
    test/primitive_test.cpp:44-94).
    """
    verts = [(0.0, radius, 0.0)]
    for i in range(1, stacks):
        phi = math.pi * i / stacks
        for j in range(slices):
            th = 2 * math.pi * j / slices
            verts.append((radius * math.sin(phi) * math.cos(th), radius * math.cos(phi),
                          radius * math.sin(phi) * math.sin(th)))
    verts.append((0.0, -radius, 0.0))
    V = np.array(verts, np.float64) + np.asarray(center, np.float64)[None, :]
    fac = []
    bot = len(verts) - 1

    def ring(i, j):
        return 1 + (i - 1) * slices + (j % slices)

    for j in range(slices):
        fac.append((0, ring(1, j + 1), ring(1, j)))
    for i in range(1, stacks - 1):
        for j in range(slices):
            a, b = ring(i, j), ring(i, j + 1)
            c, d = ring(i + 1, j), ring(i + 1, j + 1)
            fac.append((a, b, d))
            fac.append((a, d, c))
    for j in range(slices):
        fac.append((bot, ring(stacks - 1, j), ring(stacks - 1, j + 1)))
    return V.astype(F32), np.array(fac, np.uint32)


def icosphere(level: int, radius: float, center) -> tuple:
    """Icosahedron subdivided `level` times: V = 10*4^L + 2, F = 20*4^L."""
    t = (1 + math.sqrt(5)) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, np.float64)[None, :]
    return V.astype(F32), np.array(faces, np.uint32)


def obstacle(x, f, group=0) -> Mesh:
    return Mesh(x=x, facets=f, behavior=BEHAVIOR_FLOAT, mass=1.0, group=group)


# ------------------------------------------------------------------ analytic colliders
COLLIDER_SPHERE, COLLIDER_BOX, COLLIDER_CYLINDER, COLLIDER_PLANE = 1, 2, 3, 4


def quat_axis_angle(axis, angle: float) -> tuple:
    """Unit quaternion (w, x, y, z) of a rotation by `angle` about `axis`."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(0.5 * angle)
    return (math.cos(0.5 * angle), a[0] * s, a[1] * s, a[2] * s)


def quat_apply(q, v: np.ndarray) -> np.ndarray:
    """Rotate points (n, 3) by the unit quaternion q = (w, x, y, z), float64."""
    w, qv = float(q[0]), np.asarray(q[1:], np.float64)
    v = np.asarray(v, np.float64)
    t = 2.0 * np.cross(qv[None, :], v)
    return v + w * t + np.cross(qv[None, :], t)


def analytic_shape(kind: int, obj_index: int, center, half=(0.0, 0.0, 0.0), radius=0.0, half_height=0.0,
                   rot=(1.0, 0.0, 0.0, 0.0), prev_center=None, prev_rot=None, flags=1,
                   behavior=BEHAVIOR_FLOAT) -> np.ndarray:
    """One ysim_analytic_shape record (AnalyticShape field meaning,
    include/sim/scene.hpp:717-737): sphere semi-axes in `half`, box
    half-extents in `half`, cylinder `radius` / `half_height` about local +Y,
    plane: centre and rotation only (outside = local +Y)."""
    from . import abi
    r = np.zeros(1, abi.ANALYTIC)
    c = np.asarray(center, np.float32)
    r["center_radius"][0] = (c[0], c[1], c[2], radius)
    r["half_ext_height"][0] = (half[0], half[1], half[2], half_height)
    r["rot_quat"][0] = rot
    pc = c if prev_center is None else np.asarray(prev_center, np.float32)
    r["prev_center_pad"][0] = (pc[0], pc[1], pc[2], 0.0)
    r["prev_rot_quat"][0] = rot if prev_rot is None else prev_rot
    r["kind"] = kind
    r["obj_index"] = obj_index
    r["behavior_type"] = behavior
    r["flags"] = flags
    return r


def box_mesh(half, center, rot=(1.0, 0.0, 0.0, 0.0)) -> tuple:
    """Closed box surface: 8 corners, 12 facets."""
    h = np.asarray(half, np.float64)
    c = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64) * h
    x = quat_apply(rot, c) + np.asarray(center, np.float64)[None, :]
    f = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                  [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)
    return x.astype(F32), f


def cylinder_mesh(radius: float, half_height: float, center, rot=(1.0, 0.0, 0.0, 0.0), slices: int = 24) -> tuple:
    """Capped cylinder about local +Y: two rings of `slices` vertices and two cap centres."""
    th = 2 * np.pi * np.arange(slices) / slices
    ring = np.stack([radius * np.cos(th), np.zeros(slices), radius * np.sin(th)], 1)
    top = ring + [0.0, half_height, 0.0]
    bot = ring - [0.0, half_height, 0.0]
    v = np.concatenate([top, bot, [[0.0, half_height, 0.0], [0.0, -half_height, 0.0]]])
    ct, cb = 2 * slices, 2 * slices + 1
    f = []
    for j in range(slices):
        k = (j + 1) % slices
        f += [(j, k, slices + k), (j, slices + k, slices + j), (ct, k, j), (cb, slices + j, slices + k)]
    x = quat_apply(rot, v) + np.asarray(center, np.float64)[None, :]
    return x.astype(F32), np.array(f, np.uint32)


def plane_mesh(half_size: float, center, rot=(1.0, 0.0, 0.0, 0.0)) -> tuple:
    """A square of the plane (local y = 0), 4 vertices, 2 facets."""
    v = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float64) * half_size
    x = quat_apply(rot, v) + np.asarray(center, np.float64)[None, :]
    return x.astype(F32), np.array([[0, 2, 1], [0, 3, 2]], np.uint32)


def collider_mesh(x, f, kind: int, group: int = 0) -> Mesh:
    """A static mesh whose collider kind (shape lane) is `kind`."""
    return Mesh(x=x, facets=f, behavior=BEHAVIOR_FLOAT, mass=1.0, group=group, shape=kind)


# ------------------------------------------------------------------ params
def sim_params(subh, margin, morton_bits=30, rebuild_every=0, self_collision=1, swept=1,
               broad_mode=0, hash_slots_log2=22, hash_cell=0.0, radius=0.0, kd=0.01,
               kair=-0.001, contact_capacity=16):
    """SymplecticSystem + Simulator knobs.  thickness = margin = band (radius 0)."""
    return dict(subh=float(subh), kair=float(kair), kd=float(kd), margin=float(margin),
                radius=float(radius), narrow_thickness=float(margin), morton_bits=morton_bits,
                rebuild_every=rebuild_every, self_collision=self_collision, swept=swept,
                broad_mode=broad_mode, hash_slots_log2=hash_slots_log2, hash_cell=float(hash_cell),
                contact_capacity=contact_capacity)


# Per-vertex mass.  The reference's `mass` knob is per vertex
# (simulator.hpp:706-713).  A total mass of 1 spread over the grid makes
# dt*sqrt(k/m) = 9.5 at 512^2, dt = 1/1200, k = 500, far beyond the symplectic
# Euler stability bound of 2; the sheet explodes within a few substeps.  The
# configs therefore use mass 1 per vertex (the blow-up at the literal setting
# is reproducible with vertex_mass="total").
def vertex_mass(n_vertices: int, mode: str = "per_vertex") -> float:
    return 1.0 / n_vertices if mode == "total" else 1.0


# ------------------------------------------------------------------ configs
def config_c1(seed: int = 1, n: int = 64, mass_mode: str = "per_vertex", slices=32, stacks=16):
    """Plain CPU-runnable case: 64x64 cloth over a UV sphere, 30-bit LBVH, refit."""
    rng = np.random.default_rng(seed)
    l = 1.0 / (n - 1)
    cl = cloth_mesh(n, 1.0, (0.0, 1.0, 0.0), rng, mass=vertex_mass(n * n, mass_mode))
    sx, sf = uv_sphere(slices, stacks, 0.3, (0.5, 0.5, 0.5))
    p = sim_params(subh=1 / 600, margin=0.5 * l, morton_bits=30, rebuild_every=0)
    return Scene([cl, obstacle(sx, sf)], p, name="c1-64x64-cloth-uv-sphere",
                 info=dict(substeps=20))


def config_c2(seed: int = 2, n: int = 512, mass_mode: str = "per_vertex", gap=0.02):
    """512x512 self-colliding folded cloth, 63-bit LBVH rebuilt every substep."""
    rng = np.random.default_rng(seed)
    l = 1.0 / (n - 1)
    cl = cloth_mesh(n, 1.0, (0.0, 1.0, 0.0), rng, mass=vertex_mass(n * n, mass_mode),
                    fold_gap=gap)
    p = sim_params(subh=1 / 1200, margin=0.5 * l, morton_bits=63, rebuild_every=1)
    return Scene([cl], p, name=f"c2-{n}x{n}-folded-self-collision", info=dict(substeps=100))


def config_drape(seed: int = 3, n: int = 32, height=0.82, subh=1 / 600, bits=30, rebuild_every=0,
                 slices=24, stacks=12, radius=0.3):
    """Small cloth placed just above a sphere so contacts fire within a few substeps."""
    rng = np.random.default_rng(seed)
    l = 1.0 / (n - 1)
    cl = cloth_mesh(n, 1.0, (0.0, height, 0.0), rng, mass=1.0)
    sx, sf = uv_sphere(slices, stacks, radius, (0.5, 0.5, 0.5))
    p = sim_params(subh=subh, margin=0.5 * l, morton_bits=bits, rebuild_every=rebuild_every)
    return Scene([cl, obstacle(sx, sf)], p, name=f"drape-{n}")


def config_tri_drape(seed: int = 6, rings: int = 24, radius: float = 0.45, height: float = 0.83,
                     subh=1 / 600, bits=30, rebuild_every=1, slices=24, stacks=12, sphere_radius=0.3):
    """TriangularCloth disk (rings of 6r vertices) just above a UV sphere, self
    collision on; thickness = margin = band = half the ring spacing; 32
    contact slots per vertex."""
    rng = np.random.default_rng(seed)
    x, f = disk_mesh(rings, radius, (0.5, height, 0.5), rng)
    h = radius / rings
    cl = tri_cloth_mesh(x, f, thickness=0.5 * h)
    sx, sf = uv_sphere(slices, stacks, sphere_radius, (0.5, 0.5, 0.5))
    # the disk centre lands on the sphere's pole, where 24 facets meet
    p = sim_params(subh=subh, margin=0.5 * h, morton_bits=bits, rebuild_every=rebuild_every, contact_capacity=32)
    return Scene([cl, obstacle(sx, sf)], p, name=f"tri-drape-{rings}")


def config_analytic(seed: int = 7, n: int = 48, height: float = 0.8, subh=1 / 600, rebuild_every=1,
                    kinds=("sphere", "box", "cylinder", "plane"), mesh_obstacle: bool = True,
                    self_collision: int = 1) -> Scene:
    """An n x n FastGrid cloth falling onto analytic colliders: an ellipsoid,
    a rotated box, a tilted capped cylinder and a tilted half-space below them,
    next to a triangle-mesh icosphere that keeps the triangle path in play.
    Each collider also has a triangle mesh of its shape (shape lane = kind)."""
    rng = np.random.default_rng(seed)
    l = 1.0 / (n - 1)
    meshes = [cloth_mesh(n, 1.0, (0.0, height, 0.0), rng, mass=1.0)]
    shapes = []
    if "sphere" in kinds:
        c, e = (0.28, 0.55, 0.28), (0.15, 0.1, 0.13)
        sx, sf = uv_sphere(24, 12, 1.0, (0.0, 0.0, 0.0))
        meshes.append(collider_mesh((sx * np.asarray(e, F32) + np.asarray(c, F32)).astype(F32), sf, COLLIDER_SPHERE))
        shapes.append(analytic_shape(COLLIDER_SPHERE, len(meshes) - 1, c, half=e))
    if "box" in kinds:
        c, h, q = (0.72, 0.56, 0.28), (0.12, 0.08, 0.1), quat_axis_angle((0.3, 1.0, 0.2), 0.6)
        meshes.append(collider_mesh(*box_mesh(h, c, q), COLLIDER_BOX))
        shapes.append(analytic_shape(COLLIDER_BOX, len(meshes) - 1, c, half=h, rot=q))
    if "cylinder" in kinds:
        c, q = (0.28, 0.55, 0.72), quat_axis_angle((1.0, 0.0, 0.3), 0.45)
        meshes.append(collider_mesh(*cylinder_mesh(0.11, 0.09, c, q), COLLIDER_CYLINDER))
        shapes.append(analytic_shape(COLLIDER_CYLINDER, len(meshes) - 1, c, radius=0.11, half_height=0.09, rot=q))
    if "plane" in kinds:
        c, q = (0.5, 0.38, 0.5), quat_axis_angle((0.2, 0.0, 1.0), 0.12)
        meshes.append(collider_mesh(*plane_mesh(1.5, c, q), COLLIDER_PLANE))
        shapes.append(analytic_shape(COLLIDER_PLANE, len(meshes) - 1, c, rot=q))
    if mesh_obstacle:
        meshes.append(obstacle(*icosphere(2, 0.12, (0.72, 0.5, 0.72))))
    p = sim_params(subh=subh, margin=0.5 * l, morton_bits=30, rebuild_every=rebuild_every,
                   self_collision=self_collision)
    return Scene(meshes, p, name=f"analytic-{n}", analytic=np.concatenate(shapes) if shapes else None)


def config_c3(seed: int = 3, count: int = 256, lo=16, hi=256, mass_mode="per_vertex", substeps=50):
    """Batched: `count` cloth instances, side ~ log-uniform[lo, hi], each with its own
    UV sphere (r ~ U(0.15, 0.4), centre ~ U([0.2, 0.8]^2) x U(0.3, 0.7)) and its own
    k_struct ~ U(200, 1000).  Instance i is meshes (2i, 2i+1) in collision group i;
    margin <= 0 gives every instance its own margin = band = thickness (0.5 spacing)."""
    rng = np.random.default_rng(seed)
    meshes = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        n = max(lo, min(hi, n))
        r = float(rng.uniform(0.15, 0.4))
        c = (float(rng.uniform(0.2, 0.8)), float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.2, 0.8)))
        ks = float(rng.uniform(200.0, 1000.0))
        cl = cloth_mesh(n, 1.0, (0.0, 1.0, 0.0), rng, k=(ks, 50.0, 5.0),
                        mass=vertex_mass(n * n, mass_mode), group=i)
        sx, sf = uv_sphere(32, 16, r, c)
        meshes += [cl, obstacle(sx, sf, group=i)]
    p = sim_params(subh=1 / 600, margin=-1.0, morton_bits=30, rebuild_every=0)
    return Scene(meshes, p, name=f"c3-batched-{count}", info=dict(substeps=substeps))


def instance_scene(batched: Scene, i: int) -> Scene:
    """Instance i of a batched scene as a plain two-mesh scene (same params)."""
    import copy
    cl = copy.copy(batched.meshes[2 * i])
    sp = copy.copy(batched.meshes[2 * i + 1])
    cl.group = 0
    sp.group = 0
    return Scene([cl, sp], dict(batched.params), name=f"{batched.name}-instance-{i}")


def config_c4(seed: int = 4, n: int = 128, height: float = 0.81, steps: int = 60, checkpoint_every: int = 10):
    """Differentiation: n x n cloth dropped onto a UV sphere (top at y = 0.8) from
    `height`, `steps` substeps at 1/600 with checkpoints every `checkpoint_every`.
    Contacts are against the static sphere only (no self-collision), the case the
    adjoint covers.  Hidden per-spring log-stiffness ~ N(log 500, 0.3^2) defines
    the target drape; the loss is 0.5 |x_N - x_target|^2."""
    rng = np.random.default_rng(seed)
    l = 1.0 / (n - 1)
    cl = cloth_mesh(n, 1.0, (0.0, height, 0.0), rng, mass=1.0)
    sx, sf = uv_sphere(32, 16, 0.3, (0.5, 0.5, 0.5))
    p = sim_params(subh=1 / 600, margin=0.5 * l, morton_bits=30, rebuild_every=0, self_collision=0)
    return Scene([cl, obstacle(sx, sf)], p, name=f"c4-{n}x{n}-adjoint",
                 info=dict(substeps=steps, checkpoint_every=checkpoint_every, log_k_mean=math.log(500.0),
                           log_k_std=0.3, fd_eps=1e-3, fd_springs=64))


def c4_theta(scene, seed: int = 44):
    """(theta_target, theta_0): hidden per-spring log-k and the uniform start."""
    n = int(scene.meshes[0].cloth["particle_num_1d"])
    S = n * (n - 1) * 2 + 2 * (n - 1) * (n - 1) + 2 * n * (n - 2)
    rng = np.random.default_rng(seed)
    mu, sd = scene.info["log_k_mean"], scene.info["log_k_std"]
    tgt = (mu + sd * rng.standard_normal(S)).astype(F32)
    return tgt, np.full(S, mu, F32)


def config_c5(seed: int = 5, n: int = 1024, level: int = 6, broad_mode: int = 0, height: float = 1.0):
    """1024x1024 self-colliding cloth over 4 icospheres (level 6), 63-bit LBVH or hash."""
    rng = np.random.default_rng(seed)
    l = 1.0 / (n - 1)
    cl = cloth_mesh(n, 1.0, (0.0, height, 0.0), rng, mass=1.0)
    meshes = [cl]
    for _ in range(4):
        r = float(rng.uniform(0.1, 0.25))
        c = (float(rng.uniform(0.2, 0.8)), float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.2, 0.8)))
        ix, ifc = icosphere(level, r, c)
        meshes.append(obstacle(ix, ifc))
    thick = 0.5 * l
    p = sim_params(subh=1 / 2400, margin=thick, morton_bits=63, rebuild_every=1,
                   broad_mode=broad_mode, hash_slots_log2=22, hash_cell=2 * thick)
    return Scene(meshes, p, name=f"c5-{n}x{n}-4-icospheres", info=dict(substeps=50))
