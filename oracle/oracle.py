"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from ysim_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
vp = C.c_void_p


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=HERE, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    sig = {
        "yo_set_threads": (C.c_int, [C.c_int]),
        "yo_scene_box": (None, [f32p, C.c_uint32, f32p]),
        "yo_expand10": (C.c_uint32, [C.c_uint32]),
        "yo_expand21": (C.c_uint64, [C.c_uint64]),
        "yo_morton": (None, [f32p, u32p, C.c_uint32, f32p, C.c_uint32, u64p, u32p]),
        "yo_sort_pairs": (None, [u64p, u32p, C.c_uint32, C.c_uint32]),
        "yo_karras": (None, [u64p, u32p, C.c_uint32, C.c_uint32, vp, C.POINTER(C.c_int32)]),
        "yo_refit": (None, [f32p, f32p, C.c_float, C.c_uint32, u32p, C.c_uint32, vp]),
        "yo_broad_query": (C.c_uint32, [f32p, C.c_uint32, C.c_uint32, vp, C.c_uint32, u32p,
                                        C.c_uint32, C.c_float, u32p, C.c_uint32]),
        "yo_brute_candidates": (C.c_uint32, [f32p, C.c_uint32, C.c_uint32, f32p, C.c_uint32, u32p,
                                             C.c_uint32, C.c_float, u32p, C.c_uint32]),
        "yo_narrow_f32": (None, [u32p, C.c_uint32, f32p, f32p, f32p, u32p, C.c_uint32, C.c_float,
                                 C.c_float, u8p, f32p]),
        "yo_narrow_f64": (None, [u32p, C.c_uint32, f64p, f64p, f64p, u32p, C.c_uint32, C.c_double,
                                 C.c_double, u8p, f64p]),
        "yo_fastgrid_forces_f32": (None, [vp, C.c_uint32, C.c_float, C.c_float, f32p, f32p, f32p,
                                          f32p, f32p]),
        "yo_fastgrid_forces_f64": (None, [vp, C.c_uint32, C.c_double, C.c_double, f64p, f64p, f64p,
                                          f64p, f64p]),
        "yo_spring_count": (C.c_uint32, [C.c_uint32]),
        "yo_tri_adjacency": (C.c_int, [u32p, C.c_uint32, f32p, C.c_uint32, u32p, u32p, f32p, u32p, u32p, u32p,
                                       u32p, f32p]),
        "yo_tri_forces_f32": (C.c_int, [u32p, C.c_uint32, f32p, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                        C.c_float, f32p, f32p, f32p, f32p]),
        "yo_tri_forces_f64": (C.c_int, [u32p, C.c_uint32, f32p, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                        C.c_double, f64p, f64p, f64p, f64p]),
        "yo_integrate_f32": (C.c_uint32, [C.c_uint32, C.c_float, C.c_float, f32p, f32p, f32p, f32p,
                                          f32p, u32p, vp]),
        "yo_scene_create": (vp, [C.c_uint32, vp, f32p, f32p, f32p, f32p, f32p, u32p, vp]),
        "yo_scene_destroy": (None, [vp]),
        "yo_scene_enable_twin": (None, [vp]),
        "yo_scene_enable_ties": (None, [vp]),
        "yo_scene_set_collision": (None, [vp, C.c_float, C.c_float, C.c_float, C.c_uint32]),
        "yo_scene_ties": (None, [vp, u64p]),
        "yo_scene_record_times": (None, [vp, C.c_int]),
        "yo_scene_step_times": (C.c_uint32, [vp, f64p, C.c_uint32]),
        "yo_tie_bits": (C.c_uint32, [f32p, f32p, f32p, f32p, f32p, C.c_float, C.c_int]),
        "yo_scene_set_kspring": (None, [vp, f32p, C.c_uint32]),
        "yo_scene_step": (None, [vp, C.c_uint32]),
        "yo_scene_pairs": (C.c_uint64, [vp, u32p, C.c_uint64]),
        "yo_scene_get_state": (None, [vp, f32p, f32p, f32p]),
        "yo_scene_set_state": (None, [vp, f32p, f32p, f32p]),
        "yo_scene_get_twin": (None, [vp, f64p, f64p]),
        "yo_scene_num_contacts": (C.c_uint32, [vp]),
        "yo_scene_get_contacts": (None, [vp, u32p, vp]),
        "yo_scene_get_tree": (None, [vp, C.c_uint32, vp, C.POINTER(C.c_int32), u64p, u32p]),
        "yo_scene_anomaly": (C.c_uint32, [vp]),
        "yo_scene_set_analytic": (C.c_int, [vp, vp, C.c_uint32]),
        "yo_scene_set_build_mode": (None, [vp, C.c_uint32]),
        "yo_agglomerative": (None, [u64p, u32p, C.c_uint32, C.c_uint32, vp, C.POINTER(C.c_int32)]),
        "yo_scene_analytic_markers": (C.c_uint32, [vp, u32p, u8p]),
        "yo_narrow_analytic_f32": (C.c_int, [vp, f32p, f32p, C.c_float, f32p]),
        "yo_narrow_analytic_f64": (C.c_int, [vp, f64p, f64p, C.c_double, f64p]),
        "yo_analytic_bounds": (None, [vp, C.c_float, C.c_float, f32p, f32p, f32p, f64p]),
        "yo_adj_create": (vp, [C.c_uint32, vp, vp, f32p, f32p, f32p, f32p, f32p, C.c_uint32, f32p, u32p,
                               u32p]),
        "yo_adj_destroy": (None, [C.c_uint32, vp]),
        "yo_adj_num_springs": (C.c_uint32, [C.c_uint32]),
        "yo_adj_forward": (None, [C.c_uint32, vp, f64p, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, f64p,
                                  f64p]),
        "yo_adj_contact_hist": (None, [C.c_uint32, vp, u32p]),
        "yo_adj_copy_record": (None, [C.c_uint32, vp, C.c_uint32, vp]),
        "yo_adj_set_initial": (None, [C.c_uint32, vp, f64p, f64p]),
        "yo_adj_plant": (None, [C.c_uint32, vp, C.c_uint32]),
        "yo_adj_backward": (None, [C.c_uint32, vp, f64p, f64p, f64p, f64p, f64p, f64p]),
        "yo_adj_contacts": (None, [C.c_uint32, vp, u64p, u64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def P(a, t):
    return None if a is None else a.ctypes.data_as(t)


def set_threads(n: int) -> int:
    return lib().yo_set_threads(n)


# ------------------------------------------------------------------ stages
def scene_box(x):
    out = np.zeros(6, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    lib().yo_scene_box(P(x, f32p), x.size // 3, P(out, f32p))
    return out


def morton(x, facets, box6, bits):
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    facets = np.ascontiguousarray(facets, np.uint32).reshape(-1)
    n = facets.size // 3
    keys = np.zeros(n, np.uint64)
    idx = np.zeros(n, np.uint32)
    box6 = np.ascontiguousarray(box6, np.float32)
    lib().yo_morton(P(x, f32p), P(facets, u32p), n, P(box6, f32p), bits, P(keys, u64p), P(idx, u32p))
    return keys, idx


def sort_pairs(keys, vals, key_bits):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().yo_sort_pairs(P(k, u64p), P(v, u32p), k.size, key_bits)
    return k, v


def karras(keys, idx, bits):
    n = keys.size
    tree = np.zeros(max(2 * n - 1, 0), abi.BVH_NODE)
    parent = np.full(max(2 * n - 1, 0), -1, np.int32)
    k = np.ascontiguousarray(keys, np.uint64)
    i = np.ascontiguousarray(idx, np.uint32)
    lib().yo_karras(P(k, u64p), P(i, u32p), n, bits, tree.ctypes.data, P(parent, C.POINTER(C.c_int32)))
    return tree, parent


def agglomerative(keys, idx, bits):
    """Apetrei hierarchy (topology, root in slot 0) over sorted keys."""
    n = keys.size
    tree = np.zeros(max(2 * n - 1, 0), abi.BVH_NODE)
    parent = np.full(max(2 * n - 1, 0), -1, np.int32)
    k = np.ascontiguousarray(keys, np.uint64)
    i = np.ascontiguousarray(idx, np.uint32)
    lib().yo_agglomerative(P(k, u64p), P(i, u32p), n, bits, tree.ctypes.data, P(parent, C.POINTER(C.c_int32)))
    return tree, parent


def refit(tree, x, v, dt, swept, facets):
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    v = np.ascontiguousarray(v, np.float32).reshape(-1)
    facets = np.ascontiguousarray(facets, np.uint32).reshape(-1)
    n = facets.size // 3
    lib().yo_refit(P(x, f32p), P(v, f32p), dt, swept, P(facets, u32p), n, tree.ctypes.data)
    return tree


def broad_query(xq, qobj, tree, tfacets, tobj, margin):
    xq = np.ascontiguousarray(xq, np.float32).reshape(-1)
    tf = np.ascontiguousarray(tfacets, np.uint32).reshape(-1)
    nq = xq.size // 3
    L = lib()
    cnt = L.yo_broad_query(P(xq, f32p), nq, qobj, tree.ctypes.data, tf.size // 3, P(tf, u32p), tobj,
                           margin, None, 0)
    out = np.zeros(2 * cnt, np.uint32)
    L.yo_broad_query(P(xq, f32p), nq, qobj, tree.ctypes.data, tf.size // 3, P(tf, u32p), tobj,
                     margin, P(out, u32p), cnt)
    return out.reshape(-1, 2)


def brute_candidates(xq, qobj, leafbox6, tfacets, tobj, margin):
    xq = np.ascontiguousarray(xq, np.float32).reshape(-1)
    lb = np.ascontiguousarray(leafbox6, np.float32).reshape(-1)
    tf = np.ascontiguousarray(tfacets, np.uint32).reshape(-1)
    L = lib()
    cnt = L.yo_brute_candidates(P(xq, f32p), xq.size // 3, qobj, P(lb, f32p), tf.size // 3,
                                P(tf, u32p), tobj, margin, None, 0)
    out = np.zeros(2 * cnt, np.uint32)
    L.yo_brute_candidates(P(xq, f32p), xq.size // 3, qobj, P(lb, f32p), tf.size // 3, P(tf, u32p),
                          tobj, margin, P(out, u32p), cnt)
    return out.reshape(-1, 2)


def narrow(pairs, xq, xqprev, xt, tfacets, same_obj, radius, thickness, f64=False):
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1)
    tf = np.ascontiguousarray(tfacets, np.uint32).reshape(-1)
    n = pairs.size // 2
    flags = np.zeros(n, np.uint8)
    if f64:
        arrs = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in (xq, xqprev, xt)]
        nd4 = np.zeros(4 * n, np.float64)
        lib().yo_narrow_f64(P(pairs, u32p), n, P(arrs[0], f64p), P(arrs[1], f64p), P(arrs[2], f64p),
                            P(tf, u32p), int(same_obj), radius, thickness, P(flags, u8p), P(nd4, f64p))
    else:
        arrs = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (xq, xqprev, xt)]
        nd4 = np.zeros(4 * n, np.float32)
        lib().yo_narrow_f32(P(pairs, u32p), n, P(arrs[0], f32p), P(arrs[1], f32p), P(arrs[2], f32p),
                            P(tf, u32p), int(same_obj), radius, thickness, P(flags, u8p), P(nd4, f32p))
    return flags.astype(bool), nd4.reshape(-1, 4)


def fastgrid_forces(cloth: dict, x, v, ext, kair, kd, kspring=None, f64=False):
    prm = abi.fastgrid_params(cloth)
    dt = np.float64 if f64 else np.float32
    tp = f64p if f64 else f32p
    x, v, ext = (np.ascontiguousarray(a, dt).reshape(-1) for a in (x, v, ext))
    ks = None if kspring is None else np.ascontiguousarray(kspring, dt)
    f = np.zeros_like(x)
    fn = lib().yo_fastgrid_forces_f64 if f64 else lib().yo_fastgrid_forces_f32
    fn(C.byref(prm), x.size // 3, kair, kd, P(x, tp), P(v, tp), P(ext, tp), P(ks, tp), P(f, tp))
    return f


def tri_adjacency(facets, x):
    """TriangularCloth adjacency (MeshAdjacencyInitializer): dict of edges (E, 2),
    rest_edge, adj_off/adj, opp_off/opp, rest_opp; None for a rejected mesh."""
    F = np.ascontiguousarray(facets, np.uint32).reshape(-1)
    X = np.ascontiguousarray(x, np.float32).reshape(-1)
    nF, nV = F.size // 3, X.size // 3
    cnt = np.zeros(3, np.uint32)
    if lib().yo_tri_adjacency(P(F, u32p), nF, P(X, f32p), nV, P(cnt, u32p), None, None, None, None, None,
                              None, None) != 0:
        return None
    nE, nA, nO = (int(c) for c in cnt)
    out = dict(edges=np.zeros(2 * nE, np.uint32), rest_edge=np.zeros(nE, np.float32),
               adj_off=np.zeros(nV + 1, np.uint32), adj=np.zeros(nA, np.uint32),
               opp_off=np.zeros(nV + 1, np.uint32), opp=np.zeros(nO, np.uint32), rest_opp=np.zeros(nO, np.float32))
    lib().yo_tri_adjacency(P(F, u32p), nF, P(X, f32p), nV, P(cnt, u32p), P(out["edges"], u32p),
                           P(out["rest_edge"], f32p), P(out["adj_off"], u32p), P(out["adj"], u32p),
                           P(out["opp_off"], u32p), P(out["opp"], u32p), P(out["rest_opp"], f32p))
    out["edges"] = out["edges"].reshape(-1, 2)
    return out


def tri_forces(facets, x0, x, v, ext, kstretch, kbend, kair, kd, f64=False):
    """compute_tri_spring_forces at (x, v), rest lengths from x0."""
    F = np.ascontiguousarray(facets, np.uint32).reshape(-1)
    X0 = np.ascontiguousarray(x0, np.float32).reshape(-1)
    t, tp = (np.float64, f64p) if f64 else (np.float32, f32p)
    xs, vs, es = (np.ascontiguousarray(a, t).reshape(-1) for a in (x, v, ext))
    out = np.zeros_like(xs)
    fn = lib().yo_tri_forces_f64 if f64 else lib().yo_tri_forces_f32
    rc = fn(P(F, u32p), F.size // 3, P(X0, f32p), X0.size // 3, kstretch, kbend, kair, kd, P(xs, tp), P(vs, tp),
            P(es, tp), P(out, tp))
    if rc != 0:
        raise ValueError("tri_forces: mesh rejected (repeated vertex in a facet or non-manifold edge)")
    return out.reshape(-1, 3)


def tie_bits(qc, qp, t0, t1, t2, band: float, fired32: bool) -> int:
    """Threshold-tie bits of one candidate: 1 = FP32 and FP64 narrow decisions
    differ, 2 = FP64 distance within 4 FP32 ulps of the band."""
    a = [np.ascontiguousarray(t, np.float32).reshape(3) for t in (qc, qp, t0, t1, t2)]
    return int(lib().yo_tie_bits(*(P(t, f32p) for t in a), float(band), int(bool(fired32))))


def spring_count(n: int) -> int:
    return lib().yo_spring_count(n)


# ------------------------------------------------------------------ scene
def narrow_analytic(shape, pcur, pprev, band, f64: bool = False):
    """narrow_pt_analytic for one vertex: (normal, d) or None."""
    sh = np.ascontiguousarray(shape, abi.ANALYTIC).reshape(1)
    t, fn, ct = (np.float64, lib().yo_narrow_analytic_f64, f64p) if f64 else \
        (np.float32, lib().yo_narrow_analytic_f32, f32p)
    a = np.ascontiguousarray(pcur, t).reshape(3)
    b = np.ascontiguousarray(pprev, t).reshape(3)
    out = np.zeros(4, t)
    if not fn(sh.ctypes.data, P(a, ct), P(b, ct), band, P(out, ct)):
        return None
    return out[:3], out[3]


def analytic_bounds(shape, radius=0.0, thickness=0.0):
    """Unpadded swept box (lo, hi), broad rotation band, narrow rotation margin,
    narrow (theta, rmax)."""
    sh = np.ascontiguousarray(shape, abi.ANALYTIC).reshape(1)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    o2 = np.zeros(2, np.float32)
    t2 = np.zeros(2, np.float64)
    lib().yo_analytic_bounds(sh.ctypes.data, radius, thickness, P(lo, f32p), P(hi, f32p), P(o2, f32p), P(t2, f64p))
    return lo, hi, float(o2[0]), float(o2[1]), (float(t2[0]), float(t2[1]))


class OracleScene:
    """CPU twin of the HIP scene: same packed inputs, same substep."""

    def __init__(self, scene, twin: bool = False, kspring=None):
        self.scene = scene
        pk = scene.packed()
        self._keep = pk
        self.descs = abi.mesh_descs(scene)
        self.params = abi.sim_params(scene.params)
        L = lib()
        self.h = L.yo_scene_create(len(scene.meshes), C.cast(self.descs, vp), P(pk["x"], f32p),
                                   P(pk["v"], f32p), P(pk["mass"], f32p), P(pk["ext"], f32p),
                                   P(pk["fixed"], f32p), P(pk["facets"], u32p), C.byref(self.params))
        if not self.h:
            raise ValueError("oracle scene rejected (a triangular cloth is not a manifold mesh)")
        if kspring is not None:
            ks = np.ascontiguousarray(kspring, np.float32)
            L.yo_scene_set_kspring(self.h, P(ks, f32p), ks.size)
        if twin:
            L.yo_scene_enable_twin(self.h)
        self.V = scene.num_vertices
        if scene.analytic is not None and len(scene.analytic):
            self.set_analytic(scene.analytic)

    def set_analytic(self, shapes):
        a = np.ascontiguousarray(shapes, abi.ANALYTIC).reshape(-1)
        if lib().yo_scene_set_analytic(self.h, a.ctypes.data if a.size else None, a.size) != 0:
            raise ValueError("oracle: analytic shapes rejected")

    def set_build_mode(self, mode: int):
        """Builder of later dynamic rebuilds: 0 Karras, 1 agglomerative."""
        lib().yo_scene_set_build_mode(self.h, mode)

    def analytic_markers(self):
        """(cloth mesh, shape mesh) pairs and whether each was paired at the last substep."""
        L = lib()
        n = L.yo_scene_analytic_markers(self.h, None, None)
        pairs = np.zeros(2 * n, np.uint32)
        marked = np.zeros(n, np.uint8)
        L.yo_scene_analytic_markers(self.h, P(pairs, u32p), P(marked, u8p))
        return pairs.reshape(-1, 2), marked.astype(bool)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().yo_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def step(self, n: int = 1):
        lib().yo_scene_step(self.h, n)

    def state(self):
        x = np.zeros(3 * self.V, np.float32)
        v = np.zeros(3 * self.V, np.float32)
        xp = np.zeros(3 * self.V, np.float32)
        lib().yo_scene_get_state(self.h, P(x, f32p), P(v, f32p), P(xp, f32p))
        return x.reshape(-1, 3), v.reshape(-1, 3), xp.reshape(-1, 3)

    def set_state(self, x=None, v=None, xprev=None):
        a = [None if t is None else np.ascontiguousarray(t, np.float32).reshape(-1) for t in (x, v, xprev)]
        lib().yo_scene_set_state(self.h, P(a[0], f32p), P(a[1], f32p), P(a[2], f32p))

    def set_collision(self, margin: float, radius: float, narrow_thickness: float, self_collision: int):
        """Collision arguments of later substeps (the library's scene_set_collision)."""
        lib().yo_scene_set_collision(self.h, margin, radius, narrow_thickness, int(self_collision))

    def record_times(self, on: bool = True):
        """Record each later substep's per-stage wall time (clears the record)."""
        lib().yo_scene_record_times(self.h, int(on))

    def step_times(self) -> np.ndarray:
        """(substeps, 7) stage seconds (box, morton, sort, hierarchy, refit,
        query incl. narrow and contacts, integrate) of the recorded substeps."""
        L = lib()
        n = L.yo_scene_step_times(self.h, None, 0)
        out = np.zeros(7 * n, np.float64)
        L.yo_scene_step_times(self.h, P(out, f64p), n)
        return out.reshape(-1, 7)

    def enable_ties(self):
        """Classify every later candidate's narrow decision in FP64 too."""
        lib().yo_scene_enable_ties(self.h)

    def ties(self) -> dict:
        """Threshold ties since enable_ties: candidates classified, FP32/FP64
        decision differences, FP64 distances within 4 FP32 ulps of the band."""
        out = np.zeros(3, np.uint64)
        lib().yo_scene_ties(self.h, P(out, u64p))
        return dict(classified=int(out[0]), decision=int(out[1]), band=int(out[2]))

    def twin_state(self):
        x = np.zeros(3 * self.V, np.float64)
        v = np.zeros(3 * self.V, np.float64)
        lib().yo_scene_get_twin(self.h, P(x, f64p), P(v, f64p))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def contacts(self):
        L = lib()
        n = L.yo_scene_num_contacts(self.h)
        off = np.zeros(self.V + 1, np.uint32)
        rec = np.zeros(n, abi.NARROW)
        L.yo_scene_get_contacts(self.h, P(off, u32p), rec.ctypes.data)
        return off, rec

    def pairs(self):
        """Candidate set at the current state as an (n, 4) array (qobj, point, tobj, tri)."""
        L = lib()
        n = L.yo_scene_pairs(self.h, None, 0)
        out = np.zeros(4 * n, np.uint32)
        L.yo_scene_pairs(self.h, P(out, u32p), n)
        return out.reshape(-1, 4)

    def tree(self, mesh: int):
        T = self.scene.meshes[mesh].facets.shape[0]
        nodes = np.zeros(2 * T - 1, abi.BVH_NODE)
        parent = np.zeros(2 * T - 1, np.int32)
        keys = np.zeros(T, np.uint64)
        idx = np.zeros(T, np.uint32)
        lib().yo_scene_get_tree(self.h, mesh, nodes.ctypes.data, P(parent, C.POINTER(C.c_int32)),
                                P(keys, u64p), P(idx, u32p))
        return nodes, parent, keys, idx

    def anomaly(self) -> int:
        return lib().yo_scene_anomaly(self.h)


# ------------------------------------------------------------------ adjoint
class AdjointRollout:
    """CPU rollout of the scene's first cloth (mesh 0) against its static meshes,
    with the hand-derived adjoint (checkpoints every K steps, recompute per
    segment).  precision 32 follows the device's operation order; precision 64
    is the finite-difference-checked restatement."""

    def __init__(self, scene, precision: int = 64):
        from ysim_amd.scenes import BEHAVIOR_FASTGRID_CLOTH
        assert scene.meshes[0].behavior == BEHAVIOR_FASTGRID_CLOTH
        self.precision = int(precision)
        self.scene = scene
        cl = scene.meshes[0]
        pk = scene.packed()
        V = cl.x.shape[0]
        self.V = V
        self.N = int(cl.cloth["particle_num_1d"])
        self.S = lib().yo_adj_num_springs(self.N)
        self.descs = abi.mesh_descs(scene)
        self.desc = self.descs[0]
        self.params = abi.sim_params(scene.params)
        tri, obj, fid = [], [], []
        for mi, m in enumerate(scene.meshes[1:], start=1):
            if m.behavior == BEHAVIOR_FASTGRID_CLOTH or m.group != cl.group:
                continue
            tri.append(np.asarray(m.x, np.float32)[np.asarray(m.facets, np.int64)].reshape(-1, 9))
            obj.append(np.full(m.facets.shape[0], mi, np.uint32))
            fid.append(np.arange(m.facets.shape[0], dtype=np.uint32))
        obs = np.ascontiguousarray(np.concatenate(tri) if tri else np.zeros((0, 9), np.float32))
        oo = np.ascontiguousarray(np.concatenate(obj) if obj else np.zeros(0, np.uint32))
        of = np.ascontiguousarray(np.concatenate(fid) if fid else np.zeros(0, np.uint32))
        self._keep = [np.ascontiguousarray(pk[k][: (3 * V if k in ("x", "v", "ext") else V)])
                      for k in ("x", "v", "mass", "fixed", "ext")] + [obs, oo, of]
        x0, v0, mass, fixed, ext = self._keep[:5]
        self.h = lib().yo_adj_create(self.precision, C.byref(self.desc), C.byref(self.params), P(x0, f32p),
                                     P(v0, f32p), P(mass, f32p), P(fixed, f32p), P(ext, f32p), oo.size,
                                     P(obs, f32p), P(oo, u32p), P(of, u32p))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().yo_adj_destroy(self.precision, self.h)
                self.h = None
        except Exception:
            pass

    def forward(self, theta, kd: float, steps: int, K: int = 10, mode: str = "detect"):
        """mode "detect", "record" (detect and keep every step's contacts and
        gates) or "frozen" (replay the recorded contacts and gates)."""
        th = np.ascontiguousarray(theta, np.float64)
        assert th.size == self.S
        x = np.zeros(3 * self.V, np.float64)
        v = np.zeros(3 * self.V, np.float64)
        m = {"detect": 0, "record": 1, "frozen": 2}[mode]
        lib().yo_adj_forward(self.precision, self.h, P(th, f64p), float(kd), steps, K, m, P(x, f64p), P(v, f64p))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def adopt_record(self, other: "AdjointRollout"):
        """Take `other`'s recorded contacts and gates and switch to frozen mode
        (forward and backward then follow `other`'s piece of the map)."""
        lib().yo_adj_copy_record(other.precision, other.h, self.precision, self.h)

    def set_initial(self, x0=None, v0=None):
        """Initial state of later forwards (xPrev_0 = x_0); None keeps a part."""
        a = [None if t is None else np.ascontiguousarray(t, np.float64).reshape(-1) for t in (x0, v0)]
        lib().yo_adj_set_initial(self.precision, self.h, P(a[0], f64p), P(a[1], f64p))

    def plant(self, bits: int):
        """Plant a known mistake in the backward pass (tests of the checker)."""
        lib().yo_adj_plant(self.precision, self.h, int(bits))

    def contact_hist(self):
        out = np.zeros(self.V, np.uint32)
        lib().yo_adj_contact_hist(self.precision, self.h, P(out, u32p))
        return out

    def backward(self, xbar, vbar=None):
        """Adjoint of the last forward.  After a "frozen" forward (or adopt_record)
        the recorded contacts and gates are used."""
        xb = np.ascontiguousarray(xbar, np.float64).reshape(-1)
        vb = None if vbar is None else np.ascontiguousarray(vbar, np.float64).reshape(-1)
        g = np.zeros(self.S, np.float64)
        gkd = np.zeros(1, np.float64)
        gx = np.zeros(3 * self.V, np.float64)
        gv = np.zeros(3 * self.V, np.float64)
        lib().yo_adj_backward(self.precision, self.h, P(xb, f64p), P(vb, f64p), P(g, f64p), P(gkd, f64p),
                              P(gx, f64p), P(gv, f64p))
        return g, float(gkd[0]), gx.reshape(-1, 3), gv.reshape(-1, 3)

    def contacts(self):
        c = np.zeros(1, np.uint64)
        s = np.zeros(1, np.uint64)
        lib().yo_adj_contacts(self.precision, self.h, P(c, u64p), P(s, u64p))
        return int(c[0]), int(s[0])
