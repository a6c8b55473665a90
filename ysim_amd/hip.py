"""Host-side HIP backend: device context, device arrays and the resident scene.

These classes are the Python face of the `HIP` backend tag; each method names
the reference operation it stands in for.  All compute runs in libysim_hip.so.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional

import numpy as np

from . import abi
from ._lib import ERR_CAPACITY, YsimError, call, load

vp = C.c_void_p


class HipContext:
    """One device + one HIP stream (MetalGlobalContext, core_types.hpp:83-187)."""

    def __init__(self, device: int = 0):
        load()
        h = vp()
        call("ysim_hip_ctx_create", device, 0, C.byref(h))
        self.h = h
        # scenes and arrays use the context's stream: they are released first
        self._owned = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            for o in list(getattr(self, "_owned", ())):
                o.close()
            load().ysim_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        call("ysim_hip_sync", self.h)

    # ---------------------------------------------------------- device arrays
    def alloc(self, nbytes: int) -> "DeviceArray":
        return DeviceArray(self, nbytes)

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 20) -> dict:
        """Achievable HBM bandwidth of a plain device-to-device vector copy."""
        ms = C.c_float(0)
        call("ysim_hip_copy_bandwidth", self.h, int(nbytes), int(reps), C.byref(ms))
        return {"bytes": int(nbytes), "ms": ms.value, "GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9}

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.nbytes, dtype=a.dtype, shape=a.shape)
        if a.nbytes:
            call("ysim_hip_upload", self.h, d.ptr, a.ctypes.data, a.nbytes)
        return d

    def zeros(self, shape, dtype) -> "DeviceArray":
        a = np.zeros(shape, dtype)
        return self.to_device(a)


class DeviceArray:
    """A device allocation owned by Python (VectorBase<HIP, T> view + storage)."""

    def __init__(self, ctx: HipContext, nbytes: int, dtype=np.uint8, shape=None):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = vp()
        call("ysim_hip_alloc", ctx.h, max(self.nbytes, 1), C.byref(p))
        self.ptr = p
        if getattr(ctx, "_owned", None) is not None:
            ctx._owned.add(self)

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("ysim_hip_download", self.ctx.h, out.ctypes.data, self.ptr, self.nbytes)
        return out

    def set(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        call("ysim_hip_upload", self.ctx.h, self.ptr, a.ctypes.data, a.nbytes)

    def close(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            load().ysim_hip_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- stage ops
def scene_box(ctx: HipContext, x: np.ndarray, num_prims: int = 0) -> np.ndarray:
    """BVH::build scene box (bvh.hpp:1229-1231) -> (min[3], max[3])."""
    dx = ctx.to_device(np.asarray(x, np.float32).reshape(-1))
    box = ctx.zeros(8, np.float32)
    call("ysim_hip_scene_box", ctx.h, dx.ptr, dx.nbytes // 12, num_prims, box.ptr)
    b = box.get()
    return np.concatenate([b[0:3], b[4:7]])


def morton(ctx: HipContext, x, facets, box6, bits: int):
    """fillMortons_Tri (bvh.metal:42-63): keys (uint64 view) and identity indices."""
    dx = ctx.to_device(np.asarray(x, np.float32).reshape(-1))
    df = ctx.to_device(np.asarray(facets, np.uint32).reshape(-1))
    n = df.nbytes // 12
    b = np.zeros(8, np.float32)
    b[0:3] = box6[0:3]
    b[4:7] = box6[3:6]
    db = ctx.to_device(b)
    kd = np.uint64 if bits > 32 else np.uint32
    keys = ctx.zeros(n, kd)
    idx = ctx.zeros(n, np.uint32)
    call("ysim_hip_morton", ctx.h, dx.ptr, df.ptr, n, db.ptr, bits, keys.ptr, idx.ptr)
    return keys.get().astype(np.uint64), idx.get()


class RadixSorter:
    """RadixSorter<HIP, E>::sort (spatial_hash.hpp:48-70): stable LSD radix sort."""

    def __init__(self, ctx: HipContext):
        self.ctx = ctx

    def sort(self, keys: np.ndarray, vals: np.ndarray, key_bits: int, digit_bits: int = 8, items: int = 8,
             hybrid: bool = False):
        """Configuration (digit_bits, items, hybrid) changes speed only, never the result."""
        kd = np.uint64 if key_bits > 32 else np.uint32
        dk = self.ctx.to_device(np.asarray(keys, kd))
        dv = self.ctx.to_device(np.asarray(vals, np.uint32))
        if (digit_bits, items, hybrid) == (8, 8, False):
            call("ysim_hip_sort_pairs", self.ctx.h, dk.ptr, dv.ptr, len(keys), key_bits)
        else:
            call("ysim_hip_sort_pairs_ex", self.ctx.h, dk.ptr, dv.ptr, len(keys), key_bits, digit_bits, items,
                 int(hybrid))
        return dk.get().astype(np.uint64), dv.get()


def lbvh_hierarchy(ctx: HipContext, sorted_keys, sorted_idx, bits: int):
    """buildTree_Tri child links (bvh.metal:352-393) -> (nodes, parent)."""
    n = len(sorted_keys)
    kd = np.uint64 if bits > 32 else np.uint32
    dk = ctx.to_device(np.asarray(sorted_keys, kd))
    di = ctx.to_device(np.asarray(sorted_idx, np.uint32))
    tree = ctx.zeros(2 * n - 1, abi.BVH_NODE)
    parent = ctx.to_device(np.full(2 * n - 1, -1, np.int32))
    call("ysim_hip_lbvh_hierarchy", ctx.h, dk.ptr, di.ptr, n, bits, tree.ptr, parent.ptr)
    return tree.get(), parent.get()


def lbvh_agglomerative(ctx: HipContext, sorted_keys, sorted_idx, bits: int, x, facets):
    """agglomerativeBuild_Tri + agglomerativeSwapRoot (bvh.metal:884-1199) ->
    (nodes with boxes, parent)."""
    n = len(sorted_keys)
    kd = np.uint64 if bits > 32 else np.uint32
    dk = ctx.to_device(np.asarray(sorted_keys, kd))
    di = ctx.to_device(np.asarray(sorted_idx, np.uint32))
    dx = ctx.to_device(np.asarray(x, np.float32).reshape(-1))
    df = ctx.to_device(np.asarray(facets, np.uint32).reshape(-1))
    tree = ctx.zeros(2 * n - 1, abi.BVH_NODE)
    parent = ctx.to_device(np.full(2 * n - 1, -1, np.int32))
    call("ysim_hip_lbvh_agglomerative", ctx.h, dk.ptr, di.ptr, n, bits, dx.ptr, df.ptr, tree.ptr, parent.ptr)
    return tree.get(), parent.get()


def lbvh_refit(ctx: HipContext, tree, parent, x, v, dt, swept, facets):
    """buildLeaf/buildSweptLeaf + bottomUpBoxes (bvh.metal:260-586) -> nodes."""
    dt_ = ctx.to_device(np.ascontiguousarray(tree))
    dp = ctx.to_device(np.asarray(parent, np.int32))
    dx = ctx.to_device(np.asarray(x, np.float32).reshape(-1))
    dv = ctx.to_device(np.asarray(v, np.float32).reshape(-1))
    df = ctx.to_device(np.asarray(facets, np.uint32).reshape(-1))
    call("ysim_hip_lbvh_refit", ctx.h, dx.ptr, dv.ptr, float(dt), int(swept), df.ptr, df.nbytes // 12,
         dt_.ptr, dp.ptr)
    return dt_.get()


def broad_query(ctx: HipContext, xq, qobj, tree, tfacets, tobj, margin, capacity=None, segmented=False):
    """queryPoints (bvh.metal:1393-1412) for one (query mesh, target tree);
    segmented: queryPointsSegmented's per-workgroup emission (bvh.metal:1732-1827)."""
    fn = load().ysim_hip_broad_query_segmented if segmented else load().ysim_hip_broad_query
    dx = ctx.to_device(np.asarray(xq, np.float32).reshape(-1))
    dtree = ctx.to_device(np.ascontiguousarray(tree))
    df = ctx.to_device(np.asarray(tfacets, np.uint32).reshape(-1))
    nT = df.nbytes // 12
    cap = capacity if capacity is not None else 16 * (dx.nbytes // 12) + 1024
    for _ in range(2):
        out = ctx.zeros(max(cap, 1), abi.BROAD)
        cnt = ctx.zeros(1, np.uint32)
        req = C.c_uint32(0)
        rc = fn(ctx.h, dx.ptr, dx.nbytes // 12, qobj, 1, 0, dtree.ptr, nT, df.ptr, tobj,
                0, 0, float(margin), out.ptr, cap, cnt.ptr, C.byref(req))
        if rc == ERR_CAPACITY:
            cap = req.value
            continue
        if rc:
            raise YsimError(rc, load().ysim_hip_last_error().decode())
        return out.get()[: req.value]
    raise RuntimeError("broad_query: capacity retry failed")


# ---------------------------------------------------------------- the scene
class HipScene:
    """A packed scene resident in HBM plus its substep (Simulator::update)."""

    def __init__(self, ctx: HipContext, scene, kspring: Optional[np.ndarray] = None):
        self.ctx = ctx
        self.scene = scene
        pk = scene.packed()
        self.descs = abi.mesh_descs(scene)
        self.params = abi.sim_params(scene.params)
        h = vp()
        call("ysim_hip_scene_create", ctx.h, len(scene.meshes), C.cast(self.descs, vp), pk["x"].ctypes.data,
             pk["v"].ctypes.data, pk["mass"].ctypes.data, pk["ext"].ctypes.data, pk["fixed"].ctypes.data,
             pk["facets"].ctypes.data, C.byref(self.params), C.byref(h))
        self.h = h
        if getattr(ctx, "_owned", None) is not None:
            ctx._owned.add(self)
        self.V = scene.num_vertices
        if kspring is not None:
            raise NotImplementedError("per-spring stiffness is served by the adjoint API")
        if scene.analytic is not None and len(scene.analytic):
            self.set_analytic(scene.analytic)

    def deferred_builds(self) -> int:
        """Graph replays that kept a previous topology (failed order check)."""
        n = C.c_uint64(0)
        call("ysim_hip_scene_deferred_builds", self.h, C.byref(n))
        return n.value

    def set_analytic(self, shapes):
        """Install analytic colliders (a numpy array of dtype abi.ANALYTIC)."""
        a = np.ascontiguousarray(shapes, abi.ANALYTIC).reshape(-1)
        call("ysim_hip_scene_set_analytic", self.h, a.ctypes.data if a.size else None, a.size)

    def close(self):
        if getattr(self, "h", None):
            load().ysim_hip_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1, graph: bool = False):
        call("ysim_hip_scene_step", self.h, n, int(graph))

    def sync(self):
        self.ctx.sync()

    def state(self):
        x = np.empty((self.V, 3), np.float32)
        v = np.empty((self.V, 3), np.float32)
        xp = np.empty((self.V, 3), np.float32)
        call("ysim_hip_scene_get_state", self.h, x.ctypes.data, v.ctypes.data, xp.ctypes.data)
        return x, v, xp

    def set_state(self, x=None, v=None, xprev=None):
        a = [None if t is None else np.ascontiguousarray(t, np.float32).reshape(-1) for t in (x, v, xprev)]
        call("ysim_hip_scene_set_state", self.h, *(None if t is None else t.ctypes.data for t in a))

    def contacts(self):
        off = np.zeros(self.V + 1, np.uint32)
        n = C.c_uint32(0)
        call("ysim_hip_scene_get_contacts", self.h, off.ctypes.data, None, 0, C.byref(n))
        rec = np.zeros(n.value, abi.NARROW)
        call("ysim_hip_scene_get_contacts", self.h, off.ctypes.data, rec.ctypes.data, n.value, C.byref(n))
        return off, rec

    def pairs(self):
        """Candidate set at the current state as (n, 4): (qobj, point, tobj, tri)."""
        n = C.c_uint32(0)
        call("ysim_hip_scene_get_pairs", self.h, None, 0, C.byref(n))
        rec = np.zeros(n.value, abi.BROAD)
        call("ysim_hip_scene_get_pairs", self.h, rec.ctypes.data, n.value, C.byref(n))
        return np.stack([rec["query_obj"], rec["point"], rec["target_obj"], rec["triangle"]], 1)

    def set_collision(self, margin: float, radius: float, narrow_thickness: float, self_collision: int):
        """detectCollisions / narrowAndSortByVertices arguments of later substeps."""
        call("ysim_hip_scene_set_collision", self.h, float(margin), float(radius), float(narrow_thickness),
             int(self_collision))

    def set_audit(self, capacity: int):
        """Record every candidate pair of each later substep's query (the
        kernel the tuning selects, eager or graph); 0 turns it off."""
        call("ysim_hip_scene_set_audit", self.h, int(capacity))

    def audit(self):
        """Candidate pairs the last audited query reported, as (n, 4) rows
        (qobj, point, tobj, tri) sorted, duplicates kept, and the number of
        tile walks that were redone per lane."""
        n = C.c_uint32(0)
        redo = C.c_uint32(0)
        call("ysim_hip_scene_get_audit", self.h, None, 0, C.byref(n), C.byref(redo))
        rec = np.zeros(n.value, abi.BROAD)
        call("ysim_hip_scene_get_audit", self.h, rec.ctypes.data, n.value, C.byref(n), C.byref(redo))
        rows = np.stack([rec["query_obj"], rec["point"], rec["target_obj"], rec["triangle"]], 1)
        return rows.astype(np.uint32).reshape(-1, 4), int(redo.value)

    def cut(self, mesh: int) -> np.ndarray:
        """The tree cut the last Karras build of `mesh` listed (empty: none)."""
        n = C.c_uint32(0)
        out = np.zeros(512, np.int32)
        call("ysim_hip_scene_get_cut", self.h, mesh, out.ctypes.data, out.size, C.byref(n))
        return out[: n.value].copy()

    def tree(self, mesh: int, keys: bool = False):
        T = self.scene.meshes[mesh].facets.shape[0]
        nodes = np.zeros(2 * T - 1, abi.BVH_NODE)
        parent = np.zeros(2 * T - 1, np.int32)
        k = np.zeros(T, np.uint64) if keys else None
        i = np.zeros(T, np.uint32) if keys else None
        call("ysim_hip_scene_get_tree", self.h, mesh, nodes.ctypes.data, parent.ctypes.data,
             None if k is None else k.ctypes.data, None if i is None else i.ctypes.data)
        return (nodes, parent, k, i) if keys else (nodes, parent)

    def stage_timings(self, n: int):
        out = (C.c_float * len(abi.STAGES))()
        call("ysim_hip_scene_stage_timings", self.h, n, out)
        return dict(zip(abi.STAGES, list(out)))

    def snapshot(self):
        call("ysim_hip_scene_snapshot", self.h)

    def restore(self):
        call("ysim_hip_scene_restore", self.h)

    def counters(self, reset: bool = False):
        out = (C.c_uint64 * 3)()
        call("ysim_hip_scene_counters", self.h, out, int(reset))
        return dict(pairs=int(out[0]), steps=int(out[1]), contacts=int(out[2]))

    def flags(self) -> int:
        f = C.c_uint32(0)
        call("ysim_hip_scene_flags", self.h, C.byref(f))
        return f.value

    # ------------------------------------------------------------ tuning
    def tuning(self) -> dict:
        t = abi.Tuning()
        call("ysim_hip_scene_get_tuning", self.h, C.byref(t))
        return t.as_dict()

    def set_tuning(self, **kw):
        """Launch parameters (sort_digit_bits, sort_items, query_block,
        cloth_block); unspecified ones keep their current value."""
        t = abi.Tuning()
        call("ysim_hip_scene_get_tuning", self.h, C.byref(t))
        for k, v in kw.items():
            if k not in dict(abi.Tuning._fields_):
                raise KeyError(k)
            setattr(t, k, v)
        call("ysim_hip_scene_set_tuning", self.h, C.byref(t))

    def autotune(self, trial_substeps: int = 10) -> dict:
        t = abi.Tuning()
        call("ysim_hip_scene_autotune", self.h, int(trial_substeps), C.byref(t))
        return t.as_dict()

    # ------------------------------------------------------------ differentiation
    # Adjoint of the substep for a one-cloth scene with static obstacles
    # (include/ysim_hip.h, "differentiation").  The *_ptr forms take device
    # pointers (e.g. torch tensors' data_ptr()); the array forms stage numpy
    # data through DeviceArrays.
    def num_springs(self) -> int:
        n = C.c_uint32(0)
        call("ysim_hip_adj_num_springs", self.h, C.byref(n))
        return n.value

    def cloth_vertices(self) -> int:
        return int(self.scene.meshes[0].x.shape[0])

    def adj_forward_ptr(self, d_log_k: int, kd: float, steps: int, checkpoint_every: int = 10):
        call("ysim_hip_adj_forward", self.h, vp(d_log_k), float(kd), int(steps), int(checkpoint_every))

    def adj_backward_ptr(self, d_x_bar: int, d_v_bar=None, d_g_log_k=None, d_g_kd=None, d_g_x0=None,
                         d_g_v0=None):
        call("ysim_hip_adj_backward", self.h, vp(d_x_bar), vp(d_v_bar), vp(d_g_log_k), vp(d_g_kd), vp(d_g_x0),
             vp(d_g_v0))

    def adj_forward(self, log_k: np.ndarray, kd: float, steps: int, checkpoint_every: int = 10):
        lk = np.ascontiguousarray(log_k, np.float32)
        if lk.size != self.num_springs():
            raise ValueError(f"log_k has {lk.size} entries, the cloth has {self.num_springs()} springs")
        self._adj_lk = self.ctx.to_device(lk)
        self.adj_forward_ptr(self._adj_lk.ptr.value, kd, steps, checkpoint_every)

    def adj_backward(self, x_bar: np.ndarray, v_bar: Optional[np.ndarray] = None):
        """-> (dL/dlog_k, dL/dkd, dL/dx_0, dL/dv_0) as numpy arrays."""
        Vc = self.cloth_vertices()
        xb = self.ctx.to_device(np.ascontiguousarray(x_bar, np.float32).reshape(-1))
        if xb.nbytes != 12 * Vc:
            raise ValueError("x_bar must hold 3 floats per cloth vertex")
        vb = None if v_bar is None else self.ctx.to_device(np.ascontiguousarray(v_bar, np.float32).reshape(-1))
        S = self.num_springs()
        g = self.ctx.zeros(S, np.float32)
        gkd = self.ctx.zeros(1, np.float32)
        gx = self.ctx.zeros(3 * Vc, np.float32)
        gv = self.ctx.zeros(3 * Vc, np.float32)
        self.adj_backward_ptr(xb.ptr.value, None if vb is None else vb.ptr.value, g.ptr.value, gkd.ptr.value,
                              gx.ptr.value, gv.ptr.value)
        return g.get(), float(gkd.get()[0]), gx.get().reshape(-1, 3), gv.get().reshape(-1, 3)
