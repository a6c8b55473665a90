"""ctypes / numpy mirrors of the POD structs in include/ysim_hip.h."""
from __future__ import annotations

import ctypes as C

import numpy as np


class Aabb4(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("i0", C.c_int32), ("max", C.c_float * 3), ("i1", C.c_int32)]


class FastGridParams(C.Structure):
    _fields_ = [("particle_num_1d", C.c_uint32),
                ("stretch_rest_x", C.c_float), ("stretch_rest_y", C.c_float),
                ("shear_rest_a", C.c_float), ("shear_rest_b", C.c_float),
                ("bend_rest_x", C.c_float), ("bend_rest_y", C.c_float),
                ("kstretch", C.c_float), ("kshear", C.c_float), ("kbend", C.c_float),
                ("thickness", C.c_float)]


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_facets", C.c_uint32),
                ("behavior", C.c_uint32), ("collidable", C.c_uint32), ("group", C.c_uint32),
                ("shape", C.c_uint32), ("cloth", FastGridParams)]


class SimParams(C.Structure):
    _fields_ = [("subh", C.c_float), ("kair", C.c_float), ("kd", C.c_float),
                ("margin", C.c_float), ("radius", C.c_float), ("narrow_thickness", C.c_float),
                ("morton_bits", C.c_uint32), ("rebuild_every", C.c_uint32),
                ("self_collision", C.c_uint32), ("swept", C.c_uint32),
                ("broad_mode", C.c_uint32), ("hash_slots_log2", C.c_uint32),
                ("hash_cell", C.c_float), ("contact_capacity", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("sort_digit_bits", C.c_uint32), ("sort_items", C.c_uint32), ("query_block", C.c_uint32),
                ("cloth_block", C.c_uint32), ("ms_per_substep", C.c_float),
                ("ms_per_substep_default", C.c_float), ("refit_block", C.c_uint32),
                ("sort_hybrid", C.c_uint32), ("sort_incremental", C.c_uint32), ("query_waves", C.c_uint32),
                ("query_tile", C.c_uint32), ("build_mode", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


assert C.sizeof(Aabb4) == 32
assert C.sizeof(FastGridParams) == 44
assert C.sizeof(MeshDesc) == 24 + 44

BVH_NODE = np.dtype([("min", np.float32, 3), ("child_a", np.int32),
                     ("max", np.float32, 3), ("child_b", np.int32)])
BROAD = np.dtype([("point", np.uint32), ("triangle", np.uint32), ("query_obj", np.uint32),
                  ("target_obj", np.uint32), ("query_behavior", np.uint32),
                  ("target_behavior", np.uint32), ("query_shape", np.uint32),
                  ("target_shape", np.uint32)])
NARROW = np.dtype([("point", np.uint32), ("triangle", np.uint32), ("query_obj", np.uint32),
                   ("target_obj", np.uint32), ("normal_distance", np.float32, 4),
                   ("query_behavior", np.uint32), ("target_behavior", np.uint32),
                   ("query_shape", np.uint32), ("target_shape", np.uint32)])
assert BVH_NODE.itemsize == 32 and BROAD.itemsize == 32 and NARROW.itemsize == 48
# ysim_analytic_shape (AnalyticShape, include/sim/collision_pod.hpp:299-310)
ANALYTIC = np.dtype([("center_radius", np.float32, 4), ("half_ext_height", np.float32, 4),
                     ("rot_quat", np.float32, 4), ("prev_center_pad", np.float32, 4),
                     ("prev_rot_quat", np.float32, 4), ("kind", np.uint32), ("obj_index", np.uint32),
                     ("behavior_type", np.uint32), ("flags", np.uint32)])
assert ANALYTIC.itemsize == 96
COLLIDER_MESH, COLLIDER_SPHERE, COLLIDER_BOX, COLLIDER_CYLINDER, COLLIDER_PLANE = range(5)

STAGES = ("box", "morton", "sort", "hierarchy", "refit", "query", "integrate")


def mesh_descs(scene) -> "C.Array":
    arr = (MeshDesc * len(scene.meshes))()
    for i, m in enumerate(scene.meshes):
        d = arr[i]
        d.num_vertices = m.x.shape[0]
        d.num_facets = m.facets.shape[0]
        d.behavior = m.behavior
        d.collidable = int(m.collidable)
        d.group = m.group
        d.shape = m.shape
        if m.cloth:
            for k, v in m.cloth.items():
                setattr(d.cloth, k, v)
    return arr


def sim_params(p: dict) -> SimParams:
    s = SimParams()
    for k, v in p.items():
        setattr(s, k, v)
    return s


def fastgrid_params(cloth: dict) -> FastGridParams:
    s = FastGridParams()
    for k, v in cloth.items():
        setattr(s, k, v)
    return s


def ptr(a: np.ndarray, ctype=C.c_void_p):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctype)
