"""ctypes binding of libysim_hip.so (the C ABI in include/ysim_hip.h).

Fails loudly: there is no CPU fallback anywhere in the package.  If the
library is missing, importing the binding raises with the build command.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libysim_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ysim_hip.h")

vp = C.c_void_p
u32 = C.c_uint32
u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u64p = C.POINTER(C.c_uint64)


class YsimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ysim_hip error {code}: {msg}")
        self.code = code


ERR_CAPACITY = 3

_SIG = {
    "ysim_hip_last_error": (C.c_char_p, []),
    "ysim_hip_version": (C.c_int, []),
    "ysim_hip_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ysim_hip_ctx_create": (C.c_int, [C.c_int, u32, C.POINTER(vp)]),
    "ysim_hip_ctx_destroy": (C.c_int, [vp]),
    "ysim_hip_sync": (C.c_int, [vp]),
    "ysim_hip_alloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
    "ysim_hip_free": (C.c_int, [vp, vp]),
    "ysim_hip_upload": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "ysim_hip_download": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "ysim_hip_memset": (C.c_int, [vp, vp, C.c_int, C.c_size_t]),
    "ysim_hip_scene_box": (C.c_int, [vp, vp, u32, u32, vp]),
    "ysim_hip_morton": (C.c_int, [vp, vp, vp, u32, vp, u32, vp, vp]),
    "ysim_hip_sort_pairs": (C.c_int, [vp, vp, vp, u32, u32]),
    "ysim_hip_lbvh_hierarchy": (C.c_int, [vp, vp, vp, u32, u32, vp, vp]),
    "ysim_hip_lbvh_agglomerative": (C.c_int, [vp, vp, vp, u32, u32, vp, vp, vp, vp]),
    "ysim_hip_lbvh_refit": (C.c_int, [vp, vp, vp, C.c_float, u32, vp, u32, vp, vp]),
    "ysim_hip_broad_query": (C.c_int, [vp, vp, u32, u32, u32, u32, vp, u32, vp, u32, u32, u32,
                                       C.c_float, vp, u32, vp, u32p]),
    "ysim_hip_broad_query_segmented": (C.c_int, [vp, vp, u32, u32, u32, u32, vp, u32, vp, u32, u32, u32,
                                       C.c_float, vp, u32, vp, u32p]),
    "ysim_hip_narrow": (C.c_int, [vp, vp, u32, vp, vp, vp, vp, vp, vp, vp, C.c_float, C.c_float,
                                  vp, u32, vp, u32p]),
    "ysim_hip_contact_csr": (C.c_int, [vp, vp, u32, vp, u32, vp, vp]),
    "ysim_hip_cloth_step": (C.c_int, [vp, vp, u32, C.c_float, C.c_float, C.c_float, vp, vp, vp, vp,
                                      vp, vp, vp, vp, vp, vp, vp]),
    "ysim_hip_scene_create": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]),
    "ysim_hip_scene_destroy": (C.c_int, [vp]),
    "ysim_hip_scene_step": (C.c_int, [vp, u32, u32]),
    "ysim_hip_scene_refit": (C.c_int, [vp]),
    "ysim_hip_scene_query": (C.c_int, [vp]),
    "ysim_hip_scene_integrate": (C.c_int, [vp]),
    "ysim_hip_scene_get_state": (C.c_int, [vp, vp, vp, vp]),
    "ysim_hip_scene_set_state": (C.c_int, [vp, vp, vp, vp]),
    "ysim_hip_scene_set_analytic": (C.c_int, [vp, vp, C.c_uint32]),
    "ysim_hip_scene_deferred_builds": (C.c_int, [vp, vp]),
    "ysim_hip_scene_get_contacts": (C.c_int, [vp, vp, vp, u32, u32p]),
    "ysim_hip_scene_get_pairs": (C.c_int, [vp, vp, u32, u32p]),
    "ysim_hip_scene_set_audit": (C.c_int, [vp, u32]),
    "ysim_hip_scene_set_collision": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, u32]),
    "ysim_hip_scene_get_audit": (C.c_int, [vp, vp, u32, u32p, u32p]),
    "ysim_hip_scene_get_cut": (C.c_int, [vp, u32, vp, u32, u32p]),
    "ysim_hip_scene_get_tree": (C.c_int, [vp, u32, vp, vp, vp, vp]),
    "ysim_hip_scene_stage_timings": (C.c_int, [vp, u32, f32p]),
    "ysim_hip_scene_flags": (C.c_int, [vp, u32p]),
    "ysim_hip_scene_counters": (C.c_int, [vp, u64p, u32]),
    "ysim_hip_scene_snapshot": (C.c_int, [vp]),
    "ysim_hip_scene_restore": (C.c_int, [vp]),
    "ysim_hip_copy_bandwidth": (C.c_int, [vp, C.c_size_t, u32, f32p]),
    "ysim_hip_sort_pairs_ex": (C.c_int, [vp, vp, vp, u32, u32, u32, u32, u32]),
    "ysim_hip_sort_pairs_noalloc": (C.c_int, [vp, vp, vp, u32, u32]),
    "ysim_hip_scene_get_tuning": (C.c_int, [vp, vp]),
    "ysim_hip_scene_set_tuning": (C.c_int, [vp, vp]),
    "ysim_hip_scene_autotune": (C.c_int, [vp, u32, vp]),
    "ysim_hip_adj_num_springs": (C.c_int, [vp, u32p]),
    "ysim_hip_adj_forward": (C.c_int, [vp, vp, C.c_float, u32, u32]),
    "ysim_hip_adj_backward": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
}

_lib = None


def header_symbols():
    """Every ysim_hip_* function declared in include/ysim_hip.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ysim_hip_\w+)\s*\(", txt)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "(or `make -C ysim_amd/csrc`)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIG.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return load().ysim_hip_last_error().decode(errors="replace")


def check(rc: int):
    if rc != 0:
        raise YsimError(rc, last_error())
    return rc


def call(name: str, *args):
    return check(getattr(load(), name)(*args))
