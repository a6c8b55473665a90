"""The C-ABI library loads and exports every symbol include/ysim_hip.h declares.

CPU only: no compute call is made (there is no GPU here).
"""
import ctypes as C
import os
import re

from ysim_amd import _lib, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    L = _lib.load()
    syms = _lib.header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header():
    syms = set(_lib.header_symbols())
    assert set(_lib._SIG) <= syms
    assert not (syms - set(_lib._SIG)), sorted(syms - set(_lib._SIG))


def test_struct_layouts():
    assert C.sizeof(abi.Aabb4) == 32
    assert C.sizeof(abi.FastGridParams) == 44
    assert abi.BVH_NODE.itemsize == 32 and abi.BROAD.itemsize == 32 and abi.NARROW.itemsize == 48
    hdr = open(os.path.join(ROOT, "include", "ysim_hip.h")).read()
    fields = re.search(r"typedef struct ysim_sim_params \{(.*?)\} ysim_sim_params;", hdr, re.S).group(1)
    names = re.findall(r"\b(?:float|uint32_t)\s+(\w+);", fields)
    assert names == [f for f, _ in abi.SimParams._fields_]


def test_version_without_gpu():
    assert _lib.load().ysim_hip_version() >= 10000


def test_library_has_no_oracle_dependency():
    """The product library never links the checker."""
    import subprocess
    out = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    src = os.path.join(ROOT, "ysim_amd")
    for dp, _, files in os.walk(src):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                txt = open(os.path.join(dp, f)).read()
                for pat in ("from oracle", "import oracle", "liboracle", "yo_"):
                    assert pat not in txt, (f, pat)
