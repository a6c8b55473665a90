"""ysim_amd — MI355X (gfx950) HIP backend for ysim's cloth collision + spring substep.

Package layout:
  include/ysim_hip.h      C ABI (the drop-in boundary)
  ysim_amd/csrc/          hand-written HIP kernels + the C ABI implementation
  ysim_amd/host/          C++ `HIP` backend tag over the C ABI
  ysim_amd/hip.py         Python host side (ctypes) of the same ABI
  ysim_amd/scenes.py      synthetic scene generators / packed layout
"""
__version__ = "0.1.0"
