"""cpolyg_amd: MI355X-native implicit DG hot path of cpolyg.

The compute path is ``libcpolyg_hip.so`` (hand-written HIP kernels for gfx950 behind
the C ABI in ``include/cpolyg.h``); :mod:`cpolyg_amd.hip` is its ctypes binding.
:mod:`cpolyg_amd.mesh` builds the polygonal meshes and :mod:`cpolyg_amd.setups`
the initial data.
"""
from . import spec  # noqa: F401

__version__ = "0.1.0"
