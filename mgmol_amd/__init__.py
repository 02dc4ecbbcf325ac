"""mgmol_amd -- MI355X-native engine for MGmol's orbital-iteration hot path.

The compute path is ``libmgx.so`` (HIP kernels for gfx950 behind the C ABI
declared in ``include/mgx.h``); ``mgmol_amd.mgx`` binds it with ctypes and
raises if the library is missing -- there is no CPU fallback.
"""
import os

ROOT = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ROOT, "lib", "libmgx.so")
