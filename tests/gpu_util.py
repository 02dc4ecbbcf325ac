"""Helpers shared by the GPU tests (not collected: no test_ prefix).

Device buffers come from the engine's own C ABI (mgx_malloc / mgx_memcpy), so
the test process loads one HIP runtime only.
"""
import sys

import numpy as np

_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        from mgmol_amd import mgx
        _CTX = mgx.Ctx(0)
    return _CTX


def to_dev(a):
    from mgmol_amd import mgx
    return mgx.DeviceArray(ctx(), np.ascontiguousarray(a))


def to_np(t):
    return t.get()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / s)
