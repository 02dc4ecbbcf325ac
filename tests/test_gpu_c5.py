"""The full 2048-orbital workload (2048 boxes of 32^3 on a 256^3 grid, h = 0.2
bohr, 64 wells, centres at least 2 bohr apart): steps of every path, and the
Newton-Schulz refresh of S^-1 on its pair-sparse S at n = 2048.

* plain, mixed (FP32 V-cycle, split S/H) and the FP16/BF16 extension from the
  same initial orbitals: the first step's band energy and residual agree to
  the split products' accuracy; later steps (whose states differ by the
  V-cycle's rounding) to 1e-7 in energy for the FP32 V-cycle and to 1e-5 for
  the FP16-storage one, whose correction differs from the FP32 V-cycle's by
  up to 1.4e-3 of each box's largest value (measured 2.6e-6 of the energy
  after the second step); all paths lower the residual.
* After the refresh (FP64 products, and the BF16 correction), max|I - X S|
  against the Cholesky inverse of the same S (mgx_chol_inv): FP64 refresh at
  1e-15 within 1e-12 (the dense products' rounding over 2048 terms), the split
  tolerance 1e-10 within 1e-9, the BF16 correction within its tolerance 1e-10
  (the solver measures the final X's error and continues until it is below).
"""
import ctypes as C

import numpy as np
import pytest

from mgmol_amd import mgx, synth
from tests.gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


@pytest.fixture(scope="module")
def c5(ctx):
    kw = dict(synth.CONFIGS["c5"])
    kw.pop("kind")
    cfg = synth.box_config(seed=1, **kw)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, cfg.levels)
    yield cfg, bs
    bs.close()


PATHS = {
    "plain": dict(precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64, inverse_mode=mgx.INV_NS_F64),
    "mixed": dict(precond_dtype=mgx.F32, gram_prec=mgx.PREC_SPLIT, inverse_mode=mgx.INV_NS_F64),
    "ext16": dict(precond_dtype=mgx.F16, gram_prec=mgx.PREC_SPLIT, inverse_mode=mgx.INV_NS_BF16),
}


def _steps(ctx, cfg, bs, path, nsteps):
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=cfg.levels - 1, anderson_m=1,
                   precond_start=mgx.START_ZERO, **PATHS[path])
    sts = [s.step() for _ in range(nsteps)]
    return s, [(st.band_energy, st.residual) for st in sts]


def test_c5_paths_agree(ctx, c5):
    cfg, bs = c5
    assert cfg.nbox == 2048 and cfg.psi.size == 2048 * 32 ** 3
    out = {}
    for path in PATHS:
        s, out[path] = _steps(ctx, cfg, bs, path, 6)
        s.close()
        say(path, out[path])
    e0, r0 = out["plain"][0]
    for path in ("mixed", "ext16"):
        e, r = out[path][0]
        assert e == pytest.approx(e0, rel=1e-9)
        assert r == pytest.approx(r0, rel=1e-6)
        tol = 1e-7 if path == "mixed" else 1e-5
        for k in range(1, 6):
            assert out[path][k][0] == pytest.approx(out["plain"][k][0], rel=tol)
        assert out[path][-1][1] < 0.5 * r
    assert out["plain"][-1][1] < 0.5 * r0


@pytest.mark.parametrize("mode,tol,bound", [("f64", 0.0, 1e-12), ("f64", 1e-10, 1e-9), ("bf16", 1e-10, 1e-10)])
def test_c5_newton_schulz_against_cholesky(ctx, c5, mode, tol, bound):
    cfg, bs = c5
    kw = dict(PATHS["plain" if tol == 0.0 else "mixed"])
    kw["inverse_mode"] = mgx.INV_NS_F64 if mode == "f64" else mgx.INV_NS_BF16
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=cfg.levels - 1, anderson_m=1, ns_tol=tol, **kw)
    try:
        for _ in range(4):  # Cholesky on the first step, the refresh on the others
            st = s.step()
        assert st.chol_calls == 1 and st.ns_steps >= 3
        S, X = s.matrices()
        n = S.shape[0]
        # Cholesky inverse of the same S on the device
        dS = mgx.DeviceArray(ctx, np.asfortranarray(S).T.copy())
        dL = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
        dX = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
        ctx.call("mgx_chol_inv", n, mgx.ptr(dS), mgx.ptr(dL), mgx.ptr(dX))
        Xc = dX.get().T
        eye = np.eye(n)
        err_ns = float(np.max(np.abs(eye - X @ S)))
        err_ch = float(np.max(np.abs(eye - Xc @ S)))
        diff = float(np.max(np.abs(X - Xc)) / np.max(np.abs(Xc)))
        say("n=%d %s tol %g: max|I - X S| refresh %.3e, Cholesky %.3e, max|X - Xc|/max|Xc| %.3e, %d NS steps"
            % (n, mode, tol, err_ns, err_ch, diff, st.ns_steps))
        assert n == 2048
        assert err_ns <= bound
        assert err_ch <= 1e-12
        assert diff <= 10 * bound
    finally:
        s.close()
