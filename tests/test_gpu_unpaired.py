"""Kernel variants the C5 workload never takes.

* A one-level box set with odd origins and edges: the cell kernels load one
  z-point per lane (k_cell_gram<*, false>, k_cell_residual<false>,
  k_cell_density<false>) instead of z-pairs; S / H, the residual and the
  density against the oracle.
* Boxes of 40^3 points: the Anderson update takes k_box_and_update_sq (the
  normalisation in a separate pass) instead of the in-register one; the
  stateful steps against the oracle's BSolver.
"""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import mgx, synth
from tests.gpu_util import relmax, say, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def _odd_cfg():
    # levels = 1: origins on any point; odd edges; gdim a multiple of the 8^3 cells
    return synth.box_config(nbox=30, gdim=48, h=0.3, edges=(9, 11, 13, 15), nwells=4, seed=21, levels=1)


@pytest.mark.parametrize("prec", [mgx.PREC_F64, mgx.PREC_SPLIT])
def test_unpaired_gram_residual_density(ctx, prec):
    cfg = _odd_cfg()
    assert np.any(cfg.origin % 2 == 1) and np.all(cfg.edge % 2 == 1)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 1)
    try:
        hp = O.b_hloc(ob, cfg.vtot, cfg.psi)
        n = ob.nbox
        tp, th = to_dev(cfg.psi), to_dev(hp)
        tS, tH = to_dev(np.zeros((n, n))), to_dev(np.zeros((n, n)))
        ctx.call("mgx_box_gram_h", bs.h, prec, mgx.ptr(tp), mgx.ptr(th), mgx.ptr(tS), mgx.ptr(tH))
        S, H = O.b_gram_h(ob, cfg.psi, hp)
        es, eh = relmax(to_np(tS), S), relmax(to_np(tH), H)
        say("unpaired gram/H", prec, es, eh)
        tol = 1e-13 if prec == mgx.PREC_F64 else 1e-7
        assert es <= tol and eh <= tol
        th_m = np.linalg.solve(S, H)
        tt = to_dev(np.asfortranarray(th_m).T.copy())
        tr = to_dev(np.zeros_like(cfg.psi))
        ctx.call("mgx_box_residual", bs.h, mgx.ptr(tp), mgx.ptr(th), mgx.ptr(tt), mgx.ptr(tr))
        er = relmax(to_np(tr), O.b_residual(ob, cfg.psi, hp, th_m))
        say("unpaired residual", er)
        assert er <= 1e-12
        dm = 2.0 * np.linalg.inv(S)
        rho = mgx.DeviceArray(ctx, shape=tuple(cfg.gdim), dtype=np.float64)
        q = bs.density(tp, to_dev(np.asfortranarray(dm).ravel(order="F")), rho)
        ref = O.b_density(ob, cfg.psi, dm)
        ed = relmax(to_np(rho), ref)
        say("unpaired density", ed, q)
        assert ed <= 1e-13
        assert q == pytest.approx(2.0 * n, rel=1e-11)
    finally:
        bs.close()


def test_anderson_40_boxes_match_oracle(ctx):
    cfg = synth.box_config(nbox=6, gdim=88, h=0.3, edges=(40,), nwells=3, seed=12, levels=3)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64, anderson_m=1)
    ref = O.BSolver(ob, cfg.psi, cfg.vtot, 2, O.F64, 1.5, anderson_m=1)
    try:
        engaged = 0
        for k in range(14):
            st = s.step()
            e, r, _ = ref.step()
            engaged = max(engaged, ref.mm)
            say("40^3 anderson step", k, st.band_energy, e, st.residual, r, ref.accelerate)
            assert st.band_energy == pytest.approx(e, rel=1e-11)
            assert st.residual == pytest.approx(r, rel=1e-8)
        assert engaged == 1
        gp = s.psi()
        assert np.max(np.abs(gp - ref.psi)) <= 1e-10 * np.max(np.abs(ref.psi))
    finally:
        s.close()
        bs.close()
