"""The Mehrstellen discretisation on per-orbital boxes against the oracle
(pinned on the full grid in test_oracle_mehr_box):

* box H psi = -Lap_M psi + B V psi (1e-13), the Galerkin S, H (not
  symmetric) and B-matrix with the operators' one-point reach (1e-12), the
  residual (B Psi) theta - H Psi (1e-12), on a box set with paired cells
  (even origins) and one with odd origins and edges;
* the solver with lap_type = Mehrstellen: plain steps match the oracle's
  Mehrstellen steps (energy 1e-11, residual 1e-9, orbitals 1e-10), and the
  stateful steps with Anderson extrapolation match the oracle's BSolver.
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


def _cfg(kind):
    if kind == "odd":
        return synth.box_config(nbox=14, gdim=40, h=0.3, edges=(7, 9, 11), nwells=3, seed=8, levels=1)
    return synth.box_config(nbox=14, gdim=48, h=0.3, edges=(8, 12, 16), nwells=4, seed=9, levels=3)


@pytest.mark.parametrize("kind", ["even", "odd"])
def test_box_mehr_entry_points(ctx, kind):
    cfg = _cfg(kind)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    lv = 1 if kind == "odd" else 3
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, lv)
    try:
        n = ob.nbox
        tp, tv = to_dev(cfg.psi), to_dev(cfg.vtot)
        th = to_dev(np.zeros_like(cfg.psi))
        ctx.call("mgx_box_hloc_mehr", bs.h, mgx.ptr(tp), mgx.ptr(tv), mgx.ptr(th))
        hp = O.b_hloc_mehr(ob, cfg.vtot, cfg.psi)
        eh = relmax(to_np(th), hp)
        tS, tH, tB = (to_dev(np.zeros((n, n))) for _ in range(3))
        ctx.call("mgx_box_gram_hb_mehr", bs.h, mgx.ptr(tp), mgx.ptr(th), mgx.ptr(tv), mgx.ptr(tS), mgx.ptr(tH),
                 mgx.ptr(tB))
        S, H, Bm = O.b_gram_hb_mehr(ob, cfg.psi, cfg.vtot)
        # column-major on the device: entry (i, j) at [j, i] of the host copy
        es, eH, eB = relmax(to_np(tS).T, S), relmax(to_np(tH).T, H), relmax(to_np(tB).T, Bm)
        rng = np.random.default_rng(2)
        theta = rng.standard_normal((n, n))
        tt = to_dev(np.asfortranarray(theta).T.copy())
        tr = to_dev(np.zeros_like(cfg.psi))
        ctx.call("mgx_box_residual_mehr", bs.h, mgx.ptr(tp), mgx.ptr(th), mgx.ptr(tt), mgx.ptr(tr))
        er = relmax(to_np(tr), O.b_residual_mehr(ob, cfg.psi, hp, theta))
        say("mehr boxes", kind, "hloc %.2e S %.2e H %.2e B %.2e residual %.2e" % (eh, es, eH, eB, er))
        assert eh <= 1e-13 and es <= 1e-12 and eH <= 1e-12 and eB <= 1e-12 and er <= 1e-12
    finally:
        bs.close()


def test_mehr_solver_steps_match_oracle(ctx):
    cfg = synth.box_config(nbox=12, gdim=48, h=0.3, edges=(12, 16), nwells=4, seed=10, levels=3)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64, step_length=2.0,
                   precond_start=mgx.START_ZERO, lap_type=mgx.LAP_MEHRSTELLEN)
    try:
        psi = cfg.psi.copy()
        for k in range(4):
            st = s.step()
            psi, e, r = O.b_solver_step_mehr(ob, psi, cfg.vtot, 2, O.F64, 2.0)
            say("mehr step", k, st.band_energy, e, st.residual, r)
            assert st.band_energy == pytest.approx(e, rel=1e-11)
            assert st.residual == pytest.approx(r, rel=1e-9)
        assert np.max(np.abs(s.psi() - psi)) <= 1e-10 * np.max(np.abs(psi))
    finally:
        s.close()
        bs.close()


def test_mehr_solver_anderson_matches_oracle(ctx):
    cfg = synth.box_config(nbox=10, gdim=48, h=0.3, edges=(16,), nwells=3, seed=12, levels=3)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64, step_length=2.0,
                   precond_start=mgx.START_ZERO, anderson_m=1, lap_type=mgx.LAP_MEHRSTELLEN)
    ref = O.BSolver(ob, cfg.psi, cfg.vtot, 2, O.F64, 2.0, anderson_m=1, lap_type=0)
    try:
        engaged = 0
        for k in range(14):
            st = s.step()
            e, r, _ = ref.step()
            engaged = max(engaged, ref.mm)
            say("mehr anderson step", k, st.band_energy, e, st.residual, r, ref.accelerate)
            assert st.band_energy == pytest.approx(e, rel=1e-11)
            assert st.residual == pytest.approx(r, rel=1e-8)
        assert engaged == 1
    finally:
        s.close()
        bs.close()


def test_mehr_solver_rejects_unsupported(ctx):
    cfg = _cfg("even")
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        for kw in (dict(gram_prec=mgx.PREC_SPLIT), dict(precond_dtype=mgx.F16)):
            with pytest.raises(mgx.MgxError):
                mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, lap_type=mgx.LAP_MEHRSTELLEN, **kw)
    finally:
        bs.close()
