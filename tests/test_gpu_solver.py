"""GPU orbital iteration vs the oracle's composed step (FP64 plain mode),
and convergence of the mixed modes on a small box problem."""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import mgx, synth
from tests.gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def _cfg():
    return synth.box_config(nbox=20, gdim=48, h=0.3, edges=(12, 16, 20), nwells=4, seed=9, levels=3)


@pytest.mark.parametrize("start", ["gamma", "zero"])
def test_solver_steps_match_oracle(ctx, start):
    cfg = _cfg()
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    code = mgx.START_GAMMA if start == "gamma" else mgx.START_ZERO
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64, precond_start=code)
    psi = cfg.psi.copy()
    lowest = None
    for k in range(4):
        st = s.step()
        psi, e, r, lowest = O.b_solver_step(ob, psi, cfg.vtot, 2, O.F64, 1.5, lowest, zero_start=(start == "zero"))
        say("step", k, st.band_energy, e, st.residual, r)
        assert st.band_energy == pytest.approx(e, rel=1e-11)
        assert st.residual == pytest.approx(r, rel=1e-9)
    gp = s.psi()
    assert np.max(np.abs(gp - psi)) <= 1e-11 * np.max(np.abs(psi))
    s.close()
    bs.close()


@pytest.mark.parametrize("mode", ["plain", "f32precond", "split"])
def test_solver_converges(ctx, mode):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    kw = dict(mg_levels=2, precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64)
    if mode != "plain":
        kw["precond_dtype"] = mgx.F32
    if mode == "split":
        kw["gram_prec"] = mgx.PREC_SPLIT
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, **kw)
    first = s.step()
    st = s.run(1e-6, 600)
    say(mode, "iters", st.iterations, "res", st.residual, "E", st.band_energy, "t", st.seconds)
    assert st.residual < first.residual
    assert st.converged == 1
    # restart reproduces the first step
    s.reset()
    again = s.step()
    assert again.band_energy == first.band_energy
    s.close()
    bs.close()


def test_solver_profile_groups(ctx):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F32, profile=1)
    for _ in range(3):
        s.step()
    prof = s.profile(reset=True)
    say(prof)
    assert prof["jacobi0"][1] == 3 * 4
    assert prof["hloc"][1] == 3 and prof["hloc"][0] > 0
    assert s.profile()["hloc"][1] == 0
    s.close()
    bs.close()
