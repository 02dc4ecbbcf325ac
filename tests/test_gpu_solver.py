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
    # 3 fine-level sweep launches per V-cycle: the zero-start first sweep is
    # pointwise and fused with the FP64 -> FP32 cast
    assert prof["jacobi0"][1] == 3 * 3
    assert prof["hloc"][1] == 3 and prof["hloc"][0] > 0
    assert s.profile()["hloc"][1] == 0
    s.close()
    bs.close()


def test_solver_anderson_matches_oracle(ctx):
    """ABPG with Anderson extrapolation (m = 1): the engine's stateful steps,
    including the switch-on after the energy settles, follow the oracle's."""
    cfg = _cfg()
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64, anderson_m=1)
    ref = O.BSolver(ob, cfg.psi, cfg.vtot, 2, O.F64, 1.5, anderson_m=1)
    engaged = 0
    for k in range(16):
        st = s.step()
        e, r, _ = ref.step()
        engaged = max(engaged, ref.mm)
        say("anderson step", k, st.band_energy, e, st.residual, r, ref.accelerate)
        assert st.band_energy == pytest.approx(e, rel=1e-11)
        assert st.residual == pytest.approx(r, rel=1e-8)
    assert engaged == 1
    gp = s.psi()
    assert np.max(np.abs(gp - ref.psi)) <= 1e-10 * np.max(np.abs(ref.psi))
    s.close()
    bs.close()


def test_solver_deltae_stops_like_oracle(ctx):
    """Quench.conv_criterion = deltaE: the run stops at the first step after
    the second whose last two band-energy changes are below tol."""
    cfg = _cfg()
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    tol = 2e-3
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64,
                   conv_criterion=mgx.CONV_DELTAE)
    st = s.run(tol, 200)
    ref = O.BSolver(ob, cfg.psi, cfg.vtot, 2, O.F64, 1.5)
    k = 0
    while True:
        e, r, d2 = ref.step()
        if k > 1 and d2 < tol:
            break
        k += 1
    say("deltaE stop", st.iterations, k, st.band_energy, e)
    assert st.converged == 1
    # the run counts the steps that updated psi; the converged evaluation is step k
    assert st.iterations == k
    assert st.band_energy == pytest.approx(e, rel=1e-11)
    s.close()
    bs.close()


def test_solver_learned_fallback_to_vcycle(ctx):
    """A network that points uphill (y = -10 x) is rejected by the residual
    test: psi and R are restored and the V-cycle step is taken, so after the
    rejection the run retraces the plain V-cycle run, and the CNN is tried
    again after MGX_LEARNED_BACKOFF V-cycle steps, then after twice as
    many once it has been rejected again."""
    from mgmol_amd import learned
    cfg = synth.box_config(nbox=16, gdim=48, h=0.3, edges=(16,), nwells=4, seed=11, levels=3)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    p = learned.init_params(seed=0, skip=-10.0)
    for k in range(4):
        p["w%d" % k][:] = 0.0
    net = learned.Cnn(ctx, 16, 8, p)
    kw = dict(mg_levels=2, precond_dtype=mgx.F64)
    plain = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, **kw)
    lrn = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, precond_kind=mgx.PRECOND_LEARNED, cnn=net, **kw)
    ep = [plain.step().band_energy for _ in range(6)]
    sts = [lrn.step() for _ in range(18)]
    el = [st.band_energy for st in sts]
    say("plain", ep)
    say("learned", el, [(st.learned_steps, st.fallbacks) for st in sts])
    assert el[0] == pytest.approx(ep[0], rel=1e-13)
    assert el[1] > el[0]  # the uphill CNN step
    for k in range(1, 6):
        assert el[k + 1] == pytest.approx(ep[k], rel=1e-10)
    # the counters a step reports are those before its own update
    assert sts[1].fallbacks == 1 and sts[1].learned_steps == 1
    # steps 2..5 use the V-cycle; step 6 tries the network again, and step 7
    # rejects it
    assert sts[6].learned_steps == 1 and sts[6].fallbacks == 1
    assert sts[7].learned_steps == 2 and sts[7].fallbacks == 2
    # the second rejection in a row: 8 V-cycle steps (8..15), then the
    # network at step 16, rejected at step 17
    for k in range(8, 17):
        assert sts[k].learned_steps == 2 and sts[k].fallbacks == 2
    assert sts[17].learned_steps == 3 and sts[17].fallbacks == 3
    for s in (plain, lrn):
        s.close()
    net.close()
    bs.close()
