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
    # the skin's algorithmic bytes: counted by brute force on the grid
    assert prof["skin"][1] == 3
    assert prof["skin"][2] == 3 * _skin_bytes(cfg)
    assert s.profile()["hloc"][1] == 0
    s.close()
    bs.close()


def _skin_bytes(cfg):
    """Bytes the skin needs, from the grid points themselves: for every face
    of box j, the two layers beyond it over j's extent (periodic); for each
    box i < j, the points of those layers inside box i (8 bytes each), and,
    when some i meets them, the two planes of psi_j next to the face."""
    g = tuple(int(x) for x in cfg.gdim)
    nbox = len(cfg.edge)

    def box_mask(lo, hi):
        ix = [np.arange(lo[d], hi[d]) % g[d] for d in range(3)]
        m = np.zeros(g, dtype=bool)
        m[np.ix_(*ix)] = True
        return m

    boxes = [box_mask(cfg.origin[i], cfg.origin[i] + cfg.edge[i]) for i in range(nbox)]
    total = 0
    for j in range(nbox):
        o, e = cfg.origin[j], int(cfg.edge[j])
        for d in range(3):
            for side in (0, 1):
                lo, hi = o.copy(), o + e
                lo[d] = o[d] - 2 if side == 0 else o[d] + e
                hi[d] = lo[d] + 2
                slab = box_mask(lo, hi)
                counts = [int(np.count_nonzero(slab & boxes[i])) for i in range(j)]
                if any(counts):
                    total += 8 * sum(counts) + 2 * 8 * e * e
    return total


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
    """A network that points uphill (y = -10 x): the trial's network step is
    evaluated (the energy rises), the solver returns to the saved state and
    takes the V-cycle step from it, evaluates that, and keeps the V-cycle
    step.  So the learned run is the plain V-cycle run with one extra
    evaluation per trial; the network is tried again after
    MGX_LEARNED_BACKOFF V-cycle steps, then after twice as many once it has
    lost again."""
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
    ep = [plain.step().band_energy for _ in range(8)]
    sts = [lrn.step() for _ in range(19)]
    el = [st.band_energy for st in sts]
    say("plain", ep)
    say("learned", el, [(st.learned_steps, st.fallbacks) for st in sts])
    assert el[0] == pytest.approx(ep[0], rel=1e-13)
    assert el[1] > el[0]  # the uphill network step, evaluated
    # step 1 returns to the saved state and takes the V-cycle step, which
    # step 2 evaluates: from there on the plain run's states
    for k in range(2, 8):
        assert el[k] == pytest.approx(ep[k - 1], rel=1e-10)
    # a step reports the counters after its own decision
    assert sts[0].learned_steps == 1 and sts[0].fallbacks == 0
    assert sts[1].learned_steps == 1 and sts[1].fallbacks == 0
    for k in range(2, 7):  # lost at step 2, then 4 V-cycle steps (3..6)
        assert sts[k].learned_steps == 1 and sts[k].fallbacks == 1
    assert sts[7].learned_steps == 2  # the second trial
    assert el[8] > el[7]
    assert el[9] == pytest.approx(ep[7], rel=1e-10)
    # lost again at step 9: 8 V-cycle steps (10..17), the third trial at 18
    for k in range(9, 18):
        assert sts[k].learned_steps == 2 and sts[k].fallbacks == 2
    assert sts[18].learned_steps == 3
    for s in (plain, lrn):
        s.close()
    net.close()
    bs.close()


def test_solver_refresh_branches(ctx):
    """The Newton-Schulz refresh's checked branches.  With at most one
    unchecked step per refresh (ns_max_spec = 1) the solver continues with
    checked steps whenever one is not enough (ns_continued); orbitals replaced
    by very different ones make the refresh diverge from the old inverse, and
    Cholesky takes over (ns_restarts).  Energies and residuals equal those of
    a run with Cholesky + inverse every step (MGX_INV_CHOL, the reference's
    GramMatrix::updateLS / computeInverse)."""
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    kw = dict(mg_levels=2, precond_dtype=mgx.F64, anderson_m=1)
    chol = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, inverse_mode=mgx.INV_CHOL, **kw)
    capped = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, ns_max_spec=1, **kw)
    free = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, **kw)
    try:
        for k in range(12):
            a, b, c = chol.step(), capped.step(), free.step()
            say("refresh step", k, a.band_energy, b.band_energy, c.band_energy, b.ns_continued, c.ns_continued)
            for x in (b, c):
                assert x.band_energy == pytest.approx(a.band_energy, rel=1e-12)
                assert x.residual == pytest.approx(a.residual, rel=1e-9)
        assert a.chol_calls == 12 and a.ns_steps == 0
        assert b.chol_calls == 1 and b.ns_continued > 0 and b.ns_restarts == 0
        assert c.ns_restarts == 0
        # very different orbitals: the old inverse is far off, the refresh
        # diverges and Cholesky restarts it
        rng = np.random.default_rng(4)
        psi2 = cfg.psi + 0.3 * np.max(np.abs(cfg.psi)) * rng.uniform(-1.0, 1.0, cfg.psi.size)
        capped.set_psi(psi2)
        fresh = mgx.Solver(ctx, bs, cfg.vtot, psi2, **kw)
        b2, f2 = capped.step(), fresh.step()
        say("after set_psi", b2.band_energy, f2.band_energy, b2.ns_restarts, b2.chol_calls)
        assert b2.ns_restarts == 1 and b2.chol_calls == 2
        assert b2.band_energy == pytest.approx(f2.band_energy, rel=1e-12)
        assert b2.residual == pytest.approx(f2.residual, rel=1e-9)
        fresh.close()
    finally:
        for s in (chol, capped, free):
            s.close()
        bs.close()
