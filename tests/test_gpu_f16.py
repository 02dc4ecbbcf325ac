"""The V-cycle with FP16 storage (MGX_F16): the residual scaled per box to a
fixed RMS, FP16 sweeps and transfers, the FP32 correction with the scale
undone.  Checked against the oracle's FP32 V-cycle (the reference's
MGPRECONDTYPE) to a tolerance set by FP16's 11-bit mantissa, on residuals of
very different sizes (the scaling keeps them in FP16's range), and in the
solver against the plain FP64 run."""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import mgx, synth
from tests.gpu_util import relmax, say, to_dev, to_np

pytestmark = pytest.mark.gpu

# max |out - ref| / max |ref| of the FP16 V-cycle against the FP32 one
TOL_F16 = 5e-3


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def _cfg():
    # 32^3 boxes: edges that stay multiples of 8 on the three levels
    return synth.box_config(nbox=10, gdim=72, h=0.3, edges=(32,), nwells=4, seed=3, levels=3)


@pytest.mark.parametrize("gam_kind", ["zero", "gamma"])
@pytest.mark.parametrize("size", [1.0, 1e-7, 1e3])
def test_f16_precond_against_fp32_oracle(ctx, gam_kind, size):
    cfg = _cfg()
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        rng = np.random.default_rng(5)
        r = rng.standard_normal(ob.size(0)) * size
        # boxes of very different residual sizes in one call
        off = ob.offsets()
        for j in range(ob.nbox):
            r[off[j]:off[j + 1]] *= 10.0 ** (j % 5 - 2)
        gam = 0.0 if gam_kind == "zero" else O.gamma(list(ob.h), 2, 0.0, -1.2)
        tr, to = to_dev(r), to_dev(np.zeros_like(r))
        ctx.call("mgx_box_precond_mg", bs.h, 2, 2, gam, mgx.F16, mgx.ptr(tr), mgx.ptr(to))
        ref = O.b_precond(ob, 2, 2, gam, r, dtype=np.float32)
        got = to_np(to)
        worst = 0.0
        for j in range(ob.nbox):
            e = relmax(got[off[j]:off[j + 1]], ref[off[j]:off[j + 1]])
            worst = max(worst, e)
        say("f16 precond", gam_kind, size, "worst box rel err", worst)
        assert worst <= TOL_F16
    finally:
        bs.close()


def test_f16_precond_needs_edges_multiple_of_8_per_level(ctx):
    cfg = synth.box_config(nbox=8, gdim=48, h=0.3, edges=(8, 12, 16, 20), nwells=2, seed=5, levels=3)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        r = np.ones(int(cfg.psi.size))
        tr, to = to_dev(r), to_dev(np.zeros_like(r))
        with pytest.raises(mgx.MgxError):
            ctx.call("mgx_box_precond_mg", bs.h, 2, 2, 0.0, mgx.F16, mgx.ptr(tr), mgx.ptr(to))
    finally:
        bs.close()


def test_f16_solver_converges_to_the_fp64_energy(ctx):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    kw = dict(mg_levels=2, anderson_m=1, precond_start=mgx.START_ZERO, conv_criterion=mgx.CONV_RESIDUAL)
    out = {}
    for name, pk in (("plain", dict(precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64)),
                     ("f16", dict(precond_dtype=mgx.F16, gram_prec=mgx.PREC_SPLIT))):
        s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, **kw, **pk)
        st = s.run(1e-6, 1500)
        say(name, "iters", st.iterations, "res", st.residual, "E", st.band_energy)
        assert st.converged == 1
        out[name] = st
        s.close()
    bs.close()
    # two runs stopped at |R| < 1e-6 Ha agree in energy to a few 1e-6 Ha
    # (the C5 runs of the FP32 and FP64 preconditioners: 2e-6 Ha)
    assert abs(out["f16"].band_energy - out["plain"].band_energy) <= 5e-6
    # no more than a quarter more steps than the FP64 preconditioner
    assert out["f16"].iterations <= 1.25 * out["plain"].iterations + 10


def test_f16_learned_fallback_to_the_fp16_vcycle(ctx):
    """The learned path over the FP16 V-cycle: a network that points uphill
    loses to the FP16 V-cycle step from the same saved state (its box scales
    then formed from the restored R), after which the run follows the plain
    FP16 run one evaluation behind.  The scales come from sums in another
    order there, so the match is to the FP16 V-cycle's accuracy, not
    bitwise."""
    from mgmol_amd import learned
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    p = learned.init_params(seed=0, skip=-10.0)
    for k in range(4):
        p["w%d" % k][:] = 0.0
    net = learned.Cnn(ctx, 32, 8, p)
    kw = dict(mg_levels=2, precond_dtype=mgx.F16)
    plain = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, **kw)
    lrn = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, precond_kind=mgx.PRECOND_LEARNED, cnn=net, **kw)
    try:
        ep = [plain.step().band_energy for _ in range(6)]
        sts = [lrn.step() for _ in range(7)]
        el = [st.band_energy for st in sts]
        say("plain f16", ep)
        say("learned f16", el)
        assert el[0] == pytest.approx(ep[0], rel=1e-13)
        assert el[1] > el[0]  # the uphill network step
        assert sts[2].fallbacks == 1 and sts[2].learned_steps == 1
        for k in range(2, 7):
            assert el[k] == pytest.approx(ep[k - 1], rel=1e-6)
    finally:
        for s in (plain, lrn):
            s.close()
        net.close()
        bs.close()
