"""GPU parity of the Hartree multigrid (mgx_poisson_mg: pb::Mgm + pb::Vcycle)
against the oracle's restatement, through the C ABI, for the Mehrstellen,
2nd- and 4th-order operators: fixed numbers of V-cycles on periodic and on
Dirichlet / mixed grids (1e-12 relative to max |vh|; FP64 sums in another
order), a converged solve (the stopping rule and sweep count), and the
argument checks."""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import mgx
from tests.gpu_util import relmax, say, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def _case(lap_type, dims, h=0.4, seed=0, bc=(1, 1, 1)):
    ng = 2 if lap_type == 2 else 1
    og = O.Grid.make(dims, ng, [h] * 3)
    mg_ = mgx.Grid.make(dims, ng, [h] * 3, bc=bc)
    rng = np.random.default_rng(seed)
    rho = rng.standard_normal((1,) + tuple(dims))
    rho -= rho.mean()
    return og, mg_, O.embed(og, 4 * np.pi * rho)


def _run(ctx, mg_, rho, vh0, lap_type, **kw):
    trho, tvh = to_dev(rho), to_dev(vh0)
    st = mgx.PoissonStats()
    ctx.call("mgx_poisson_mg", lap_type, mgx.block(trho, mg_), mgx.block(tvh, mg_), kw.get("nu1", 2),
             kw.get("nu2", 2), kw.get("max_sweeps", 10), kw.get("tol", 1e-16), kw.get("max_nlevels", 10), st)
    return to_np(tvh), st


@pytest.mark.parametrize("lap_type,dims", [(0, (32, 32, 32)), (1, (16, 24, 32)), (2, (32, 16, 16))])
def test_fixed_cycles_match_oracle(ctx, lap_type, dims):
    og, mg_, rho = _case(lap_type, dims)
    rng = np.random.default_rng(1)
    vh0 = O.embed(og, 0.1 * rng.standard_normal((1,) + tuple(dims)))
    got, st = _run(ctx, mg_, rho, vh0, lap_type, max_sweeps=3, tol=0.0)
    ref = vh0.copy()
    conv, fin, rel, red, sweeps = O.poisson_mgm(og, lap_type, rho, ref, max_sweeps=3, tol=0.0)
    err = relmax(O.interior(og, got), O.interior(og, ref))
    say("poisson type %d dims %s: rel err %.3e, residual %.3e vs %.3e" % (lap_type, dims, err, st.final_residual,
                                                                         fin))
    assert err <= 1e-12
    assert st.nb_sweeps == 3 and not st.converged
    assert st.final_residual == pytest.approx(fin, rel=1e-9)
    assert st.residual_reduction == pytest.approx(red, rel=1e-9)
    # ghosts refreshed: periodic copies of the interior
    g = O.trade(og, got.copy())
    assert np.array_equal(g, got)


@pytest.mark.parametrize("lap_type,dims,bc", [(0, (32, 32, 32), (0, 0, 0)), (1, (16, 24, 32), (1, 0, 1)),
                                             (2, (32, 16, 16), (0, 1, 1)), (0, (16, 32, 24), (1, 1, 0))])
def test_fixed_cycles_dirichlet_match_oracle(ctx, lap_type, dims, bc):
    og, mg_, rho = _case(lap_type, dims, seed=5, bc=bc)
    rng = np.random.default_rng(2)
    vh0 = O.embed(og, 0.1 * rng.standard_normal((1,) + tuple(dims)))
    got, st = _run(ctx, mg_, rho, vh0, lap_type, max_sweeps=3, tol=0.0)
    ref = vh0.copy()
    conv, fin, rel, red, sweeps = O.poisson_mgm(og, lap_type, rho, ref, max_sweeps=3, tol=0.0, bc=bc)
    err = relmax(O.interior(og, got), O.interior(og, ref))
    say("poisson type %d dims %s bc %s: rel err %.3e, residual %.3e vs %.3e" % (lap_type, dims, bc, err,
                                                                                st.final_residual, fin))
    assert err <= 1e-12
    assert st.nb_sweeps == 3 and not st.converged
    assert st.final_residual == pytest.approx(fin, rel=1e-9)
    assert st.residual_reduction == pytest.approx(red, rel=1e-9)
    # ghosts refreshed, the boundary planes zero, no mean removed
    g = O.trade(og, got.copy(), bc)
    assert np.array_equal(g, got)
    assert abs(float(np.mean(O.interior(og, got)))) > 1e-6 * np.max(np.abs(got))


def test_converged_solve(ctx):
    og, mg_, rho = _case(0, (32, 32, 32), seed=3)
    vh0 = O.empty_block(og, 1)
    got, st = _run(ctx, mg_, rho, vh0, 0, max_sweeps=30, tol=1e-11)
    ref = vh0.copy()
    conv, fin, rel, red, sweeps = O.poisson_mgm(og, 0, rho, ref, max_sweeps=30, tol=1e-11)
    say("converged: sweeps %d vs %d, rel residual %.3e" % (st.nb_sweeps, sweeps, st.final_relative_residual))
    assert st.converged and conv and st.nb_sweeps == sweeps
    assert st.final_relative_residual < 1e-11
    assert relmax(O.interior(og, got), O.interior(og, ref)) <= 1e-10
    assert abs(float(np.mean(O.interior(og, got)))) < 1e-13 * np.max(np.abs(got))


def test_rejects_bad_arguments(ctx):
    og, mg_, rho = _case(2, (16, 16, 16))
    g1 = mgx.Grid.make((16, 16, 16), 1, [0.4] * 3)
    r1 = to_dev(O.embed(O.Grid.make((16, 16, 16), 1, [0.4] * 3), np.zeros((1, 16, 16, 16))))
    st = mgx.PoissonStats()
    # 4th order needs 2 ghosts
    with pytest.raises(mgx.MgxError):
        ctx.call("mgx_poisson_mg", 2, mgx.block(r1, g1), mgx.block(r1, g1), 2, 2, 1, 0.0, 10, st)
    # multipole boundary values (bc 2) are not built
    gd = mgx.Grid.make((16, 16, 16), 1, [0.4] * 3, bc=(2, 1, 1))
    with pytest.raises(mgx.MgxError):
        ctx.call("mgx_poisson_mg", 0, mgx.block(r1, gd), mgx.block(r1, gd), 2, 2, 1, 0.0, 10, st)
