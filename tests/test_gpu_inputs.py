"""The box solver on the reference's own test inputs (tests/EnergyAndForces:
six localization centres of N2 on a 64^3 mesh, Gaussian initial orbitals of
width 1.5 bohr; wells at the atoms stand in for the pseudopotentials): the
first steps match the oracle's, the residual falls, and the density of the
result holds two electrons per orbital."""
import os

import numpy as np
import pytest

import oracle as O
from mgmol_amd import inputs, mgx
from tests.gpu_util import relmax, say

pytestmark = pytest.mark.gpu
D = os.path.join(os.path.dirname(__file__), "golden", "energy_and_forces")


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def test_solver_on_reference_inputs(ctx):
    cfg = inputs.box_setup(os.path.join(D, "mgmol.cfg"), os.path.join(D, "lrs.in"), os.path.join(D, "coords.in"),
                           edge=28, levels=3)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64, step_length=2.0,
                   precond_start=mgx.START_ZERO)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    psi = cfg.psi.copy()
    for k in range(3):
        st = s.step()
        psi, e, r, _ = O.b_solver_step(ob, psi, cfg.vtot, 2, O.F64, 2.0, zero_start=True)
        say("step %d: E %.12f vs %.12f, |R| %.3e vs %.3e" % (k, st.band_energy, e, st.residual, r))
        assert st.band_energy == pytest.approx(e, rel=1e-11)
        assert st.residual == pytest.approx(r, rel=1e-9)
    r0 = st.residual
    for _ in range(60):
        st = s.step()
    say("after 63 steps: E %.10f |R| %.3e (from %.3e)" % (st.band_energy, st.residual, r0))
    assert st.residual < 0.2 * r0
    rho = mgx.DeviceArray(ctx, shape=cfg.gdim, dtype=np.float64)
    q = s.density(rho)
    assert q == pytest.approx(2.0 * cfg.nbox, rel=1e-11)
    s.close()
    bs.close()


def test_solver_on_reference_inputs_with_their_fdtype(ctx):
    """The same inputs with their own discretisation, FDtype=Mehrstellen
    (H = -Lap_M + B V, Galerkin B-matrix, residual (B Psi) theta - H Psi,
    V-cycle with the Mehrstellen fine operator): the steps match the
    oracle's Mehrstellen steps, the residual falls."""
    cfgp = os.path.join(D, "mgmol.cfg")
    lap = inputs.fd_type(cfgp)
    assert lap == mgx.LAP_MEHRSTELLEN
    cfg = inputs.box_setup(cfgp, os.path.join(D, "lrs.in"), os.path.join(D, "coords.in"), edge=28, levels=3)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64, step_length=2.0,
                   precond_start=mgx.START_ZERO, lap_type=lap)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    psi = cfg.psi.copy()
    for k in range(3):
        st = s.step()
        psi, e, r = O.b_solver_step_mehr(ob, psi, cfg.vtot, 2, O.F64, 2.0)
        say("Mehrstellen step %d: E %.12f vs %.12f, |R| %.3e vs %.3e" % (k, st.band_energy, e, st.residual, r))
        assert st.band_energy == pytest.approx(e, rel=1e-11)
        assert st.residual == pytest.approx(r, rel=1e-9)
    r0 = st.residual
    for _ in range(60):
        st = s.step()
    say("Mehrstellen after 63 steps: E %.10f |R| %.3e (from %.3e)" % (st.band_energy, st.residual, r0))
    assert st.residual < 0.2 * r0
    s.close()
    bs.close()
