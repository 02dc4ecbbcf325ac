"""GPU parity of the electron density (mgx_box_density, mgx_solver_density)
against the oracle (Rho::computeRhoSubdomainUsingBlas3 and
rescaleTotalCharge, src/Rho.cc:122-208, 360-449), through the C ABI.
Tolerance: 1e-13 relative to max |rho| (FP64 MFMA sums in another order)."""
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


def _setup(ctx, seed=5, nbox=24, gdim=48, edges=(8, 12, 16, 20)):
    cfg = synth.box_config(nbox=nbox, gdim=gdim, h=0.3, edges=edges, nwells=2, seed=seed, levels=3)
    b = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    return cfg, b, bs


def test_density_random_dm_and_two_inverse_s(ctx):
    cfg, b, bs = _setup(ctx)
    n = cfg.nbox
    rng = np.random.default_rng(2)
    a = rng.standard_normal((n, n))
    psi = to_dev(cfg.psi)
    rho = mgx.DeviceArray(ctx, shape=tuple(cfg.gdim), dtype=np.float64)
    vel = cfg.h ** 3
    for dm in (a + a.T, 2.0 * np.linalg.inv(O.b_gram_h(b, cfg.psi, cfg.psi)[0])):
        q = bs.density(psi, to_dev(np.asfortranarray(dm).ravel(order="F")), rho)
        ref = O.b_density(b, cfg.psi, dm)
        err = relmax(to_np(rho), ref)
        say("density rel err %.3e, charge %.15g vs %.15g" % (err, q, O.charge(ref, vel)))
        assert err <= 1e-13
        assert q == pytest.approx(O.charge(ref, vel), rel=1e-12)
    # dm = 2 S^-1: 2 electrons per orbital
    assert q == pytest.approx(2.0 * n, rel=1e-11)
    bs.close()


def test_density_rescale_and_error(ctx):
    cfg, b, bs = _setup(ctx, seed=6)
    S = O.b_gram_h(b, cfg.psi, cfg.psi)[0]
    dm = to_dev(2.0 * np.linalg.inv(S))
    psi = to_dev(cfg.psi)
    rho = mgx.DeviceArray(ctx, shape=tuple(cfg.gdim), dtype=np.float64)
    q = bs.density(psi, dm, rho)
    ref, q0 = O.rescale_charge(O.b_density(b, cfg.psi, 2.0 * np.linalg.inv(S)), cfg.h ** 3, 1.015 * q)
    q1 = bs.density(psi, dm, rho, nel=1.015 * q)
    assert q1 == pytest.approx(q0, rel=1e-12)
    assert relmax(to_np(rho), ref) <= 1e-13
    with pytest.raises(mgx.MgxError, match="rc=-4"):
        bs.density(psi, dm, rho, nel=1.05 * q)
    bs.close()


def test_density_needs_cells(ctx):
    # a 36^3 grid is not a whole number of 8^3 cells
    cfg = synth.box_config(nbox=4, gdim=36, h=0.3, edges=(8,), nwells=1, seed=1, levels=3)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    rho = mgx.DeviceArray(ctx, shape=tuple(cfg.gdim), dtype=np.float64)
    with pytest.raises(mgx.MgxError, match="rc=-1"):
        bs.density(to_dev(cfg.psi), to_dev(np.eye(4)), rho)
    bs.close()


def test_solver_density(ctx):
    cfg, b, bs = _setup(ctx, seed=7, nbox=16, edges=(12, 16))
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, precond_dtype=mgx.F64)
    for _ in range(3):
        s.step()
    rho = mgx.DeviceArray(ctx, shape=tuple(cfg.gdim), dtype=np.float64)
    q = s.density(rho)
    psi = s.psi()
    S = O.b_gram_h(b, psi, psi)[0]
    ref = O.b_density(b, psi, 2.0 * np.linalg.inv(S))
    err = relmax(to_np(rho), ref)
    say("solver density rel err %.3e, charge %.15g" % (err, q))
    assert err <= 1e-12
    assert q == pytest.approx(2.0 * cfg.nbox, rel=1e-11)
    # the solver carries on after the density call
    st = s.step()
    assert np.isfinite(st.band_energy)
    s.close()
    bs.close()
