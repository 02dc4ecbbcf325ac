"""Pins of the oracle's Hartree multigrid (oracle.poisson_mgm, restating
pb::Mgm / pb::Vcycle) on periodic grids:
  * the converged potential equals the exact discrete solution A v = rhs,
    computed by FFT from the operator's own stencil (A applied to a delta;
    the stencils are pinned by their Fourier symbols in
    test_oracle_stencils), with the mean removed as GridFunc::average0 does;
  * every V-cycle contracts the residual (a multigrid property, not a
    restatement);
  * the stopping rule: a tolerance met before any cycle leaves vh unchanged
    up to the mean.
The individual iterates follow the restatement only: no reference vectors
exist for them (parity unpinned beyond the converged solution)."""
import numpy as np
import pytest

import oracle as O


def _fft_solution(g, lap_type, rhs_int):
    delta = np.zeros((1,) + tuple(g.dim))
    delta[0, 0, 0, 0] = 1.0
    blk = O.embed(g, delta)
    O.trade(g, blk)
    kern = O.interior(g, O.lap(g, lap_type, blk))[0]
    lam = np.fft.fftn(kern)
    rh = np.fft.fftn(rhs_int)
    lam[0, 0, 0] = 1.0
    vh = rh / lam
    vh[0, 0, 0] = 0.0
    return np.real(np.fft.ifftn(vh))


def _setup(lap_type, dims=(16, 16, 16), h=0.4, seed=0):
    ng = 2 if lap_type == 2 else 1
    g = O.Grid.make(dims, ng, [h] * 3)
    rng = np.random.default_rng(seed)
    rho = rng.standard_normal((1,) + tuple(dims))
    rho -= rho.mean()
    return g, O.embed(g, rho)


@pytest.mark.parametrize("lap_type", [0, 1, 2])
def test_mgm_converges_to_fft_solution(lap_type):
    g, rho = _setup(lap_type)
    vh = O.empty_block(g, 1)
    conv, final, rel, red, sweeps = O.poisson_mgm(g, lap_type, rho, vh, max_sweeps=40, tol=1e-12)
    assert conv and sweeps < 40 and rel < 1e-12
    r = rho.copy()
    O.trade(g, r)
    rhs = O.rhs_mehr(g, r) if lap_type == 0 else r
    ref = _fft_solution(g, lap_type, O.interior(g, rhs)[0])
    got = O.interior(g, vh)[0]
    assert abs(got.mean()) < 1e-14 * np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("lap_type,bound", [(0, 0.15), (1, 0.3), (2, 0.4)])
def test_vcycle_contracts(lap_type, bound):
    """Measured factors per cycle on this 16^3 case: Mehrstellen 0.08-0.11,
    2nd order 0.19-0.25, 4th order over 2nd-order coarse levels 0.27-0.33."""
    g, rho = _setup(lap_type, seed=1)
    prev = None
    for k in range(1, 6):
        vh = O.empty_block(g, 1)
        _, final, _, _, sweeps = O.poisson_mgm(g, lap_type, rho, vh, max_sweeps=k, tol=0.0)
        assert sweeps == k
        if prev is not None:
            assert final < bound * prev
        prev = final


def test_tolerance_met_before_any_cycle():
    g, rho = _setup(0, seed=2)
    vh = O.embed(g, np.full((1, 16, 16, 16), 0.5))
    conv, _, _, _, sweeps = O.poisson_mgm(g, 0, rho, vh, max_sweeps=5, tol=1e3)
    assert conv and sweeps == 0
    assert np.max(np.abs(O.interior(g, vh))) == 0.0
