"""Pins of the oracle's Hartree multigrid (oracle.poisson_mgm, restating
pb::Mgm / pb::Vcycle):
  * on periodic grids the converged potential equals the exact discrete
    solution A v = rhs, computed by FFT from the operator's own stencil (A
    applied to a delta; the stencils are pinned by their Fourier symbols in
    test_oracle_stencils), with the mean removed as GridFunc::average0 does;
  * with Dirichlet directions the index-0 layer is the boundary plane (zero,
    with the ghosts, as GridFunc::setBoundaryValues sets it) and the
    converged potential equals the solution of the dense system over the
    free points, built column by column from the operator; no mean is
    removed; for the 2nd-order operator with all directions Dirichlet, a sine
    mode sin(pi k i / n), i = 0 .. n-1, is an eigenvector with the closed-form
    eigenvalue sum_i (2 - 2 cos(pi k_i / n_i)) / h_i^2 (a cell-centred
    reading, zero ghosts only, would need n + 1 and fails it);
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


def _free(dims, bc):
    m = np.ones(dims, dtype=bool)
    for k in range(3):
        if bc[k] != 1:
            sl = [slice(None)] * 3
            sl[k] = 0
            m[tuple(sl)] = False
    return m


def _dense_solution(g, lap_type, bc, rhs_int):
    free = _free(tuple(g.dim), bc)
    idx = np.flatnonzero(free)
    A = np.zeros((idx.size, idx.size))
    for c, j in enumerate(idx):
        e = np.zeros(int(np.prod(g.dim)))
        e[j] = 1.0
        blk = O.embed(g, e.reshape((1,) + tuple(g.dim)))
        O.trade(g, blk, bc)
        A[:, c] = O.interior(g, O.lap(g, lap_type, blk)).reshape(-1)[idx]
    out = np.zeros(tuple(g.dim))
    out.reshape(-1)[idx] = np.linalg.solve(A, rhs_int.reshape(-1)[idx])
    return out


@pytest.mark.parametrize("lap_type", [0, 1, 2])
@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 0, 1), (0, 1, 1)])
def test_mgm_dirichlet_matches_dense_solve(lap_type, bc):
    dims = (8, 8, 8)
    g, rho = _setup(lap_type, dims=dims, h=0.5, seed=4)
    rng = np.random.default_rng(5)
    rho = O.embed(g, rng.standard_normal((1,) + dims))
    vh = O.empty_block(g, 1)
    conv, final, rel, red, sweeps = O.poisson_mgm(g, lap_type, rho, vh, max_sweeps=60, tol=1e-12, bc=bc)
    assert conv and rel < 1e-12
    r = rho.copy()
    O.trade(g, r, bc)
    rhs = O.rhs_mehr(g, r) if lap_type == 0 else r
    ref = _dense_solution(g, lap_type, bc, O.interior(g, rhs)[0])
    got = O.interior(g, vh)[0]
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    # the boundary plane holds zero; no mean removal (the mean is nonzero)
    assert np.all(got[~_free(dims, bc)] == 0.0)
    assert abs(got.mean()) > 1e-3 * np.max(np.abs(ref))


def test_mgm_dirichlet_sine_mode_closed_form():
    dims, h, k = (16, 12, 8), (0.3, 0.4, 0.5), (1, 2, 3)
    g = O.Grid.make(dims, 1, list(h))
    ax = [np.sin(np.pi * k[i] * np.arange(dims[i]) / dims[i]) for i in range(3)]
    mode = np.einsum("i,j,k->ijk", *ax)
    lam = sum((2 - 2 * np.cos(np.pi * k[i] / dims[i])) / h[i] ** 2 for i in range(3))
    vh = O.empty_block(g, 1)
    conv, _, rel, _, _ = O.poisson_mgm(g, 1, O.embed(g, lam * mode[None]), vh, max_sweeps=60, tol=1e-13,
                                       bc=(0, 0, 0))
    assert conv
    assert np.max(np.abs(O.interior(g, vh)[0] - mode)) <= 1e-11


def test_dirichlet_trade_zeroes_boundary_plane():
    """The meshes and ghost widths of tests/testSetGhostValues.cc; the check
    of GridFunc::test_setBoundaryValues (src/pb/GridFunc.cc:3062-3095): after
    the Dirichlet boundary values are set on a function of ones, its dot
    product with itself is gsize - n0 n1 - n0 (n2 - 1) - (n1 - 1)(n2 - 1) =
    (n0 - 1)(n1 - 1)(n2 - 1); the free points 1 .. n-1 keep their value."""
    for ng in (1, 2):
        for dims in ((16, 24, 20), (20, 32, 24), (24, 20, 32)):
            g = O.Grid.make(dims, ng, [1.0 / d for d in dims])
            u = O.embed(g, np.full((1,) + dims, 1.111))
            O.trade(g, u, (0, 0, 0))
            inner = O.interior(g, u)[0]
            assert np.all(inner[1:, 1:, 1:] == 1.111)
            assert np.count_nonzero(u) == (dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1)
