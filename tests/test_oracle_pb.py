"""Pins of the oracle's Poisson-Boltzmann multigrid (oracle.pb_mgm, restating
SolverPB<PBh4M>: the dielectric operator -div(eps grad) in the transformed
form sqrt(eps)^-1 (-Lap_M + B(pp .)) sqrt(eps)^-1, pp = Lap_4 sqrt(eps) /
sqrt(eps)):
  * the dielectric function and its derivative: the limits (eps -> e0 at low
    density, -> 1 at high), eps(rho0) = (1 + e0)/2, and d eps / d rho equal to
    a central difference of eps;
  * pp equals the 4th-order stencil's Fourier symbol applied to sqrt(eps)
    spectrally, divided by sqrt(eps) (sign and operator order);
  * the converged potential on a Dirichlet grid equals the dense solve of the
    fine-level transformed operator over the free points (the V-cycle only
    decides how fast it gets there);
  * with a constant density (eps constant, pp ~ 0) the potential is the
    Mehrstellen Poisson potential divided by eps;
  * the 4th-order gradient of get_vepsilon on a sine mode has the symbol
    (8 sin kh - sin 2kh) / (6h), and vks = d eps/d rho |grad vh|^2 (-1/(8 pi)).
The coarse levels' pp follows the restatement only (parity unpinned beyond the
converged solution, which they do not change).

The gradient forms (oracle.pbg_mgm, restating SolverPB<PBh4> and
SolverPB<PBh2>: eps (-Lap x) - grad x . grad eps, PBh2 on the coarse levels):
  * on eps = 1 + a cos(k1 r) and x = sin(k2 r) the operator equals its closed
    form eps lam(k2) sin(k2 r) + a cos(k2 r) sin(k1 r) sum_d s_d(k2) s_d(k1)
    (lam, s_d: the Fourier symbols of -Lap and of the first difference of
    each order), which fixes the sign, the pairing of the directions and the
    order of each term;
  * the converged potential equals the dense solve of the fine operator over
    the free points (Dirichlet and mixed grids);
  * with a constant density the potential is the Poisson potential of the
    same order divided by eps;
  * the 2nd-order gradient of get_vepsilon has the symbol sin(kh) / h."""
import numpy as np
import pytest

import oracle as O

E0, RHO0, DRHO0 = 78.36, 0.0004, 1.3


def _sym4(k, h):
    return sum((30 - 32 * np.cos(kk * hh) + 2 * np.cos(2 * kk * hh)) / (12 * hh * hh) for kk, hh in zip(k, h))


def _density(dims, h, seed):
    """a smooth positive density with values across rho0 (two Gaussian blobs)."""
    rng = np.random.default_rng(seed)
    r = [np.arange(n) * hh for n, hh in zip(dims, h)]
    X, Y, Z = np.meshgrid(*r, indexing="ij")
    L = [n * hh for n, hh in zip(dims, h)]
    out = np.zeros(dims)
    for _ in range(2):
        c = rng.uniform(0.3, 0.7, 3) * L
        out += 0.01 * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / 0.8)
    return out


def test_dielectric_function_and_derivative():
    rho = np.array([1e-9, RHO0, 1e3 * RHO0])
    e = O.pb_epsilon(rho, E0, RHO0, DRHO0)
    assert e[0] == pytest.approx(E0, rel=1e-9)
    assert e[1] == pytest.approx(0.5 * (1 + E0), rel=1e-14)
    assert e[2] == pytest.approx(1.0, abs=1e-5)
    x = np.linspace(0.002 * RHO0, 5 * RHO0, 50)
    dx = 1e-7 * RHO0
    fd = (O.pb_epsilon(x + dx, E0, RHO0, DRHO0) - O.pb_epsilon(x - dx, E0, RHO0, DRHO0)) / (2 * dx)
    assert np.allclose(O.pb_depsilon(x, E0, RHO0, DRHO0), fd, rtol=1e-6, atol=1e-6 * np.max(np.abs(fd)))
    assert O.pb_depsilon(np.array([0.0005 * RHO0]), E0, RHO0, DRHO0)[0] == 0.0


def test_pp_is_the_spectral_laplacian_of_sqrt_eps():
    dims, h = (16, 12, 20), (0.3, 0.35, 0.25)
    g = O.Grid.make(dims, 2, list(h))
    rhod = _density(dims, h, 1)
    sa, isa, pp = O.pb_init(g, rhod, E0, RHO0, DRHO0)
    s = O.interior(g, sa)[0]
    k = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(n, d=hh) for n, hh in zip(dims, h)], indexing="ij")
    lam = _sym4(k, h)
    lap_s = np.real(np.fft.ifftn(-lam * np.fft.fftn(s)))
    ref = lap_s / s
    got = O.interior(g, pp)[0]
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert np.max(np.abs(got)) > 1.0  # the density varies: pp is not small


def _dense_pb(g, pp, isa, rho, bc):
    free = np.ones(tuple(g.dim), dtype=bool)
    for k in range(3):
        if bc[k] != 1:
            sl = [slice(None)] * 3
            sl[k] = 0
            free[tuple(sl)] = False
    idx = np.flatnonzero(free)
    A = np.zeros((idx.size, idx.size))
    for c, j in enumerate(idx):
        e = np.zeros(int(np.prod(g.dim)))
        e[j] = 1.0
        col = O.pb_apply(g, pp, O.embed(g, e.reshape((1,) + tuple(g.dim))), bc)
        A[:, c] = O.interior(g, col)[0].reshape(-1)[idx]
    res = rho * isa
    O.trade(g, res, bc)
    rhs = O.interior(g, O.rhs_mehr(g, res))[0].reshape(-1)
    u = np.zeros(int(np.prod(g.dim)))
    u[idx] = np.linalg.solve(A, rhs[idx])
    return u.reshape(tuple(g.dim)) * O.interior(g, isa)[0]


@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 0, 1)])
def test_converged_potential_is_the_dense_solution(bc):
    dims, h = (8, 8, 8), (0.5, 0.5, 0.5)
    g = O.Grid.make(dims, 2, list(h))
    rhod = _density(dims, h, 2)
    rho = O.embed(g, np.random.default_rng(3).standard_normal((1,) + dims))
    vh = O.empty_block(g, 1)
    conv, fin, rel, red, sweeps, vks = O.pb_mgm(g, rho, rhod, vh, E0, RHO0, DRHO0, max_sweeps=200, tol=1e-12, bc=bc)
    assert conv and rel < 1e-12 and sweeps > 1
    _, isa, pp = O.pb_init(g, rhod, E0, RHO0, DRHO0)
    ref = _dense_pb(g, pp, isa, rho, bc)
    got = O.interior(g, vh)[0]
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_constant_density_gives_the_poisson_potential_over_eps():
    dims, h = (16, 16, 16), (0.4, 0.4, 0.4)
    g = O.Grid.make(dims, 2, list(h))
    bc = (0, 0, 0)
    rhod = np.full(dims, 0.5 * RHO0)
    eps = float(O.pb_epsilon(np.array([0.5 * RHO0]), E0, RHO0, DRHO0)[0])
    rng = np.random.default_rng(4)
    rho = O.embed(g, rng.standard_normal((1,) + dims))
    vh = O.empty_block(g, 1)
    conv, *_ = O.pb_mgm(g, rho, rhod, vh, E0, RHO0, DRHO0, max_sweeps=60, tol=1e-13, bc=bc)
    vp = O.empty_block(g, 1)
    convp, *_ = O.poisson_mgm(g, 0, rho, vp, max_sweeps=60, tol=1e-13, bc=bc)
    assert conv and convp
    got, ref = O.interior(g, vh)[0], O.interior(g, vp)[0] / eps
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_vepsilon_gradient_symbol():
    dims, h = (16, 12, 20), (0.3, 0.35, 0.25)
    g = O.Grid.make(dims, 2, list(h))
    kv = [2 * np.pi * 2 / (dims[0] * h[0]), 2 * np.pi * 1 / (dims[1] * h[1]), 2 * np.pi * 3 / (dims[2] * h[2])]
    r = np.meshgrid(*[np.arange(n) * hh for n, hh in zip(dims, h)], indexing="ij")
    ph = kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2]
    u = O.embed(g, np.sin(ph)[None])
    O.trade(g, u)
    for d in range(3):
        sym = (8 * np.sin(kv[d] * h[d]) - np.sin(2 * kv[d] * h[d])) / (6 * h[d])
        assert np.max(np.abs(O.pb_del1(g, u, d) - sym * np.cos(ph))) < 1e-12


def _sym2(k, h):
    return sum((2 - 2 * np.cos(kk * hh)) / (hh * hh) for kk, hh in zip(k, h))


def _d1sym(order, k, h):
    if order == 4:
        return (8 * np.sin(k * h) - np.sin(2 * k * h)) / (6 * h)
    return np.sin(k * h) / h


@pytest.mark.parametrize("order", [4, 2])
def test_gradient_operator_closed_form(order):
    dims, h = (16, 12, 20), (0.3, 0.35, 0.25)
    g = O.Grid.make(dims, 2, list(h))
    r = np.meshgrid(*[np.arange(n) * hh for n, hh in zip(dims, h)], indexing="ij")
    k1 = [2 * np.pi * m / (n * hh) for m, n, hh in zip((1, 2, 1), dims, h)]
    k2 = [2 * np.pi * m / (n * hh) for m, n, hh in zip((2, 1, 3), dims, h)]
    p1 = sum(k * x for k, x in zip(k1, r))
    p2 = sum(k * x for k, x in zip(k2, r))
    a = 0.3
    e = 1.0 + a * np.cos(p1)
    eps = O.embed(g, e[None])
    O.trade(g, eps)
    x = O.embed(g, np.sin(p2)[None])
    got = O.interior(g, O.pbg_apply(g, eps, x, (1, 1, 1), order))[0]
    lam = _sym4(k2, h) if order == 4 else _sym2(k2, h)
    cross = sum(_d1sym(order, kb, hh) * _d1sym(order, ka, hh) for ka, kb, hh in zip(k1, k2, h))
    ref = e * lam * np.sin(p2) + a * np.cos(p2) * np.sin(p1) * cross
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))


def _dense_pbg(g, eps, rho, bc, order):
    free = np.ones(tuple(g.dim), dtype=bool)
    for k in range(3):
        if bc[k] != 1:
            sl = [slice(None)] * 3
            sl[k] = 0
            free[tuple(sl)] = False
    idx = np.flatnonzero(free)
    A = np.zeros((idx.size, idx.size))
    for c, j in enumerate(idx):
        e = np.zeros(int(np.prod(g.dim)))
        e[j] = 1.0
        col = O.pbg_apply(g, eps, O.embed(g, e.reshape((1,) + tuple(g.dim))), bc, order)
        A[:, c] = O.interior(g, col)[0].reshape(-1)[idx]
    rhs = O.interior(g, rho)[0].reshape(-1)
    u = np.zeros(int(np.prod(g.dim)))
    u[idx] = np.linalg.solve(A, rhs[idx])
    return u.reshape(tuple(g.dim))


@pytest.mark.parametrize("order", [4, 2])
@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 0, 1)])
def test_gradient_converged_potential_is_the_dense_solution(order, bc):
    # eps from 1 to 10 (with e0 = 78.36 on this 8^3 grid of h = 0.5 the
    # 4th-order gradient form's cycles grow instead of converging)
    e0 = 10.0
    dims, h = (8, 8, 8), (0.5, 0.5, 0.5)
    g = O.Grid.make(dims, 2, list(h))
    rhod = _density(dims, h, 2)
    rho = O.embed(g, np.random.default_rng(3).standard_normal((1,) + dims))
    vh = O.empty_block(g, 1)
    conv, fin, rel, red, sweeps, vks = O.pbg_mgm(g, rho, rhod, vh, e0, RHO0, DRHO0, order=order, max_sweeps=100,
                                                 tol=1e-12, bc=bc)
    assert conv and rel < 1e-12 and sweeps > 1
    eps = O.embed(g, O.pb_epsilon(rhod, e0, RHO0, DRHO0)[None])
    O.trade(g, eps)
    ref = _dense_pbg(g, eps, rho, bc, order)
    got = O.interior(g, vh)[0]
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


@pytest.mark.parametrize("order", [4, 2])
def test_gradient_constant_density_gives_the_poisson_potential_over_eps(order):
    dims, h = (16, 16, 16), (0.4, 0.4, 0.4)
    g = O.Grid.make(dims, 2, list(h))
    bc = (0, 0, 0)
    rhod = np.full(dims, 0.5 * RHO0)
    eps = float(O.pb_epsilon(np.array([0.5 * RHO0]), E0, RHO0, DRHO0)[0])
    rng = np.random.default_rng(4)
    rho = O.embed(g, rng.standard_normal((1,) + dims))
    vh = O.empty_block(g, 1)
    conv, *_ = O.pbg_mgm(g, rho, rhod, vh, E0, RHO0, DRHO0, order=order, max_sweeps=120, tol=1e-13, bc=bc)
    vp = O.empty_block(g, 1)
    convp, *_ = O.poisson_mgm(g, 2 if order == 4 else 1, rho, vp, max_sweeps=120, tol=1e-13, bc=bc)
    assert conv and convp
    got, ref = O.interior(g, vh)[0], O.interior(g, vp)[0] / eps
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_vepsilon_2nd_order_gradient_symbol():
    dims, h = (16, 12, 20), (0.3, 0.35, 0.25)
    g = O.Grid.make(dims, 1, list(h))
    kv = [2 * np.pi * 2 / (dims[0] * h[0]), 2 * np.pi * 1 / (dims[1] * h[1]), 2 * np.pi * 3 / (dims[2] * h[2])]
    r = np.meshgrid(*[np.arange(n) * hh for n, hh in zip(dims, h)], indexing="ij")
    ph = kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2]
    u = O.embed(g, np.sin(ph)[None])
    O.trade(g, u)
    for d in range(3):
        assert np.max(np.abs(O.pb_del1_2nd(g, u, d) - np.sin(kv[d] * h[d]) / h[d] * np.cos(ph))) < 1e-12
