"""Pins for the oracle's multigrid pieces (CPU only).

Restriction/prolongation: tests/testMGkernels.cc restated, R = P^T / 8 as a
matrix identity, exact tensor-product symbol of full weighting on a mode,
exact reproduction of linear data by prolongation.  Jacobi: amplification per
Fourier mode.  V-cycle: linearity, masks of ones == no masks, and the
closed-form result for a constant right-hand side, which fixes the sweep
counts per level, the sign of the coarse correction and the coarse factors.
"""
import numpy as np
import pytest

import oracle as O


def _g(dims, ng=2, ll=2.0):
    return O.Grid.make(dims, ng, [ll / d for d in dims])


def test_mgkernels_reference_unit_test():
    """tests/testMGkernels.cc: restrict then extend constants ifunc*1.11 (1e-8)."""
    g = _g((32, 24, 20))
    gc = g.coarse()
    u = np.zeros((5, 32, 24, 20))
    for i in range(5):
        u[i] = i * 1.11
    blk = O.trade(g, O.embed(g, u))
    c = O.restrict(g, blk, gc)
    O.trade(gc, c)
    f2 = O.extend(gc, c, g)
    out = O.interior(g, f2)
    for i in range(5):
        assert np.max(np.abs(out[i] - i * 1.11)) < 1e-8


def _matrix(op, n_in, shape_in):
    cols = []
    for k in range(n_in):
        e = np.zeros(n_in)
        e[k] = 1.0
        cols.append(op(e.reshape(shape_in)).reshape(-1))
    return np.array(cols).T


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_restriction_is_scaled_adjoint_of_prolongation(dt):
    g = _g((8, 8, 8), ng=1)
    gc = g.coarse()

    def P(c):
        cb = O.trade(gc, O.embed(gc, c[None].astype(dt)))
        return O.interior(g, O.extend(gc, cb, g))[0]

    def R(f):
        fb = O.trade(g, O.embed(g, f[None].astype(dt)))
        return O.interior(gc, O.restrict(g, fb, gc))[0]

    Pm = _matrix(P, 64, (4, 4, 4))
    Rm = _matrix(R, 512, (8, 8, 8))
    assert np.array_equal(Rm, Pm.T / 8.0)
    assert np.all(np.sum(Pm, axis=1) == 1.0)


def test_restriction_symbol_on_mode():
    """Full weighting = tensor product of [1/4,1/2,1/4]: on cos(k.r) sampled at
    even fine points the result is prod_d (1+cos(k_d h))/2 * cos(k.R)."""
    g = _g((16, 12, 8), ng=2, ll=3.0)
    gc = g.coarse()
    h = list(g.h)
    m = (2, 1, 3)
    k = [2 * np.pi * m[i] / (g.dim[i] * h[i]) for i in range(3)]
    x = [np.arange(g.dim[i]) * h[i] for i in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    u = np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    c = O.interior(gc, O.restrict(g, O.trade(g, O.embed(g, u[None])), gc))[0]
    fac = np.prod([(1 + np.cos(k[i] * h[i])) / 2 for i in range(3)])
    assert np.max(np.abs(c - fac * u[::2, ::2, ::2])) < 1e-14


def test_prolongation_exact_on_linear_data():
    g = _g((8, 8, 8), ng=1)
    gc = g.coarse()
    ic = np.arange(4)
    CX, CY, CZ = np.meshgrid(ic, ic, ic, indexing="ij")
    lin = lambda x, y, z: 0.3 + 1.5 * x - 0.25 * y + 2.0 * z
    cb = O.trade(gc, O.embed(gc, lin(2 * CX, 2 * CY, 2 * CZ)[None].astype(float)))
    f = O.interior(g, O.extend(gc, cb, g))[0]
    fi = np.arange(8)
    FX, FY, FZ = np.meshgrid(fi, fi, fi, indexing="ij")
    # the last fine plane per direction interpolates towards the periodic image
    sl = (slice(0, 7),) * 3
    assert np.max(np.abs(f[sl] - lin(FX, FY, FZ)[sl])) < 1e-13


@pytest.mark.parametrize("kind", [1, 2])
def test_jacobi_amplification(kind):
    g = _g((16, 16, 16), ng=2, ll=4.0)
    h = list(g.h)
    k = [2 * np.pi * 3 / 4.0] * 3
    x = [np.arange(16) * h[i] for i in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    u = np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    f = O.embed(g, 0.4 * u[None])
    xb = O.embed(g, u[None])
    w = O.empty_block(g, 1)
    jf = O.jacobi_factor(kind, h)
    O.jacobi(g, kind, f, xb, w, jf)
    th = [k[i] * h[i] for i in range(3)]
    if kind == 2:
        lam = sum((30 - 32 * np.cos(t) + 2 * np.cos(2 * t)) / (12 * hh * hh) for t, hh in zip(th, h))
    else:
        lam = sum((2 - 2 * np.cos(t)) / (hh * hh) for t, hh in zip(th, h))
    exp = u - jf * (lam * u - 0.4 * u)
    assert np.max(np.abs(O.interior(g, xb)[0] - exp)) < 1e-12
    # omega / diag: 2/3 on Laph4, 6/7 on Laph2
    om = {2: 2.0 / 3.0, 1: 6.0 / 7.0}[kind]
    assert jf == pytest.approx(om / O.diag_el(kind, h), rel=1e-15)


def _mg_const(dt, masks=None, max_levels=2):
    g = _g((16, 16, 16), ng=2, ll=4.8)
    c = 0.75
    f = O.embed(g, np.full((2, 16, 16, 16), c), dt)
    gam = 0.01
    v = np.zeros_like(f)
    O.axpy(gam, f, v)
    O.mg(g, 2, max_levels, v, f, masks=masks)
    return g, O.interior(g, v), c, gam


def test_vcycle_constant_closed_form():
    """L(const) = 0, so every sweep adds jf*(f - 0); the stale residual handed
    to restriction is -f.  Result: (gamma + 4 jf0 + 4 jf1 + 4 jf2) * c for the
    3-level schedule 2 pre/2 post, 4 coarsest sweeps, correction v -= P e."""
    g, out, c, gam = _mg_const(np.float64)
    h0 = list(g.h)
    jf0 = O.jacobi_factor(2, h0)
    jf1 = O.jacobi_factor(1, [2 * x for x in h0])
    jf2 = O.jacobi_factor(1, [4 * x for x in h0])
    expect = (gam + 4 * jf0 + 4 * jf1 + 4 * jf2) * c
    assert np.max(np.abs(out - expect)) < 1e-15 * 64 * abs(expect)
    g1, out1, _, _ = _mg_const(np.float64, max_levels=1)
    jf1 = O.jacobi_factor(1, [2 * x for x in h0])
    expect1 = (gam + 4 * jf0 + 4 * jf1) * c
    assert np.max(np.abs(out1 - expect1)) < 1e-13 * abs(expect1)


def test_vcycle_masks_of_ones_equal_no_mask():
    g, a, _, _ = _mg_const(np.float64)
    ones = [np.ones((2, (16 >> l) ** 3), np.uint8) for l in range(3)]
    _, b, _, _ = _mg_const(np.float64, masks=ones)
    assert np.array_equal(a, b)


def test_vcycle_linear_and_float_close():
    rng = np.random.default_rng(7)
    g = _g((16, 16, 16), ng=2, ll=4.8)
    f1 = O.embed(g, rng.standard_normal((1, 16, 16, 16)))
    f2 = O.embed(g, rng.standard_normal((1, 16, 16, 16)))
    def run(f, dt=np.float64):
        f = f.astype(dt)
        v = np.zeros_like(f)
        O.axpy(0.02, f, v)
        return O.interior(g, O.mg(g, 2, 2, v, f)).astype(np.float64)
    a, b = 0.7, -1.3
    lhs = run(a * f1 + b * f2)
    rhs = a * run(f1) + b * run(f2)
    assert np.max(np.abs(lhs - rhs)) < 1e-12 * np.max(np.abs(lhs))
    lo = run(f1, np.float32)
    hi = run(f1)
    assert np.max(np.abs(lo - hi)) < 1e-5 * np.max(np.abs(hi))


def test_precond_constant_and_gamma():
    g = _g((16, 16, 16), ng=2, ll=4.8)
    h0 = list(g.h)
    gam = O.gamma(h0, 2, -0.1, -1.4)
    d = 1.0 / O.diag_el(2, h0) * 4 ** 3
    assert gam == pytest.approx(1.0 / (2.0 / d + 1.3), rel=1e-15)
    res = np.full((1, g.size), 0.5)
    out64 = O.precond(g, 2, 2, gam, res, dtype=np.float64)
    jf0 = O.jacobi_factor(2, h0)
    jf1 = O.jacobi_factor(1, [2 * x for x in h0])
    jf2 = O.jacobi_factor(1, [4 * x for x in h0])
    assert np.max(np.abs(out64 - 0.5 * (gam + 4 * (jf0 + jf1 + jf2)))) < 1e-13
    out32 = O.precond(g, 2, 2, gam, res, dtype=np.float32)
    assert np.max(np.abs(out32 - out64)) < 1e-6 * np.max(np.abs(out64))
