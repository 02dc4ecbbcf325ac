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


# ---------------------------------------------------------------------------
# Two-level V-cycles on one Fourier mode, in closed form.
#
# On a periodic grid every stencil is diagonal in Fourier space, full
# weighting maps the fine mode cos(k.r) to the coarse mode cos(k.r_c) times
# W_0 = prod_d (1 + cos k_d h) / 2, and trilinear prolongation maps the coarse
# mode to the 8 fine modes cos((k + pi/h (.) a).r), a in {0,1}^3, with weights
# W_a = prod_d (1 + (-1)^a_d cos k_d h) / 2 (the tensor product of the 1-D
# prolongation symbol).  A two-level cycle on f = cos(k.r) from x = gamma f is
# then: Jacobi sweeps c <- c - jf (lam c - 1) on the mode; the restricted
# residual r_c = W_0 w, with w = lam c - 1 taken either before the last
# pre-sweep's update (the stale residual the last `jacobi` call leaves in w,
# Preconditioning::mg, src/Preconditioning.cc:170-186) or after it (the fresh
# residual A x - rhs of pb::Vcycle, src/pb/Vcycle.h:28-249); four coarse
# sweeps from zero with the coarse operator's symbol; x -= P e, which puts
# -W_a v_c on mode a; and the post-sweeps on every mode.  A dropped sweep, a
# wrong restriction or prolongation weight, a sign of the correction, or a
# fresh residual where a stale one belongs each change the predicted output by
# far more than the 1e-12 tolerance (checked below for the residual).
# ---------------------------------------------------------------------------
def _sym4(k, h):
    return sum((30 - 32 * np.cos(kk * hh) + 2 * np.cos(2 * kk * hh)) / (12 * hh * hh) for kk, hh in zip(k, h))


def _sym2(k, h):
    return sum((2 - 2 * np.cos(kk * hh)) / (hh * hh) for kk, hh in zip(k, h))


def _two_level_modes(k, h, gam, stale, nu1=2, nu2=2, ncoarse=4):
    """[(coefficient, wave vector)] of the cycle's output on f = cos(k.r)."""
    hc = [2 * x for x in h]
    jf = O.jacobi_factor(2, h)
    jfc = O.jacobi_factor(1, hc)
    lam0 = _sym4(k, h)
    c, w = gam, None
    for _ in range(nu1):
        w = lam0 * c - 1.0
        c = c - jf * w
    if not stale:
        w = lam0 * c - 1.0
    w0 = np.prod([(1 + np.cos(kk * hh)) / 2 for kk, hh in zip(k, h)])
    rc = w0 * w
    lamc = _sym2(k, hc)
    vc = 0.0
    for _ in range(ncoarse):
        vc = vc - jfc * (lamc * vc - rc)
    out = []
    for a in np.ndindex(2, 2, 2):
        wa = np.prod([(1 + (-1) ** ad * np.cos(kk * hh)) / 2 for ad, kk, hh in zip(a, k, h)])
        ka = [kk + np.pi / hh * ad for ad, kk, hh in zip(a, k, h)]
        lam = _sym4(ka, h)
        ca = (c if a == (0, 0, 0) else 0.0) - wa * vc
        for _ in range(nu2):
            ca = ca - jf * (lam * ca - (1.0 if a == (0, 0, 0) else 0.0))
        out.append((ca, ka))
    return out


def _eval_modes(modes, dims, h):
    x = [np.arange(dims[i]) * h[i] for i in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    return sum(c * np.cos(ka[0] * X + ka[1] * Y + ka[2] * Z) for c, ka in modes)


@pytest.mark.parametrize("m", [(1, 2, 0), (1, 1, 1), (3, 0, 2)])
@pytest.mark.parametrize("gam", [0.0, 0.013])
def test_vcycle_two_level_fourier_mode_stale_residual(m, gam):
    dims = (16, 16, 16)
    g = _g(dims, ng=2, ll=4.8)
    h = list(g.h)
    k = [2 * np.pi * mm / (dims[i] * h[i]) for i, mm in enumerate(m)]
    x = [np.arange(dims[i]) * h[i] for i in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    u = np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    f = O.embed(g, u[None])
    v = np.zeros_like(f)
    O.axpy(gam, f, v)
    O.mg(g, 2, 1, v, f)
    out = O.interior(g, v)[0]
    stale = _eval_modes(_two_level_modes(k, h, gam, stale=True), dims, h)
    fresh = _eval_modes(_two_level_modes(k, h, gam, stale=False), dims, h)
    scale = np.max(np.abs(stale))
    assert np.max(np.abs(out - stale)) < 1e-12 * scale
    # the closed form tells the two schedules apart by far more than its tolerance
    assert np.max(np.abs(fresh - stale)) > 1e-4 * scale


@pytest.mark.parametrize("m", [(1, 2, 0), (2, 1, 1)])
def test_poisson_vcycle_two_level_fourier_mode_fresh_residual(m):
    """pb::Vcycle with one coarsening (max_nlevels = 1): nu1 = nu2 = 2 fine
    sweeps around the fresh residual's coarse correction, nu1 + nu2 = 4
    sweeps from zero on the coarsest level (its operator Laph2)."""
    dims = (16, 16, 16)
    g = _g(dims, ng=2, ll=4.8)
    h = list(g.h)
    k = [2 * np.pi * mm / (dims[i] * h[i]) for i, mm in enumerate(m)]
    x = [np.arange(dims[i]) * h[i] for i in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    u = np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    rhs = O.embed(g, u[None])
    O.trade(g, rhs)
    w = np.zeros_like(rhs)
    O._vcycle(g, 2, w, rhs, 0, 1, 2, 2)
    out = O.interior(g, w)[0]
    fresh = _eval_modes(_two_level_modes(k, h, 0.0, stale=False), dims, h)
    stale = _eval_modes(_two_level_modes(k, h, 0.0, stale=True), dims, h)
    scale = np.max(np.abs(fresh))
    assert np.max(np.abs(out - fresh)) < 1e-12 * scale
    assert np.max(np.abs(fresh - stale)) > 1e-4 * scale
