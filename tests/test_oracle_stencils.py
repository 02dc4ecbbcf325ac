"""Pins for the oracle's finite-difference operators (CPU only).

Each check is independent of the oracle's coefficient tables: reference unit
tests restated (tests/testLaph4.cc, tests/testBatchLaph4.cc), discrete
Fourier symbols derived from the stencil shape, convergence order, operator
symmetry, and the diagonal read back from a delta function.
"""
import numpy as np
import pytest

import oracle as O


def _grid(dims=(32, 24, 20), ll=2.0, ng=2):
    h = [ll / d for d in dims]
    return O.Grid.make(dims, ng, h), h


def _coords(g, h, origin=0.0):
    x = [origin + (np.arange(g.dim[i]) + g.ghosts) * h[i] for i in range(3)]
    return np.meshgrid(*x, indexing="ij")


def test_laph4_sin_reference_unit_test():
    """tests/testLaph4.cc:7-85: -Lap of sum of sines, margin 2e-3."""
    g, h = _grid()
    X, Y, Z = _coords(g, h)
    cx = cy = cz = 2 * np.pi / 2.0
    u = np.sin(cx * X) + np.sin(cy * Y) + np.sin(cz * Z)
    blk = O.embed(g, u[None])
    O.trade(g, blk)
    out = O.interior(g, O.lap(g, 2, blk))[0]
    expect = cx * cx * np.sin(cx * X) + cy * cy * np.sin(cy * Y) + cz * cz * np.sin(cz * Z)
    assert np.max(np.abs(out - expect)) < 2e-3


def test_batch_laph4_reference_unit_test():
    """tests/testBatchLaph4.cc: same check for 10 copies in one call."""
    g, h = _grid()
    X, Y, Z = _coords(g, h)
    c = np.pi
    u = np.sin(c * X) + np.sin(c * Y) + np.sin(c * Z)
    blk = O.embed(g, np.repeat(u[None], 10, axis=0))
    O.trade(g, blk)
    out = O.interior(g, O.lap(g, 2, blk))
    expect = c * c * (np.sin(c * X) + np.sin(c * Y) + np.sin(c * Z))
    assert np.max(np.abs(out - expect[None])) < 2e-3
    assert np.array_equal(out[0], out[9])


def _mode(g, h, m):
    L = [g.dim[i] * h[i] for i in range(3)]
    k = [2 * np.pi * m[i] / L[i] for i in range(3)]
    X, Y, Z = _coords(g, h, origin=0.3)
    return np.cos(k[0] * X + k[1] * Y + k[2] * Z), [k[i] * h[i] for i in range(3)]


def _sym(kind, th, h):
    if kind == 2:  # 4th order: (30 - 32 cos t + 2 cos 2t) / (12 h^2)
        return sum((30 - 32 * np.cos(t) + 2 * np.cos(2 * t)) / (12 * hh * hh) for t, hh in zip(th, h))
    if kind == 1:  # 2nd order
        return sum((2 - 2 * np.cos(t)) / (hh * hh) for t, hh in zip(th, h))
    raise ValueError(kind)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("m", [(1, 0, 0), (2, 3, 1), (5, 4, 3)])
def test_symbol_exact(kind, m):
    """The discrete operator is diagonal on Fourier modes of the periodic grid
    with the symbol of its 1-D stencil; this pins neighbour offsets and the
    pairing of coefficients with directions."""
    g, h = _grid()
    u, th = _mode(g, h, m)
    blk = O.trade(g, O.embed(g, u[None]))
    out = O.interior(g, O.lap(g, kind, blk))[0]
    lam = _sym(kind, th, h)
    assert np.max(np.abs(out - lam * u)) <= 1e-11 * max(1.0, abs(lam))


@pytest.mark.parametrize("kind,order", [(1, 2), (2, 4)])
def test_convergence_order(kind, order):
    """|symbol - |k|^2| shrinks like h^order: independent of the coefficient table."""
    errs = []
    for n in (16, 32):
        g, h = _grid((n, n, n), ll=2 * np.pi)
        u, th = _mode(g, h, (1, 1, 1))
        blk = O.trade(g, O.embed(g, u[None]))
        out = O.interior(g, O.lap(g, kind, blk))[0]
        errs.append(np.max(np.abs(out - 3.0 * u)))
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.15


def test_mehrstellen_consistency_order():
    """Mehrstellen: L_M u = B(-Lap u) + O(h^4), with B u = u + h^2/12 Lap u + O(h^4).
    Checked on a mode as the h^4 decay of L_M - B*|k|^2 (a wrong coefficient
    in either operator breaks the 4th-order consistency)."""
    errs = []
    for n in (16, 32):
        g, h = _grid((n, n, n), ll=2 * np.pi)
        u, th = _mode(g, h, (1, 1, 1))
        blk = O.trade(g, O.embed(g, u[None]))
        lm = O.interior(g, O.lap(g, 0, blk))[0]
        b = O.interior(g, O.rhs_mehr(g, blk))[0]
        errs.append(np.max(np.abs(lm - 3.0 * b)))
    assert abs(np.log2(errs[0] / errs[1]) - 4) < 0.25


def test_rhs_mehr_constant_and_ghost_widths():
    g, h = _grid((8, 8, 8), ll=1.0, ng=2)
    blk = O.trade(g, O.embed(g, np.full((1, 8, 8, 8), 3.0)))
    for rg in (0, 1, 2):
        out = O.rhs_mehr(g, blk, rg)
        gi = O.Grid.make((8, 8, 8), rg, h)
        assert np.all(O.interior(gi, out) == 3.0)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_symmetric(kind):
    rng = np.random.default_rng(3)
    g, h = _grid((12, 10, 8), ll=1.7)
    a = O.trade(g, O.embed(g, rng.standard_normal((1, 12, 10, 8))))
    b = O.trade(g, O.embed(g, rng.standard_normal((1, 12, 10, 8))))
    la = O.interior(g, O.lap(g, kind, a))
    lb = O.interior(g, O.lap(g, kind, b))
    s1 = np.sum(O.interior(g, a) * lb)
    s2 = np.sum(la * O.interior(g, b))
    assert abs(s1 - s2) <= 1e-11 * abs(s1)


@pytest.mark.parametrize("kind,diag", [(1, 2.0), (2, 2.5), (0, 4.0 / 3.0)])
def test_diag_matches_stencil_centre(kind, diag):
    """Laph*::diagEl equals the stencil's own centre weight (read from a delta)."""
    g, h = _grid((8, 8, 8), ll=1.6)
    d = np.zeros((1, 8, 8, 8))
    d[0, 3, 4, 5] = 1.0
    out = O.interior(g, O.lap(g, kind, O.trade(g, O.embed(g, d))))[0]
    s = sum(1.0 / (x * x) for x in h)
    assert out[3, 4, 5] == pytest.approx(diag * s, rel=1e-15)
    assert O.diag_el(kind, h) == pytest.approx(out[3, 4, 5], rel=1e-15)


def test_withpot_and_float_path():
    rng = np.random.default_rng(5)
    g, h = _grid((10, 12, 8), ll=2.0)
    u = rng.standard_normal((3, 10, 12, 8))
    v = rng.standard_normal((10, 12, 8))
    blk = O.trade(g, O.embed(g, u))
    hp = O.lap4_withpot(g, blk, v)
    ref = O.interior(g, O.lap(g, 2, blk)).reshape(3, -1) + v.reshape(1, -1) * u.reshape(3, -1)
    assert np.max(np.abs(hp - ref)) < 1e-11 * np.max(np.abs(ref))
    blk32 = blk.astype(np.float32)
    hp32 = O.lap4_withpot(g, blk32, v)
    assert hp32.dtype == np.float32
    assert np.max(np.abs(hp32 - hp)) < 1e-5 * np.max(np.abs(hp))


def test_trade_periodic_and_dirichlet():
    g, h = _grid((6, 5, 4), ll=1.0, ng=2)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((2, 6, 5, 4))
    blk = O.trade(g, O.embed(g, u))
    full = np.pad(u, ((0, 0), (2, 2), (2, 2), (2, 2)), mode="wrap")
    assert np.array_equal(blk, full)
    # Dirichlet y (GridFunc::setBoundaryValues, src/pb/GridFunc.cc:2189-2340):
    # the ghosts and the boundary plane y = 0 hold zero
    blk2 = O.trade(g, O.embed(g, u), bc=(1, 0, 1))
    uz = u.copy()
    uz[:, :, 0, :] = 0.0
    exp = np.pad(uz, ((0, 0), (2, 2), (0, 0), (2, 2)), mode="wrap")
    exp = np.pad(exp, ((0, 0), (0, 0), (2, 2), (0, 0)))
    assert np.array_equal(blk2, exp)


def test_hloc_mehrstellen_constant_potential():
    """Mehrstellen H psi = L_M psi + B(V psi): for constant V the symbol is
    sym(L_M) + v0 * sym(B) on every mode."""
    g, h = _grid((16, 16, 16), ll=3.0)
    u, th = _mode(g, h, (1, 2, 3))
    v0 = -0.7
    blk = O.embed(g, u[None])
    hp = O.hloc(g, 0, blk, np.full(g.size, v0))[0].reshape(u.shape)
    t = O.trade(g, blk.copy())
    ref = O.interior(g, O.lap(g, 0, t))[0] + v0 * O.interior(g, O.rhs_mehr(g, t))[0]
    assert np.max(np.abs(hp - ref)) < 1e-12 * np.max(np.abs(ref))
    symB = 0.5 + sum(np.cos(x) for x in th) / 6.0
    assert np.max(np.abs(O.interior(g, O.rhs_mehr(g, t))[0] - symB * u)) < 1e-13
