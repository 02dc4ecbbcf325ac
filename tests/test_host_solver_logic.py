"""CPU checks of the solver's host-side rules: the Anderson coefficient
bounds of AndersonMix::update, the energy-change bookkeeping of
DFTsolver::checkConvergenceEnergy, and the seeded centre generator."""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import synth


@pytest.mark.parametrize("mat,rhs,expect", [
    (2.0, 0.5, 0.25),      # inside [-3, 0.5]: rhs/mat
    (1.0, 0.5, 0.5),       # on the upper bound
    (1.0, 0.75, 0.0),      # 0.5 < theta <= 1: simple steepest descent
    (1.0, 1.5, -0.5),      # theta > 1: reset to -0.5
    (1.0, -5.0, -3.0),     # below min_theta: clamped
    (0.0, 1.0, 0.0),       # degenerate history
])
def test_anderson_theta_bounds(mat, rhs, expect):
    assert O._anderson_theta(mat, rhs) == expect


def test_anderson_single_history_is_secant_minimiser():
    """With one history vector, theta minimises |(1-t) f + t f0| in the
    inner product: theta = <f - f0, f> / |f - f0|^2 (checked by brute force)."""
    rng = np.random.default_rng(0)
    f, f0 = rng.standard_normal(50), rng.standard_normal(50)
    df = f - f0
    th = O._anderson_theta(float(df @ df), float(df @ f))
    ts = np.linspace(-3, 0.5, 35001)
    best = ts[np.argmin([np.sum(((1 - t) * f + t * f0) ** 2) for t in ts])]
    if -3 <= float(df @ f) / float(df @ df) <= 0.5:
        assert th == pytest.approx(best, abs=2e-4)


def test_anderson_update_is_a_step_from_the_extrapolated_point():
    """For a linear residual map f(x) = alpha (b - A x) the combination of
    AndersonMix::update is one preconditioned step taken from the
    extrapolated point: xb + beta f(xb), with xb = (1 - t) x + t x0 (fb =
    f(xb) by linearity); and the secant theta makes |f(xb)| no larger than
    |f(x)| (t = 0 is a candidate of the minimisation)."""
    rng = np.random.default_rng(3)
    n = 40
    M = rng.standard_normal((n, n))
    A = M @ M.T / n + np.eye(n)
    b = rng.standard_normal(n)
    alpha = 0.2

    def fmap(x):
        return alpha * (b - A @ x)
    x0 = rng.standard_normal(n)
    x = x0 + fmap(x0)
    f, f0 = fmap(x), fmap(x0)
    df = f - f0
    t = O._anderson_theta(float(df @ df), float(df @ f))
    for beta in (1.0, 0.5):
        nxt = O._anderson_update(x, x0, f, f0, t, beta)
        xb = (1 - t) * x + t * x0
        assert np.allclose(nxt, xb + beta * fmap(xb), rtol=1e-12, atol=1e-12)
    assert np.linalg.norm(fmap((1 - t) * x + t * x0)) <= np.linalg.norm(f) * (1 + 1e-12)
    # without history (x0 = x, f0 = f) the update is the plain step x + beta f
    assert np.array_equal(O._anderson_update(x, x, f, f, 0.0, 1.0), x + f)


def test_centres_min_separation():
    rng = np.random.default_rng(5)
    L = np.array([20.0, 20.0, 20.0])
    c = synth.centers(200, L, rng, min_sep=2.0)
    d = c[:, None, :] - c[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt(np.sum(d * d, axis=2)) + np.eye(200) * 1e9
    assert r.min() >= 2.0
    assert np.all((c >= 0) & (c < L))
    # without a separation the draw is the plain uniform one
    a = synth.centers(10, L, np.random.default_rng(1))
    b = np.random.default_rng(1).uniform(0.0, 1.0, (10, 3)) * L
    assert np.array_equal(a, b)
