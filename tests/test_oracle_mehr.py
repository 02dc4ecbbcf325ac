"""Pins of the oracle's Mehrstellen projected matrices (mehr_bmatrix,
mehr_projected, mehr_residual) by closed forms on periodic Fourier modes:
for psi = cos / sin modes the B-matrix is diagonal with the B symbol
beta(k) = 1/2 + (1/6) sum_d cos(k_d h_d); with a constant potential v0 the
modes solve H psi = (lambda_M(k) + v0 beta(k)) B-weighted, so the residual
(B psi) theta - H psi vanishes and the eigenvalue sum is
tr(theta) = sum_k (lambda_M(k) + v0 beta(k)) / beta(k)."""
import numpy as np
import pytest

import oracle as O


def _modes(dims, h, ks):
    x = [np.arange(d) * h[i] for i, d in enumerate(dims)]
    L = [d * h[i] for i, d in enumerate(dims)]
    out = []
    for (m, kind) in ks:
        ph = sum(2 * np.pi * m[i] * x[i].reshape([-1 if j == i else 1 for j in range(3)]) / L[i] for i in range(3))
        out.append((np.cos(ph) if kind == "c" else np.sin(ph)).reshape(-1))
    return np.array(out)


def _beta(dims, h, m):
    return 0.5 + sum(np.cos(2 * np.pi * m[i] / dims[i]) for i in range(3)) / 6.0


DIMS, H = (16, 12, 20), (0.3, 0.35, 0.25)
KS = [((1, 0, 0), "c"), ((0, 2, 1), "s"), ((1, 1, 1), "c"), ((3, 0, 2), "s"), ((2, 3, 4), "c")]


def test_bmatrix_modes():
    g = O.Grid.make(DIMS, 1, H)
    psi = _modes(DIMS, H, KS)
    B = O.mehr_bmatrix(g, psi)
    vel = H[0] * H[1] * H[2]
    ref = np.diag([_beta(DIMS, H, m) * vel * np.sum(p * p) for (m, _), p in zip(KS, psi)])
    assert np.max(np.abs(B - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_residual_vanishes_and_eigsum():
    g = O.Grid.make(DIMS, 2, H)
    psi = _modes(DIMS, H, KS)
    v0 = -0.7
    vtot = np.full(int(np.prod(DIMS)), v0)
    hpsi = O.hloc(g, 0, O.embed(g, psi.reshape((len(KS),) + DIMS)), vtot)
    vel = H[0] * H[1] * H[2]
    S = vel * psi @ psi.T
    Hm = vel * psi @ hpsi.T
    B = O.mehr_bmatrix(g, psi)
    _, Sinv = O.chol_inv(S)
    invB, th, HB, es = O.mehr_projected(S, Sinv, B, Hm)
    res = O.mehr_residual(g, psi, hpsi, th)
    assert np.max(np.abs(res)) <= 1e-11 * np.max(np.abs(hpsi))
    # eigenvalue sum: tr(theta) = sum (lambda_M + v0 beta) / beta, lambda_M from H psi itself
    lam = np.array([(hp @ p) / (p @ p) for hp, p in zip(hpsi, psi)])
    beta = np.array([_beta(DIMS, H, m) for m, _ in KS])
    assert es == pytest.approx(float(np.sum(lam / beta)), rel=1e-12)
    assert np.allclose(HB, HB.T, rtol=0, atol=0)
