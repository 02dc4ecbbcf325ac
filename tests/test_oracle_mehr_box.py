"""Pins of the oracle's Mehrstellen functions on per-orbital boxes against the
same operators applied on the full periodic grid to each embedded orbital
(CPU only).

* box-truncated H psi_j equals the full-grid (-Lap_M + B V) psi_j inside box j;
* the Galerkin S, H and B-matrix equal vel E_i . (op E_j) on the full grid
  (H is not symmetric, since B V is not);
* the residual (B Psi) theta - H Psi equals the full-grid one masked to each
  box;
* with V = 0 the Galerkin H is symmetric and B is symmetric positive definite.
"""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import synth


def _cfg(levels):
    if levels == 1:
        return synth.box_config(nbox=7, gdim=24, h=0.35, edges=(5, 7), nwells=3, seed=4, levels=1)
    return synth.box_config(nbox=8, gdim=24, h=0.35, edges=(8,), nwells=3, seed=5, levels=3)


def _full(b, cfg):
    g = O.Grid.make(tuple(cfg.gdim), 1, [cfg.h] * 3)
    E = O.b_embed(b, cfg.psi)
    HF = O.hloc(g, 0, O.embed(g, E), cfg.vtot).reshape(E.shape)
    BF = O.interior(g, O.rhs_mehr(g, O.trade(g, O.embed(g, E)), 1))
    return g, E, HF, BF


@pytest.mark.parametrize("levels", [1, 3])
def test_box_mehr_against_full_grid(levels):
    cfg = _cfg(levels)
    b = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    g, E, HF, BF = _full(b, cfg)
    n = b.nbox
    vel = cfg.h ** 3
    hp = O.b_hloc_mehr(b, cfg.vtot, cfg.psi)
    off = b.offsets(0)
    ref = O.b_extract(b, HF)
    assert np.allclose(hp, ref, rtol=1e-13, atol=1e-13 * np.max(np.abs(ref)))
    bp = O.b_bpsi_mehr(b, cfg.psi)
    assert np.allclose(bp, O.b_extract(b, BF), rtol=1e-13, atol=1e-15)
    S, H, Bm = O.b_gram_hb_mehr(b, cfg.psi, cfg.vtot)
    Ef, Hf, Bf = E.reshape(n, -1), HF.reshape(n, -1), BF.reshape(n, -1)
    for got, want in ((S, vel * Ef @ Ef.T), (H, vel * Ef @ Hf.T), (Bm, vel * Ef @ Bf.T)):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # the reach outside the boxes matters, and V makes H non-symmetric
    assert np.max(np.abs(H - H.T)) > 1e-8 * np.max(np.abs(H))
    rng = np.random.default_rng(1)
    th = rng.standard_normal((n, n))
    R = O.b_residual_mehr(b, cfg.psi, hp, th)
    m = O.b_masks(b)
    Rf = (th.T @ Bf - Hf) * m
    assert np.allclose(O.b_embed(b, R).reshape(n, -1), Rf, rtol=1e-12, atol=1e-12 * np.max(np.abs(Rf)))
    del off


def test_box_mehr_zero_potential_symmetric():
    cfg = _cfg(3)
    b = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    S, H, Bm = O.b_gram_hb_mehr(b, cfg.psi, np.zeros_like(cfg.vtot))
    assert np.max(np.abs(H - H.T)) <= 1e-13 * np.max(np.abs(H))
    assert np.max(np.abs(Bm - Bm.T)) <= 1e-14 * np.max(np.abs(Bm))
    assert np.min(np.linalg.eigvalsh(0.5 * (Bm + Bm.T))) > 0.0


def test_box_mehr_solver_step_lowers_residual():
    cfg = synth.box_config(nbox=6, gdim=32, h=0.3, edges=(8,), nwells=2, seed=6, levels=3)
    b = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    psi = cfg.psi.copy()
    out = []
    for _ in range(6):
        psi, e, r = O.b_solver_step_mehr(b, psi, cfg.vtot, 2, O.F64, 2.0)
        out.append((e, r))
    assert out[-1][1] < 0.5 * out[0][1]
    assert out[-1][0] < out[0][0]
