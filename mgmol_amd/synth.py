"""Seeded synthetic inputs for the orbital-iteration workloads.

Conventions (stated once, used by tests, bench and smoke):
  * RNG: numpy PCG64 (``np.random.default_rng(seed)``).
  * Positions ~ U([0, L)^3), L = gdim*h, periodic minimum-image distances;
    optionally redrawn until at least ``min_sep`` bohr from every earlier
    centre (random sequential addition).
  * Orbital psi_j(r) = exp(-|r - c_j|^2 / (2 sigma^2)), cut at 6 sigma, plus
    U(-1e-3, 1e-3) noise on the points inside the cut, normalised so that
    h^3 sum psi^2 = 1.  (MGmol's own initial Gaussians are exp(-r^2/rc^2),
    src/LocGridOrbitals.cc:565; the 2 sigma^2 convention is used here.)
  * Wells V(r) = sum_k d_k exp(-|r - c_k|^2 / (2 w^2)), d_k ~ U(-2, -0.5) Ha,
    w = 1.5 bohr, cut at 6 w; the Hamiltonian uses V in Rydberg (x2), as
    MGmol does (src/Potentials.cc:101-102).
  * Boxes: cubic, edge drawn from ``edges``, origin = the centre's grid point
    minus edge/2, rounded down to a multiple of 2^(levels-1) so that every
    multigrid level of the box is aligned with the coarse grid.
"""
from dataclasses import dataclass, field

import numpy as np


def _mi(d, L):
    return d - L * np.round(d / L)


def _gauss1d(coords, c, L, s, cut):
    d = _mi(coords - c, L)
    g = np.exp(-d * d / (2.0 * s * s))
    g[np.abs(d) > cut] = 0.0
    return g, np.abs(d)


def wells(gdim, h, nwells, rng, width=1.5, depth=(-2.0, -0.5)):
    """Local potential on the full periodic grid, Rydberg, float64 (gx,gy,gz)."""
    gdim = tuple(int(x) for x in gdim)
    L = np.array(gdim) * h
    cen = rng.uniform(0.0, 1.0, (nwells, 3)) * L
    dep = rng.uniform(depth[0], depth[1], nwells)
    v = np.zeros(gdim)
    cut = 6.0 * width
    r = int(np.ceil(cut / h)) + 1
    for k in range(nwells):
        idx = []
        fac = []
        for d in range(3):
            ic = int(np.floor(cen[k, d] / h))
            ii = np.arange(ic - r, ic + r + 1)
            g, _ = _gauss1d(ii * h, cen[k, d], L[d], width, cut)
            idx.append(np.mod(ii, gdim[d]))
            fac.append(g)
        blk = np.einsum("i,j,k->ijk", *fac)
        # a window wider than the grid wraps onto itself: accumulate
        np.add.at(v, np.ix_(*idx), dep[k] * blk)
    return 2.0 * v, cen, dep


@dataclass
class BoxConfig:
    gdim: tuple
    h: float
    origin: np.ndarray      # (nbox, 3) int32
    edge: np.ndarray        # (nbox,) int32
    offset: np.ndarray      # (nbox+1,) int64, level 0
    psi: np.ndarray         # packed float64
    vtot: np.ndarray        # (gx, gy, gz) float64, Rydberg
    centers: np.ndarray
    levels: int
    meta: dict = field(default_factory=dict)

    @property
    def nbox(self):
        return len(self.edge)


def centers(n, L, rng, min_sep=None):
    """n positions ~ U([0, L)^3); with ``min_sep`` (bohr), candidates closer
    than min_sep (minimum image) to an accepted centre are redrawn (random
    sequential addition), so that no two orbitals start nearly identical."""
    L = np.asarray(L, dtype=np.float64)
    if not min_sep:
        return rng.uniform(0.0, 1.0, (n, 3)) * L
    out = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        c = rng.uniform(0.0, 1.0, 3) * L
        tries += 1
        if tries > 1000 * n:
            raise RuntimeError("cannot place %d centres %.2f bohr apart" % (n, min_sep))
        if k and np.min(np.sum(_mi(out[:k] - c, L) ** 2, axis=1)) < min_sep * min_sep:
            continue
        out[k] = c
        k += 1
    return out


def box_config(nbox, gdim, h, edges, nwells, seed, levels=3, sigma=1.0, noise=1e-3, min_sep=None):
    gdim = (int(gdim),) * 3 if np.isscalar(gdim) else tuple(int(x) for x in gdim)
    rng = np.random.default_rng(seed)
    L = np.array(gdim) * h
    cen = centers(nbox, L, rng, min_sep)
    edge = rng.choice(np.asarray(edges, dtype=np.int32), nbox).astype(np.int32)
    m = 1 << (levels - 1)
    origin = np.zeros((nbox, 3), dtype=np.int32)
    for d in range(3):
        start = np.floor(cen[:, d] / h).astype(np.int64) - edge // 2
        origin[:, d] = np.mod((start // m) * m, gdim[d])
    off = np.zeros(nbox + 1, dtype=np.int64)
    off[1:] = np.cumsum(edge.astype(np.int64) ** 3)
    psi = np.zeros(int(off[-1]))
    cut = 6.0 * sigma
    vel = h ** 3
    for j in range(nbox):
        e = int(edge[j])
        fac, dist = [], []
        for d in range(3):
            coords = (origin[j, d] + np.arange(e)) * h
            g, ad = _gauss1d(coords, cen[j, d], L[d], sigma, np.inf)
            fac.append(g)
            dist.append(ad)
        val = np.einsum("i,j,k->ijk", *fac)
        r2 = dist[0][:, None, None] ** 2 + dist[1][None, :, None] ** 2 + dist[2][None, None, :] ** 2
        inside = r2 <= cut * cut
        val = np.where(inside, val + rng.uniform(-noise, noise, val.shape), 0.0)
        val /= np.sqrt(vel * np.sum(val * val))
        psi[off[j]:off[j + 1]] = val.reshape(-1)
    v, wc, wd = wells(gdim, h, nwells, rng)
    return BoxConfig(gdim, h, origin, edge, off, psi, v, cen, levels,
                     {"seed": seed, "sigma": sigma, "nwells": nwells, "well_centers": wc, "well_depths": wd})


def grid_orbitals(n, gdim, h, seed, sigma=1.0, noise=1e-3, nwells=16):
    """Full-grid (C1-style) orbitals: (n, gx*gy*gz) float64 + potential."""
    gdim = (int(gdim),) * 3 if np.isscalar(gdim) else tuple(int(x) for x in gdim)
    rng = np.random.default_rng(seed)
    L = np.array(gdim) * h
    cen = rng.uniform(0.0, 1.0, (n, 3)) * L
    psi = np.zeros((n,) + gdim)
    cut = 6.0 * sigma
    for j in range(n):
        fac, dist = [], []
        for d in range(3):
            g, ad = _gauss1d(np.arange(gdim[d]) * h, cen[j, d], L[d], sigma, np.inf)
            fac.append(g)
            dist.append(ad)
        val = np.einsum("i,j,k->ijk", *fac)
        r2 = dist[0][:, None, None] ** 2 + dist[1][None, :, None] ** 2 + dist[2][None, None, :] ** 2
        inside = r2 <= cut * cut
        val = np.where(inside, val + rng.uniform(-noise, noise, val.shape), 0.0)
        val /= np.sqrt(h ** 3 * np.sum(val * val))
        psi[j] = val
    v, _, _ = wells(gdim, h, nwells, rng)
    return psi.reshape(n, -1), v


# The five workloads of the measurement plan.
CONFIGS = {
    "c1": dict(kind="grid", n=64, gdim=96, h=0.3, nwells=16, levels=3),
    # c2, c5: centres at least 2 bohr apart: uniform centres put
    # near-duplicate orbitals in the set (c5: overlap 0.98, cond(S) ~ 225 at
    # seed 1), with which the reference's update direction never settles
    # (residual stalls at 1e-3 .. 1e-2)
    "c2": dict(kind="box", nbox=512, gdim=192, h=0.25, edges=(24, 28, 32, 36, 40), nwells=16, levels=3,
               min_sep=2.0),
    "c3": dict(kind="dense", n=1024, gdim=256, h=0.2),
    "c5": dict(kind="box", nbox=2048, gdim=256, h=0.2, edges=(32,), nwells=64, levels=3, min_sep=2.0),
}
