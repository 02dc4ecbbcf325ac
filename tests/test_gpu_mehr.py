"""GPU parity of the Mehrstellen projected-matrix entry points
(mgx_mehr_bmatrix, mgx_mehr_projected, mgx_mehr_residual) against the
oracle, through the C ABI; 1e-12 relative (FP64, sums in another order)."""
import ctypes

import numpy as np
import pytest

import oracle as O
from mgmol_amd import mgx
from tests.gpu_util import relmax, say, to_dev, to_np

pytestmark = pytest.mark.gpu

DIMS, H = (16, 12, 20), (0.3, 0.35, 0.25)


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def _psi(n, ld, seed):
    rng = np.random.default_rng(seed)
    numpt = int(np.prod(DIMS))
    psi = np.zeros((n, ld))
    psi[:, :numpt] = rng.standard_normal((n, numpt))
    return psi


@pytest.mark.parametrize("bc", [(1, 1, 1), (1, 0, 1)])
def test_bmatrix_and_residual(ctx, bc):
    n, numpt = 8, int(np.prod(DIMS))
    ld = numpt + 24
    psi = _psi(n, ld, 1)
    og = O.Grid.make(DIMS, 2, H)
    mg_ = mgx.Grid.make(DIMS, 2, H, bc)
    tp = to_dev(psi)
    tB = to_dev(np.zeros((n, n)))
    ctx.call("mgx_mehr_bmatrix", ctypes.byref(mg_), mgx.ptr(tp), ld, n, mgx.ptr(tB))
    ref = O.mehr_bmatrix(og, psi[:, :numpt], bc)
    # column-major B(i, j) = vel <psi_i, B psi_j>; with a Dirichlet direction
    # it is not symmetric (psi keeps its boundary plane, B psi reads it as 0)
    err = relmax(to_np(tB).T, ref)
    say("bmatrix bc", bc, "rel err", err)
    assert err <= 1e-12
    rng = np.random.default_rng(2)
    hp = np.zeros((n, ld))
    hp[:, :numpt] = rng.standard_normal((n, numpt))
    th = rng.standard_normal((n, n))
    tt = to_dev(np.asfortranarray(th).T.copy())
    tr = to_dev(np.zeros((n, ld)))
    ctx.call("mgx_mehr_residual", ctypes.byref(mg_), mgx.ptr(tp), mgx.ptr(to_dev(hp)), mgx.ptr(tt), ld, n,
             mgx.ptr(tr))
    rref = O.mehr_residual(og, psi[:, :numpt], hp[:, :numpt], th, bc)
    err = relmax(to_np(tr)[:, :numpt], rref)
    say("  residual rel err", err)
    assert err <= 1e-12


def test_projected(ctx):
    n = 24
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, 3 * n))
    S = a @ a.T / (3 * n)
    b = rng.standard_normal((n, 3 * n))
    B = b @ b.T / (3 * n) + 0.5 * np.eye(n)
    h = rng.standard_normal((n, n))
    Hm = h + h.T
    _, Sinv = O.chol_inv(S)
    invB, th, HB, es = O.mehr_projected(S, Sinv, B, Hm)
    dev = [to_dev(np.ascontiguousarray(m.T)) for m in (S, Sinv, B, Hm)]
    out = [to_dev(np.zeros((n, n))) for _ in range(3)]
    e = ctypes.c_double()
    ctx.call("mgx_mehr_projected", n, *[mgx.ptr(d) for d in dev], *[mgx.ptr(o) for o in out], ctypes.byref(e))
    # device matrices are column-major: transpose back
    got = [to_np(o).T for o in out]
    for name, g_, r_ in zip(("invB", "theta", "HB"), got, (invB, th, HB)):
        err = relmax(g_, r_)
        say(name, "rel err", err)
        assert err <= 1e-11
    assert np.array_equal(got[2], got[2].T)
    assert e.value == pytest.approx(es, rel=1e-11)
