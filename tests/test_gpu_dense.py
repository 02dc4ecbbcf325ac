"""GPU parity of the dense layer: MFMA Gram / local products (FP64 and
hi/lo split), slab accumulation, Cholesky / inverse, theta, trace, Ritz,
residual and its norm -- against the oracle / numpy, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import mgx
from tests.gpu_util import relmax, say, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def test_mfma_single_tile_layout(ctx):
    """One 16x16 output from 4 rows: checks the MFMA operand / result lane
    layout (any transposition or row mix-up fails)."""
    rng = np.random.default_rng(0)
    psi = rng.standard_normal((16, 4))
    tp, ts = to_dev(psi), to_dev(np.zeros((1, 16, 16)))
    ctx.call("mgx_local_product", mgx.PREC_F64, mgx.ptr(tp), mgx.ptr(tp), 4, 4, 16, 1, 1.0, mgx.ptr(ts))
    got = to_np(ts)[0]
    # column-major: got[j, i] = <psi_i, psi_j>
    assert np.allclose(got.T, psi @ psi.T, rtol=1e-14, atol=1e-14)
    b = rng.standard_normal((16, 4))
    tb = to_dev(b)
    ctx.call("mgx_local_product", mgx.PREC_F64, mgx.ptr(tp), mgx.ptr(tb), 4, 4, 16, 1, 1.0, mgx.ptr(ts))
    got = to_np(ts)[0]
    assert np.allclose(got.T, psi @ b.T, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n,ld,lnp,sub", [(5, 1000, 500, 2), (64, 4096, 4096, 1), (37, 3000, 1000, 3)])
def test_local_gram_and_product_f64(ctx, n, ld, lnp, sub):
    rng = np.random.default_rng(n)
    psi = rng.standard_normal((n, ld))
    hp = rng.standard_normal((n, ld))
    vel = 0.027
    tp, th = to_dev(psi), to_dev(hp)
    ts = to_dev(np.zeros((sub, n, n)))
    ctx.call("mgx_local_gram", mgx.PREC_F64, mgx.ptr(tp), ld, lnp, n, sub, vel, mgx.ptr(ts))
    ref = O.local_gram(psi, lnp, sub, vel)
    g = to_np(ts)
    assert relmax(g, ref) <= 1e-13
    assert np.array_equal(g, np.transpose(g, (0, 2, 1)))
    ctx.call("mgx_local_product", mgx.PREC_F64, mgx.ptr(tp), mgx.ptr(th), ld, lnp, n, sub, vel, mgx.ptr(ts))
    assert relmax(to_np(ts), O.local_product(psi, hp, lnp, sub, vel)) <= 1e-13


def test_local_gram_split_error(ctx):
    """hi/lo split on FP32 MFMA: error vs FP64 stated and bounded."""
    rng = np.random.default_rng(11)
    n, ld = 64, 200000
    psi = rng.standard_normal((n, ld)) * 1e-2
    tp = to_dev(psi)
    ts = to_dev(np.zeros((1, n, n)))
    ctx.call("mgx_local_gram", mgx.PREC_SPLIT, mgx.ptr(tp), ld, ld, n, 1, 1.0, mgx.ptr(ts))
    ref = O.local_gram(psi, ld, 1, 1.0)
    err = relmax(to_np(ts), ref)
    say("split gram rel err (max-norm)", err)
    # FP32 accumulation over 64-row chunks bounds the error (measured 5.6e-9)
    assert err < 2e-8


def test_wide_products(ctx):
    """n >= 256: the library dgemm in FP64 (K-batches summed in FP64); the
    split on the hand-written wide kernel (hi / lo formed in registers, FP32
    MFMA sums over 64-row chunks, FP64 over the chunks).  Two slabs of 35,000
    rows each with an odd leading dimension (unaligned columns load one row
    at a time); the last chunk of a split is partial."""
    rng = np.random.default_rng(21)
    n, lnp, sub = 256, 35000, 2
    ld = sub * lnp + 17
    psi = rng.standard_normal((n, ld)) * 1e-2
    hp = psi + 1e-3 * rng.standard_normal((n, ld))
    vel = 0.008
    tp, th = to_dev(psi), to_dev(hp)
    ts = to_dev(np.zeros((sub, n, n)))
    gref = O.local_gram(psi, lnp, sub, vel)
    pref = O.local_product(psi, hp, lnp, sub, vel)
    ctx.call("mgx_local_gram", mgx.PREC_F64, mgx.ptr(tp), ld, lnp, n, sub, vel, mgx.ptr(ts))
    g = to_np(ts)
    assert relmax(g, gref) <= 1e-13
    assert np.array_equal(g, np.transpose(g, (0, 2, 1)))
    ctx.call("mgx_local_product", mgx.PREC_F64, mgx.ptr(tp), mgx.ptr(th), ld, lnp, n, sub, vel, mgx.ptr(ts))
    assert relmax(to_np(ts), pref) <= 1e-13
    ctx.call("mgx_local_gram", mgx.PREC_SPLIT, mgx.ptr(tp), ld, lnp, n, sub, vel, mgx.ptr(ts))
    g = to_np(ts)
    eg = relmax(g, gref)
    assert np.array_equal(g, np.transpose(g, (0, 2, 1)))
    ctx.call("mgx_local_product", mgx.PREC_SPLIT, mgx.ptr(tp), mgx.ptr(th), ld, lnp, n, sub, vel, mgx.ptr(ts))
    ep = relmax(to_np(ts), pref)
    say("wide split rel err (max-norm): gram %.3e product %.3e" % (eg, ep))
    # hi*hi of every 4 rows (one MFMA step) added in FP64, the cross terms
    # (2^-24 of it) in FP32 over 64 rows: measured 1.4e-9 / 1.9e-9
    assert eg < 1e-8 and ep < 1e-8


def test_local_to_global(ctx):
    rng = np.random.default_rng(12)
    sub, m, numst = 3, 4, 7
    ss = rng.standard_normal((sub, m, m))
    ss[0, 1, 2] = 1e-16
    gids = np.array([[0, 3, -1, 5], [1, 3, 6, -1], [0, 2, 4, 5]], dtype=np.int32)
    tss, tg = to_dev(ss), to_dev(gids)
    tglob = to_dev(np.zeros((numst, numst)))
    ctx.call("mgx_local_to_global", mgx.ptr(tss), mgx.ptr(tg), sub, m, numst, 1e-14, mgx.ptr(tglob))
    # oracle returns [row, col] with column-major storage mapped to numpy order
    ref = O.local_to_global(ss, gids, numst)
    assert relmax(to_np(tglob), ref) <= 1e-15


def test_chol_inv_theta_eigsum_ritz(ctx):
    """tests/testGramMatrix.cc identities plus tr(S^-1 H) = sum of Ritz values."""
    n = 213
    rng = np.random.default_rng(13)
    A = rng.uniform(-1.0 / n, 1.0 / n, (n, n))
    B = 0.5 * (A + A.T)
    np.fill_diagonal(B, 2.0)
    H = rng.standard_normal((n, n))
    H = H + H.T
    tS, tL, tI = to_dev(B), to_dev(np.zeros_like(B)), to_dev(np.zeros_like(B))
    ctx.call("mgx_chol_inv", n, mgx.ptr(tS), mgx.ptr(tL), mgx.ptr(tI))
    L = to_np(tL).T  # column-major -> numpy row-major lower
    assert np.max(np.abs(L @ L.T - B)) < 1e-14
    assert np.allclose(np.triu(L, 1), 0.0)
    Sinv = to_np(tI).T
    assert np.max(np.abs(Sinv @ B - np.eye(n))) < 1e-13
    tH, tT = to_dev(np.asfortranarray(H).T.copy()), to_dev(np.zeros_like(H))
    ctx.call("mgx_theta", n, mgx.ptr(tI), mgx.ptr(tH), mgx.ptr(tT))
    th = to_np(tT).T
    assert relmax(th, Sinv @ H) <= 1e-13
    import ctypes
    tr = ctypes.c_double()
    ctx.call("mgx_eigsum", n, mgx.ptr(tH), mgx.ptr(tI), ctypes.byref(tr))
    w, _ = O.ritz(H, B)
    assert tr.value == pytest.approx(np.sum(w), rel=1e-12)
    te = to_dev(np.zeros(n))
    ctx.call("mgx_ritz", n, mgx.ptr(tH), mgx.ptr(tS), mgx.ptr(te), None)
    assert relmax(to_np(te), w) <= 1e-12


def test_chol_not_spd(ctx):
    n = 8
    S = -np.eye(n)
    tS, tL, tI = to_dev(S), to_dev(np.zeros_like(S)), to_dev(np.zeros_like(S))
    rc = mgx.lib().mgx_chol_inv(ctx.h, n, mgx.ptr(tS), mgx.ptr(tL), mgx.ptr(tI))
    assert rc == mgx.MGX_ENOTSPD


def test_residual_and_norm(ctx):
    rng = np.random.default_rng(14)
    n, ld = 6, 4000
    psi = rng.standard_normal((n, ld))
    hp = rng.standard_normal((n, ld))
    th = rng.standard_normal((n, n))
    tp, th_d, tt = to_dev(psi), to_dev(hp), to_dev(np.asfortranarray(th).T.copy())
    tr = to_dev(np.zeros_like(psi))
    ctx.call("mgx_residual", mgx.ptr(tp), mgx.ptr(th_d), mgx.ptr(tt), ld, ld, n, mgx.ptr(tr))
    ref = O.residual(psi, hp, th)
    assert relmax(to_np(tr), ref) <= 1e-13
    Sinv = np.diag(rng.uniform(0.5, 1.5, n))
    tI = to_dev(Sinv)
    import ctypes
    out = ctypes.c_double()
    ctx.call("mgx_residual_norm", mgx.ptr(tr), ld, ld, n, 0.01, mgx.ptr(tI), ctypes.byref(out))
    d = 0.01 * np.sum(ref * ref, axis=1)
    assert out.value == pytest.approx(O.residual_norm(d, Sinv), rel=1e-13)


def test_synth_separable_against_oracle_operator(ctx):
    """The dense workload's device generator: psi_j = fx fy fz on z-planes
    [z0, z0 + nzs) and H psi_j from the 1D second differences equal the
    oracle's (-Delta_h4 + V) psi_j on the full periodic grid (1e-12)."""
    from mgmol_amd.measure import _separable_factors
    import ctypes as C
    n, gd, h, z0, nzs = 3, 16, 0.3, 5, 6
    f, lf, rng = _separable_factors(n, gd, h, seed=7)
    v = rng.standard_normal((gd, gd, gd))
    full = np.einsum("ji,jk,jl->jikl", f[0], f[1], f[2])  # (n, x, y, z)
    g = O.Grid.make((gd, gd, gd), 2, (h, h, h))
    HF = O.lap4_withpot(g, O.trade(g, O.embed(g, full)), v).reshape(n, gd, gd, gd)
    dev = [to_dev(np.ascontiguousarray(a)) for a in f + lf]
    K = gd * gd * nzs
    tp, th = to_dev(np.zeros((n, K))), to_dev(np.zeros((n, K)))
    gdim = (C.c_int * 3)(gd, gd, gd)
    ctx.check(mgx.lib().mgx_synth_separable(ctx.h, n, gdim, z0, nzs, *[mgx.ptr(d) for d in dev], mgx.ptr(to_dev(v)),
                                            mgx.ptr(tp), mgx.ptr(th)), "mgx_synth_separable")
    ps = full[:, :, :, z0:z0 + nzs].reshape(n, K)
    hs = HF[:, :, :, z0:z0 + nzs].reshape(n, K)
    assert relmax(to_np(tp), ps) <= 1e-14
    assert relmax(to_np(th), hs) <= 1e-12
    # normalised over the full grid
    assert np.allclose(h ** 3 * np.sum(full * full, axis=(1, 2, 3)), 1.0, rtol=1e-12)
