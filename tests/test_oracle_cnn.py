"""Pins of the learned-preconditioner oracle: the vectorised convolution
against a direct loop sum, the backward pass against central finite
differences (float64), and the bfloat16 rounding rule."""
import numpy as np
import pytest

from oracle import cnn


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -7, 3.0e38], dtype=np.float32)
    r = cnn.bf16(v)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even (1 + 2^-6)
    assert r[2] == np.float32(1.0 + 2 ** -6)
    assert r[3] == -2.5 and r[4] == np.float32(1.0 + 2 ** -7)
    assert np.isfinite(r[5])
    u = r.view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 3, 5, 3))
    w = rng.standard_normal((27, 2, 3))
    b = rng.standard_normal(2)
    assert np.allclose(cnn.conv(x, w, b), cnn.conv_direct(x, w, b), rtol=1e-13, atol=1e-13)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = cnn.init_params(seed=3, skip=0.7)
    for k in range(4):
        p["b%d" % k] = (rng.standard_normal(p["b%d" % k].shape) * 0.1).astype(np.float32)
    p = {k: np.asarray(v, dtype=np.float64) for k, v in p.items()}
    x = rng.standard_normal((2, 4, 4, 4))
    t = rng.standard_normal((2, 4, 4, 4))
    L, g = cnn.backward(p, x, t)
    eps = 1e-6
    for name in ["s", "b0", "b3", "w0", "w1", "w2", "w3", "b1"]:
        flat = np.asarray(p[name]).reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            pp = {k: np.array(v, copy=True) for k, v in p.items()}
            pm = {k: np.array(v, copy=True) for k, v in p.items()}
            pp[name].reshape(-1)[idx] += eps
            pm[name].reshape(-1)[idx] -= eps
            num = (cnn.loss(cnn.forward(pp, x)[0], t) - cnn.loss(cnn.forward(pm, x)[0], t)) / (2 * eps)
            ana = np.asarray(g[name]).reshape(-1)[idx]
            assert ana == pytest.approx(num, rel=1e-5, abs=1e-9), (name, idx)
