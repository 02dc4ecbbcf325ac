"""GPU parity of the learned preconditioner's kernels (forward, backward,
Adam training, box apply) against the CPU oracle in its BF16 mode, through
the C ABI."""
import numpy as np
import pytest

from mgmol_amd import learned, mgx
from oracle import cnn as OC
from tests.gpu_util import say, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def _params(seed, bias=0.05):
    p = learned.init_params(seed=seed, skip=0.3)
    rng = np.random.default_rng(seed + 100)
    for k in range(4):
        p["b%d" % k] = (rng.standard_normal(p["b%d" % k].shape) * bias).astype(np.float32)
    return p


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


@pytest.mark.parametrize("edge,nb", [(16, 2), (20, 3), (5, 4)])
def test_cnn_forward_matches_oracle(ctx, edge, nb):
    """Any transposed MFMA operand, wrong tap or channel order gives O(1)
    errors; BF16 rounding of activations bounds the rest."""
    p = _params(1)
    rng = np.random.default_rng(edge)
    x = rng.standard_normal((nb, edge, edge, edge)).astype(np.float32)
    net = learned.Cnn(ctx, edge, nb, p)
    tx, ty = to_dev(x), to_dev(np.zeros_like(x))
    net.forward(tx, nb, ty)
    y = to_np(ty)
    ref, _ = OC.forward(p, x, bf16_mode=True)
    r64, _ = OC.forward({k: np.asarray(v, dtype=np.float64) for k, v in p.items()}, x.astype(np.float64))
    e_bf, e_64 = _rel(y, ref), _rel(y, r64)
    say("cnn forward", edge, nb, "rel err vs bf16 oracle", e_bf, "vs fp64", e_64)
    assert e_bf < 5e-3
    assert e_64 < 3e-2
    net.close()


def test_cnn_grad_matches_oracle(ctx):
    p = _params(2)
    rng = np.random.default_rng(7)
    nb, e = 3, 12
    x = rng.standard_normal((nb, e, e, e)).astype(np.float32)
    t = (0.5 * x + 0.1 * rng.standard_normal(x.shape)).astype(np.float32)
    net = learned.Cnn(ctx, e, nb, p)
    loss, g = net.grad(to_dev(x), to_dev(t), nb)
    Lr, gr = OC.backward(p, x, t, bf16_mode=True)
    say("cnn loss", loss, Lr)
    assert loss == pytest.approx(Lr, rel=2e-3)
    for name in learned.NAMES:
        err = _rel(g[name], gr[name])
        say("grad", name, err)
        assert err < 3e-2, name
    net.close()


def test_cnn_training_reduces_loss(ctx):
    """Adam on a smoothing target (t = local 3^3 average of x): the loss falls
    well below the untrained network's."""
    rng = np.random.default_rng(3)
    nb, e = 8, 16
    x = rng.standard_normal((nb, e, e, e)).astype(np.float32)
    w = np.full((27, 1, 1), 1.0 / 27.0)
    t = OC.conv(x[..., None].astype(np.float64), w, np.zeros(1))[..., 0].astype(np.float32)
    net = learned.Cnn(ctx, e, nb, learned.init_params(seed=4, skip=0.0))
    tx, tt = to_dev(x), to_dev(t)
    first = net.train_step(tx.data_ptr(), tt.data_ptr(), nb, 3e-3)
    last = first
    for _ in range(150):
        last = net.train_step(tx.data_ptr(), tt.data_ptr(), nb, 3e-3)
    say("train loss", first, "->", last)
    assert last < 0.2 * first
    # the parameters moved and the reported loss is the one the oracle sees
    p = net.params()
    L_ref = OC.loss(OC.forward(p, x, bf16_mode=True)[0], t)
    l2, _ = net.grad(tx, tt, nb)
    assert l2 == pytest.approx(L_ref, rel=5e-3)
    net.close()


def test_cnn_apply_scales_per_box(ctx):
    p = _params(5)
    rng = np.random.default_rng(9)
    nbox, e = 5, 8
    scales = np.array([1e-6, 3.0, 1.0, 0.02, 50.0])
    R = rng.standard_normal((nbox, e, e, e)) * scales[:, None, None, None]
    net = learned.Cnn(ctx, e, 2, p)  # batches of 2 exercise the batching
    tR, tout = to_dev(R), to_dev(np.zeros_like(R))
    net.apply(nbox, tR, tout)
    out = to_np(tout)
    for j in range(nbox):
        rms = np.sqrt(np.mean(R[j] ** 2))
        ref = rms * OC.forward(p, (R[j] / rms)[None].astype(np.float32), bf16_mode=True)[0][0]
        assert _rel(out[j], ref) < 5e-3, j
    net.close()
