"""mgx_local_gram_noalloc (the Gram entry point for callers that cannot
allocate device memory) against mgx_local_gram, through the C ABI.

* Two calls in a row, the second with a smaller input (the fixed scratch
  array keeps the first call's partials): bit-identical to mgx_local_gram
  (same kernels, same chunking), FP64 and hi/lo split.
* A size whose chunk partials exceed the scratch array at the usual
  chunking: fewer chunks, so the sums are ordered differently: within
  1e-13 of mgx_local_gram and of the oracle.
* Sizes beyond the limits (n >= 256, subdivx * n * n * 8 > 16 MiB) are
  rejected with MGX_EINVAL before any launch.
"""
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


def _both(ctx, prec, psi, ld, lnp, n, sub, vel):
    tp = to_dev(psi)
    ta, tb = to_dev(np.zeros((sub, n, n))), to_dev(np.full((sub, n, n), np.nan))
    ctx.call("mgx_local_gram", prec, mgx.ptr(tp), ld, lnp, n, sub, vel, mgx.ptr(ta))
    ctx.call("mgx_local_gram_noalloc", prec, mgx.ptr(tp), ld, lnp, n, sub, vel, mgx.ptr(tb))
    return to_np(ta), to_np(tb)


@pytest.mark.parametrize("prec", [mgx.PREC_F64, mgx.PREC_SPLIT])
def test_noalloc_gram_two_calls(ctx, prec):
    rng = np.random.default_rng(5)
    vel = 0.027
    for n, ld, lnp, sub in [(64, 20000, 10000, 2), (37, 3000, 1000, 3)]:
        psi = rng.standard_normal((n, ld))
        say("noalloc gram: prec %d n %d ld %d lnp %d subdivx %d" % (prec, n, ld, lnp, sub))
        a, b = _both(ctx, prec, psi, ld, lnp, n, sub, vel)
        assert np.array_equal(a, b)
        ref = O.local_gram(psi, lnp, sub, vel)
        assert relmax(b, ref) <= (1e-13 if prec == mgx.PREC_F64 else 1e-7)


def test_noalloc_gram_fewer_chunks(ctx):
    rng = np.random.default_rng(6)
    n, lnp, sub, vel = 255, 8192, 4, 0.027
    ld = lnp * sub
    psi = rng.standard_normal((n, ld))
    say("noalloc gram, chunks capped: n %d lnp %d subdivx %d" % (n, lnp, sub))
    a, b = _both(ctx, mgx.PREC_F64, psi, ld, lnp, n, sub, vel)
    assert relmax(b, a) <= 1e-13
    assert relmax(b, O.local_gram(psi, lnp, sub, vel)) <= 1e-13
    assert np.array_equal(b, np.transpose(b, (0, 2, 1)))


def test_noalloc_gram_rejects_beyond_limits(ctx):
    tp = to_dev(np.zeros((1, 8)))
    ts = to_dev(np.zeros(8))
    for n, sub in [(256, 1), (255, 33)]:
        rc = mgx.lib().mgx_local_gram_noalloc(ctx.h, mgx.PREC_F64, mgx.ptr(tp), 100000, 1000, n, sub, 1.0,
                                              mgx.ptr(ts))
        assert rc == mgx.MGX_EINVAL
        say("rejected n %d subdivx %d:" % (n, sub), mgx.lib().mgx_last_error(ctx.h).decode())
