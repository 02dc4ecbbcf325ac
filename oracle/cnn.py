"""CPU oracle of the learned orbital preconditioner (test infrastructure).

Only tests, the smoke check and the bench's CPU baseline import this module.
The reference has no neural preconditioner; this restates the network the
measurement plan defines, so that the GPU forward, backward and training
kernels have a CPU checker:

  4 conv3d layers 1 -> 16 -> 16 -> 16 -> 1, 3^3 taps, zero padding ("same"
  on each box, the box's outside being zero as for the orbitals), bias,
  ReLU after layers 1-3, and a scalar skip:  y = s * x + conv4(h3).

Storage conventions shared with the engine:
  * activations: [sample][x][y][z][channel], channel fastest;
  * weights: w[tap][cout][cin], tap = (dx*3 + dy)*3 + dz for offsets
    (dx-1, dy-1, dz-1); bias b[cout]; the skip s is a scalar;
  * the GPU keeps FP32 master weights, computes with BF16 weights and
    BF16 activations, and accumulates in FP32: ``bf16=True`` rounds the
    same quantities here (round to nearest even).

Pins (tests/test_oracle_cnn.py): the vectorised convolution equals a direct
loop sum on tiny inputs; the backward pass equals central finite
differences of the loss in float64.
"""
import numpy as np

CH = 16
LAYERS = ((1, CH), (CH, CH), (CH, CH), (CH, 1))


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def taps():
    return [(dx - 1, dy - 1, dz - 1) for dx in range(3) for dy in range(3) for dz in range(3)]


def init_params(seed=0, skip=1.0):
    """He-normal weights (fan-in 27*cin), zero biases; float32."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, (ci, co) in enumerate(LAYERS):
        std = np.sqrt(2.0 / (27 * ci)) if k < 3 else np.sqrt(1.0 / (27 * ci)) * 0.1
        p["w%d" % k] = (rng.standard_normal((27, co, ci)) * std).astype(np.float32)
        p["b%d" % k] = np.zeros(co, dtype=np.float32)
    p["s"] = np.float32(skip)
    return p


def _shift(x, d):
    """out[v] = x[v + d] with zeros outside (x: [n, e, e, e, c])."""
    out = np.zeros_like(x)
    src = [slice(None)]
    dst = [slice(None)]
    for ax, o in enumerate(d):
        e = x.shape[ax + 1]
        if o >= 0:
            src.append(slice(o, e))
            dst.append(slice(0, e - o))
        else:
            src.append(slice(0, e + o))
            dst.append(slice(-o, e))
    src.append(slice(None))
    dst.append(slice(None))
    out[tuple(dst)] = x[tuple(src)]
    return out


def conv(x, w, b):
    """y[v, co] = b[co] + sum_tap sum_ci w[tap, co, ci] x[v + off(tap), ci]."""
    y = np.zeros(x.shape[:4] + (w.shape[1],), dtype=x.dtype)
    for t, d in enumerate(taps()):
        y += _shift(x, d) @ w[t].T.astype(x.dtype)
    return y + b.astype(x.dtype)


def conv_direct(x, w, b):
    """Loop restatement of conv() for tiny inputs (the pin)."""
    n, e0, e1, e2, ci = x.shape
    co = w.shape[1]
    y = np.zeros((n, e0, e1, e2, co))
    for s in range(n):
        for i in range(e0):
            for j in range(e1):
                for k in range(e2):
                    for o in range(co):
                        acc = float(b[o])
                        for t, (dx, dy, dz) in enumerate(taps()):
                            a, bb, c = i + dx, j + dy, k + dz
                            if 0 <= a < e0 and 0 <= bb < e1 and 0 <= c < e2:
                                acc += float(np.dot(w[t, o], x[s, a, bb, c]))
                        y[s, i, j, k, o] = acc
    return y


def forward(p, x, bf16_mode=False):
    """x: [n, e, e, e] -> (y [n, e, e, e], cache).  float64 unless bf16_mode,
    which rounds weights and the stored activations as the engine does."""
    dt = np.float32 if bf16_mode else np.float64
    q = (lambda a: bf16(a)) if bf16_mode else (lambda a: np.asarray(a, dtype=dt))
    h = [q(x)[..., None].astype(dt)]
    pre = []
    for k in range(4):
        z = conv(h[-1], q(p["w%d" % k]).astype(dt), p["b%d" % k].astype(dt))
        pre.append(z)
        if k < 3:
            h.append(q(np.maximum(z, 0.0)).astype(dt))
    y = dt(p["s"]) * np.asarray(x, dtype=dt) + pre[3][..., 0]
    return y, (h, pre)


def loss(y, t):
    """Mean squared error over all voxels."""
    d = y - t
    return float(np.mean(d * d))


def backward(p, x, t, bf16_mode=False):
    """Loss and gradients {name: array} of loss(forward(p, x), t)."""
    dt = np.float32 if bf16_mode else np.float64
    q = (lambda a: bf16(a)) if bf16_mode else (lambda a: np.asarray(a, dtype=dt))
    y, (h, pre) = forward(p, x, bf16_mode)
    L = loss(y, np.asarray(t, dtype=dt))
    dy = 2.0 * (y - t) / y.size
    g = {"s": dt(np.sum(dy * x))}
    dz = dy[..., None].astype(dt)
    for k in range(3, -1, -1):
        w = q(p["w%d" % k]).astype(dt)
        g["b%d" % k] = np.sum(dz, axis=(0, 1, 2, 3)).astype(dt)
        gw = np.zeros_like(w)
        for ti, d in enumerate(taps()):
            hs = _shift(h[k], d)
            gw[ti] = np.einsum("nxyzo,nxyzi->oi", dz, hs)
        g["w%d" % k] = gw
        if k == 0:
            break
        # dh[v, ci] = sum_tap sum_co w[tap, co, ci] dz[v - off(tap), co]
        dh = np.zeros(h[k].shape, dtype=dt)
        for ti, d in enumerate(taps()):
            dh += _shift(dz, tuple(-o for o in d)) @ w[ti]
        dz = dh * (pre[k - 1] > 0)
    return L, g
