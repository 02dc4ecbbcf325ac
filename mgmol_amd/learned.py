"""The learned orbital preconditioner: parameter handling and training loop
over the engine's CNN entry points (mgx_cnn_*).

The network: 4 conv3d layers 1 -> 16 -> 16 -> 16 -> 1, 3^3 taps, zero padding
per box, bias, ReLU after the first three, scalar skip y = s x + conv(h3).
It is trained on pairs (x, t) = (R_j, K R_j) / rms(R_j), where K is the
V-cycle preconditioner (OrbitalsPreconditioning::precond_mg,
src/OrbitalsPreconditioning.cc:87-117) and R_j a box residual produced by the
solver itself.  Parameters are FP32 on the host side and in the engine's
master copy; the engine computes with BF16 weights and activations.
"""
import ctypes as C

import numpy as np

from . import mgx

CH = 16
LAYERS = ((1, CH), (CH, CH), (CH, CH), (CH, 1))
NAMES = ["w0", "b0", "w1", "b1", "w2", "b2", "w3", "b3", "s"]


def shapes():
    out = {}
    for k, (ci, co) in enumerate(LAYERS):
        out["w%d" % k] = (27, co, ci)
        out["b%d" % k] = (co,)
    out["s"] = ()
    return out


def pack(p):
    """dict -> flat FP32 vector in the engine's order."""
    return np.concatenate([np.asarray(p[n], dtype=np.float32).reshape(-1) for n in NAMES])


def unpack(v):
    v = np.asarray(v, dtype=np.float32)
    out, o = {}, 0
    for n in NAMES:
        shp = shapes()[n]
        size = int(np.prod(shp)) if shp else 1
        out[n] = v[o:o + size].reshape(shp) if shp else np.float32(v[o])
        o += size
    assert o == v.size
    return out


def init_params(seed=0, skip=1.0):
    """He-normal weights (fan-in 27*cin; the last layer scaled by 0.1),
    zero biases, skip weight ``skip``."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, (ci, co) in enumerate(LAYERS):
        std = np.sqrt(2.0 / (27 * ci)) if k < 3 else np.sqrt(1.0 / (27 * ci)) * 0.1
        p["w%d" % k] = (rng.standard_normal((27, co, ci)) * std).astype(np.float32)
        p["b%d" % k] = np.zeros(co, dtype=np.float32)
    p["s"] = np.float32(skip)
    return p


class Cnn:
    """One network on the engine (mgx_cnn)."""

    def __init__(self, ctx, edge, max_batch, params=None):
        self.ctx, self.edge, self.max_batch = ctx, int(edge), int(max_batch)
        self.h = C.c_void_p()
        ctx.check(mgx.lib().mgx_cnn_create(ctx.h, self.edge, self.max_batch, C.byref(self.h)), "mgx_cnn_create")
        self.nparams = mgx.lib().mgx_cnn_nparams()
        self.set_params(params if params is not None else init_params())

    def set_params(self, p):
        v = pack(p) if isinstance(p, dict) else np.ascontiguousarray(p, dtype=np.float32)
        assert v.size == self.nparams
        self.ctx.check(mgx.lib().mgx_cnn_set_params(self.h, v.ctypes.data_as(C.c_void_p)), "mgx_cnn_set_params")

    def params(self):
        v = np.zeros(self.nparams, dtype=np.float32)
        self.ctx.check(mgx.lib().mgx_cnn_get_params(self.h, v.ctypes.data_as(C.c_void_p)), "mgx_cnn_get_params")
        return unpack(v)

    def forward(self, x, nb, y):
        """device FP32 x -> y (DeviceArray or raw pointers)."""
        self.ctx.check(mgx.lib().mgx_cnn_forward(self.h, mgx.ptr(x), nb, mgx.ptr(y)), "mgx_cnn_forward")

    def grad(self, x, t, nb):
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_double()
        self.ctx.check(mgx.lib().mgx_cnn_grad(self.h, mgx.ptr(x), mgx.ptr(t), nb, C.byref(loss),
                                              g.ctypes.data_as(C.c_void_p)), "mgx_cnn_grad")
        return loss.value, unpack(g)

    def train_step(self, x_ptr, t_ptr, nb, lr):
        """one Adam step on nb samples at raw device addresses; returns the loss."""
        loss = C.c_double()
        self.ctx.check(mgx.lib().mgx_cnn_train_step(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr), nb,
                                                    C.c_float(lr), C.byref(loss)), "mgx_cnn_train_step")
        return loss.value

    def apply(self, nbox, R, out):
        self.ctx.check(mgx.lib().mgx_cnn_apply(self.h, nbox, mgx.ptr(R), mgx.ptr(out)), "mgx_cnn_apply")

    def make_pairs(self, nbox, R_ptr, KR_ptr, x_ptr, t_ptr):
        self.ctx.check(mgx.lib().mgx_cnn_make_pairs(self.h, nbox, C.c_void_p(R_ptr), C.c_void_p(KR_ptr),
                                                    C.c_void_p(x_ptr), C.c_void_p(t_ptr)), "mgx_cnn_make_pairs")

    def close(self):
        if self.h:
            mgx.lib().mgx_cnn_destroy(self.h)
            self.h = C.c_void_p()


def train(cnn, x, t, nsamples, epochs, batch, lr, seed=0, log=None):
    """Adam over shuffled mini-batches of the nsamples pairs held in the
    device arrays x, t (FP32, nsamples * edge^3 each).  Returns the per-epoch
    mean losses."""
    rng = np.random.default_rng(seed)
    m = cnn.edge ** 3
    bytes_per = 4 * m
    hist = []
    nbat = max(1, nsamples // batch)
    for ep in range(epochs):
        # contiguous batches at a shuffled start order (samples are
        # shuffled once when the set is built)
        order = rng.permutation(nbat)
        tot = 0.0
        for i, b in enumerate(order):
            s0 = int(b) * batch
            nb = min(batch, nsamples - s0)
            tot += cnn.train_step(x.data_ptr() + s0 * bytes_per, t.data_ptr() + s0 * bytes_per, nb, lr)
            if log and (i + 1) % 500 == 0:
                log("epoch %d batch %d/%d running loss %.6e" % (ep, i + 1, nbat, tot / (i + 1)))
        hist.append(tot / nbat)
        if log:
            log("epoch %d loss %.6e" % (ep, hist[-1]))
    return hist


def build_dataset(ctx, seeds, box_kw, steps, capture_every, precond_dtype=None, log=None):
    """Training pairs generated by the plain solver itself: for each seeded
    box configuration (``synth.box_config(seed=s, **box_kw)``, one box
    edge), run the V-cycle solver and, every ``capture_every`` steps, turn
    each box's residual R_j and its V-cycle correction K R_j into the pair
    (R_j, K R_j) / rms(R_j).  Returns (x, t, nsamples, edge) with x, t FP32
    device arrays of nsamples * edge^3."""
    from . import synth
    precond_dtype = mgx.F32 if precond_dtype is None else precond_dtype
    edges = set(box_kw["edges"])
    if len(edges) != 1:
        raise ValueError("training boxes need one edge")
    edge = edges.pop()
    ncap = steps // capture_every
    nbox = box_kw["nbox"]
    nsamples = len(seeds) * ncap * nbox
    m = edge ** 3
    x = mgx.DeviceArray(ctx, shape=(nsamples * m,), dtype=np.float32, zero=False)
    t = mgx.DeviceArray(ctx, shape=(nsamples * m,), dtype=np.float32, zero=False)
    helper = Cnn(ctx, edge, 1)
    k = 0
    for seed in seeds:
        cfg = synth.box_config(seed=int(seed), **box_kw)
        bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, cfg.levels)
        sol = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=cfg.levels - 1, precond_dtype=precond_dtype)
        sol.set_keep_residual(True)  # the residuals are read between steps
        kr = mgx.DeviceArray(ctx, shape=(cfg.psi.size,), dtype=np.float64, zero=False)
        pres = C.c_void_p()
        for it in range(steps):
            sol.step()
            if (it + 1) % capture_every:
                continue
            ctx.check(mgx.lib().mgx_solver_buffers(sol.h, None, None, None, None, C.byref(pres)), "buffers")
            ctx.check(mgx.lib().mgx_box_precond_mg(ctx.h, bs.h, mgx.LAP_4TH, cfg.levels - 1, 0.0, precond_dtype,
                                                   pres, mgx.ptr(kr)), "mgx_box_precond_mg")
            off = k * nbox * m * 4
            helper.make_pairs(nbox, pres.value, kr.data_ptr(), x.data_ptr() + off, t.data_ptr() + off)
            k += 1
        if log:
            log("dataset: seed %d done (%d captures)" % (seed, k))
        kr.free()
        sol.close()
        bs.close()
    helper.close()
    return x, t, nsamples, edge
