"""Train the learned orbital preconditioner on the GPU.

    python -m mgmol_amd.train_cnn [--out mgmol_amd/weights/cnn_e32.npz] ...

Training pairs come from the plain V-cycle solver itself: seeded box
configurations of 32^3 boxes at the density of the 2048-orbital workload
(256 boxes on a 128^3 grid, h = 0.2 bohr, 8 wells, centres at least 2 bohr
apart), run for ``--steps`` orbital steps with every box's (R, K R) / rms(R)
captured every ``--every`` steps (8 seeds x 100 captures x 256 boxes =
204,800 pairs by default).  The network is trained with Adam over shuffled
mini-batches and the weights saved as FP32 arrays, with the training log and
the relative error on held-out pairs from another seed.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import learned, mgx

BOX_KW = dict(nbox=256, gdim=128, h=0.2, edges=(32,), nwells=8, levels=3, min_sep=2.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_error(cnn, x, t, nsamples, batch):
    """||net(x) - t|| / ||t|| over the held-out pairs."""
    m = cnn.edge ** 3
    y = mgx.DeviceArray(cnn.ctx, shape=(batch * m,), dtype=np.float32, zero=False)
    num = den = 0.0
    ta = t.get()
    for s0 in range(0, nsamples, batch):
        nb = min(batch, nsamples - s0)
        cnn.ctx.check(mgx.lib().mgx_cnn_forward(cnn.h, C.c_void_p(x.data_ptr() + 4 * s0 * m), nb,
                                                C.c_void_p(y.data_ptr())), "mgx_cnn_forward")
        yv = y.get()[:nb * m]
        tv = ta[s0 * m:(s0 + nb) * m]
        num += float(np.sum((yv - tv) ** 2))
        den += float(np.sum(tv.astype(np.float64) ** 2))
    y.free()
    return float(np.sqrt(num / den))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights", "cnn_e32.npz"))
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=0)
    a = ap.parse_args(argv)
    ctx = mgx.Ctx(0)
    t0 = time.perf_counter()
    x, t, ns, edge = learned.build_dataset(ctx, range(1000, 1000 + a.seeds), BOX_KW, a.steps, a.every, log=log)
    t_data = time.perf_counter() - t0
    log("dataset: %d pairs of %d^3 in %.1f s" % (ns, edge, t_data))
    vx, vt, vns, _ = learned.build_dataset(ctx, [2000], dict(BOX_KW), 40, 20, log=log)
    cnn = learned.Cnn(ctx, edge, a.batch, learned.init_params(seed=a.init_seed, skip=0.0))
    e0 = rel_error(cnn, vx, vt, vns, a.batch)
    t1 = time.perf_counter()
    hist = learned.train(cnn, x, t, ns, a.epochs, a.batch, a.lr, log=log)
    t_train = time.perf_counter() - t1
    e1 = rel_error(cnn, vx, vt, vns, a.batch)
    log("held-out relative error: untrained %.4f -> trained %.4f" % (e0, e1))
    p = cnn.params()
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    meta = dict(pairs=ns, edge=edge, epochs=a.epochs, batch=a.batch, lr=a.lr, seeds=a.seeds, steps=a.steps,
                every=a.every, box=dict(BOX_KW, edges=list(BOX_KW["edges"])), loss=hist, heldout_rel_error=e1,
                heldout_rel_error_untrained=e0, seconds_dataset=t_data, seconds_train=t_train)
    np.savez(a.out, **{k: np.asarray(v, dtype=np.float32) for k, v in p.items()}, meta=json.dumps(meta))
    print(json.dumps(meta))
    cnn.close()
    for d in (x, t, vx, vt):
        d.free()
    ctx.close()


def load(path):
    """weights saved by main() -> (params dict, meta dict)."""
    z = np.load(path, allow_pickle=False)
    p = {k: z[k] for k in learned.NAMES}
    p["s"] = np.float32(p["s"])
    return p, json.loads(str(z["meta"]))


if __name__ == "__main__":
    main()
