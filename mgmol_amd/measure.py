"""Measurements of the other workloads of the measurement plan, one JSON line
each (the converged C5 solve is bench.py's):

    python -m mgmol_amd.measure c2     # batched uneven boxes: a converged solve
    python -m mgmol_amd.measure c3     # tall-skinny S / H contractions, FP64 and split
    python -m mgmol_amd.measure c4     # the CNN: inference and training throughput

c2: 512 orbitals in boxes of edge {24..40} on a 192^3 grid, h = 0.25 bohr,
    16 wells, centres at least 2 bohr apart; ABPG (Anderson m = 1) to
    |R| < 1e-6 on the plain and mixed paths, kernel time per group.
c3: one z-slab of the dense 256^3 x 1024 problem (8 planes, K = 524,288 grid
    points, N = 1024 Gaussian orbitals with sigma ~ U(0.8, 1.5) bohr, h = 0.2):
    S = h^3 Psi^T Psi and H = h^3 Psi^T (Psi') in FP64 (dsyrk / dgemm) and
    with the hi/lo split (FP32 chunk products summed in FP64);
    FP64-equivalent flops K N (N+1) and 2 K N^2; error of the split against
    the FP64 result.
c4: inference over 2048 boxes of 32^3 (29,376 flop per voxel) and Adam steps
    on 64-box batches (forward + backward + weight gradients).
poisson: the Hartree multigrid (mgx_poisson_mg, Mehrstellen, nu1 = nu2 = 2) on
    the 256^3 grid of c5 (h = 0.2 bohr) from zero to |r| / |rhs| < 1e-10, for a
    seeded zero-mean 4 pi rho.
density: rho = sum_ij (2 S^-1)_ij psi_i psi_j on the c5 boxes (2048 x 32^3 on
    a 256^3 grid), mgx_box_density; algorithmic bytes: psi read once (8 per
    orbital point) and rho zeroed and written (16 per grid point).
"""
import ctypes as C
import json
import sys
import time

import numpy as np

from . import learned, mgx, synth


def _timeit(fn, reps):
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return (time.perf_counter() - t0) / reps


def c2(ctx):
    kw = dict(synth.CONFIGS["c2"])
    kw.pop("kind")
    cfg = synth.box_config(seed=1, **kw)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, cfg.levels)
    out = {"config": "c2", "boxes": cfg.nbox, "points": int(cfg.psi.size), "pairs": bs.npairs()}
    for mode, pk in (("plain", dict(precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64)),
                     ("mixed", dict(precond_dtype=mgx.F32, gram_prec=mgx.PREC_SPLIT))):
        s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=cfg.levels - 1, profile=1, anderson_m=1,
                       precond_start=mgx.START_ZERO, conv_criterion=mgx.CONV_RESIDUAL, **pk)
        s.run(1e-6, 3)
        s.reset()
        s.profile(reset=True)
        t0 = time.perf_counter()
        st = s.run(1e-6, 3000)
        dt = time.perf_counter() - t0
        prof = s.profile()
        out[mode] = {"seconds": round(dt, 3), "iterations": st.iterations, "converged": bool(st.converged),
                     "residual": st.residual, "energy": st.band_energy,
                     "ms_per_iteration": round(1e3 * dt / max(st.iterations, 1), 3),
                     "kernel_ms_per_iteration": {g: round(v[0] / max(st.iterations, 1), 3) for g, v in prof.items()}}
        s.close()
    bs.close()
    return out


def _slab_orbitals(n, gdim, h, nz, seed):
    """Gaussian orbitals (sigma ~ U(0.8, 1.5)) on the z-planes 0..nz-1 of a
    periodic gdim^3 grid, normalised over the full grid: (n, gdim*gdim*nz)."""
    rng = np.random.default_rng(seed)
    L = gdim * h
    cen = rng.uniform(0.0, 1.0, (n, 3)) * L
    sig = rng.uniform(0.8, 1.5, n)
    x = np.arange(gdim) * h
    z = np.arange(nz) * h
    psi = np.zeros((n, gdim, gdim, nz))
    for j in range(n):
        f = []
        for d, coords in enumerate((x, x, z)):
            dd = coords - cen[j, d]
            dd -= L * np.round(dd / L)
            f.append(np.exp(-dd * dd / (2.0 * sig[j] ** 2)))
        psi[j] = np.einsum("i,j,k->ijk", *f)
        # normalisation of the full-grid Gaussian: h^3 sum = (sqrt(pi) sigma)^3
        psi[j] /= np.sqrt((np.sqrt(np.pi) * sig[j]) ** 3)
    return psi.reshape(n, -1)


def c3(ctx):
    n, gdim, h, nz = 1024, 256, 0.2, 8
    psi = _slab_orbitals(n, gdim, h, nz, seed=3)
    K = psi.shape[1]
    # psi' = psi with a perturbation, standing in for H psi in the product
    rng = np.random.default_rng(4)
    hp = psi + 1e-3 * rng.standard_normal(psi.shape) * (psi != 0)
    vel = h ** 3
    dp, dh = mgx.DeviceArray(ctx, psi), mgx.DeviceArray(ctx, hp)
    del psi, hp
    S = {p: mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64) for p in (mgx.PREC_F64, mgx.PREC_SPLIT)}
    H = {p: mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64) for p in (mgx.PREC_F64, mgx.PREC_SPLIT)}
    out = {"config": "c3", "K": K, "N": n, "note": "one 8-plane z-slab of the 256^3 grid"}
    for p, name in ((mgx.PREC_F64, "f64"), (mgx.PREC_SPLIT, "split")):
        ts = _timeit(lambda: ctx.check(mgx.lib().mgx_local_gram(ctx.h, p, mgx.ptr(dp), K, K, n, 1, vel,
                                                                mgx.ptr(S[p])), "gram"), 5)
        th = _timeit(lambda: ctx.check(mgx.lib().mgx_local_product(ctx.h, p, mgx.ptr(dp), mgx.ptr(dh), K, K, n, 1,
                                                                   vel, mgx.ptr(H[p])), "product"), 5)
        fs, fh = K * n * (n + 1.0), 2.0 * K * n * n
        out[name] = {"S_ms": round(1e3 * ts, 3), "H_ms": round(1e3 * th, 3),
                     "S_tflops_fp64_equiv": round(fs / ts / 1e12, 2), "H_tflops_fp64_equiv": round(fh / th / 1e12, 2)}
    s64, s32 = S[mgx.PREC_F64].get(), S[mgx.PREC_SPLIT].get()
    h64, h32 = H[mgx.PREC_F64].get(), H[mgx.PREC_SPLIT].get()
    out["split_error"] = {"S_max_abs": float(np.max(np.abs(s32 - s64))),
                          "S_max_rel": float(np.max(np.abs(s32 - s64)) / np.max(np.abs(s64))),
                          "H_max_abs": float(np.max(np.abs(h32 - h64))),
                          "H_max_rel": float(np.max(np.abs(h32 - h64)) / np.max(np.abs(h64)))}
    for a in list(S.values()) + list(H.values()) + [dp, dh]:
        a.free()
    return out


def c4(ctx):
    e, nbox, batch = 32, 2048, 64
    m = e ** 3
    net = learned.Cnn(ctx, e, 256, learned.init_params(seed=0))
    rng = np.random.default_rng(5)
    R = mgx.DeviceArray(ctx, rng.standard_normal(nbox * m))
    out_d = mgx.DeviceArray(ctx, shape=(nbox * m,), dtype=np.float64, zero=False)
    t_inf = _timeit(lambda: net.apply(nbox, R, out_d), 5)
    flop = 2 * 27 * (16 + 256 + 256 + 16) * m * nbox
    tnet = learned.Cnn(ctx, e, batch, learned.init_params(seed=1))
    x = mgx.DeviceArray(ctx, rng.standard_normal(batch * m).astype(np.float32))
    t = mgx.DeviceArray(ctx, (0.03 * rng.standard_normal(batch * m)).astype(np.float32))
    t_step = _timeit(lambda: tnet.train_step(x.data_ptr(), t.data_ptr(), batch, 1e-4), 10)
    # training flops: forward + data gradient (layers 1-3) + weight gradients (all)
    fwd = 2 * 27 * (16 + 256 + 256 + 16) * m * batch
    train = fwd + 2 * 27 * (256 + 256 + 16) * m * batch + fwd
    out = {"config": "c4", "inference_ms_2048_boxes": round(1e3 * t_inf, 3),
           "inference_tflops": round(flop / t_inf / 1e12, 1),
           "train_step_ms_batch64": round(1e3 * t_step, 3), "train_samples_per_s": round(batch / t_step, 1),
           "train_tflops": round(train / t_step / 1e12, 1)}
    for a in (R, out_d, x, t):
        a.free()
    net.close()
    tnet.close()
    return out


def density(ctx):
    kw = dict(synth.CONFIGS["c5"])
    kw.pop("kind")
    cfg = synth.box_config(seed=1, **kw)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, cfg.levels)
    n = cfg.nbox
    psi = mgx.DeviceArray(ctx, cfg.psi)
    S = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
    H = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
    ctx.check(mgx.lib().mgx_box_gram_h(ctx.h, bs.h, mgx.PREC_F64, mgx.ptr(psi), mgx.ptr(psi), mgx.ptr(S),
                                       mgx.ptr(H)), "gram")
    dm = mgx.DeviceArray(ctx, 2.0 * np.linalg.inv(S.get()))
    ng = int(np.prod(cfg.gdim))
    rho = mgx.DeviceArray(ctx, shape=(ng,), dtype=np.float64, zero=False)
    q = [0.0]

    def run():
        q[0] = bs.density(psi, dm, rho)

    t = _timeit(run, 10)
    nbytes = 8 * cfg.psi.size + 16 * ng
    out = {"config": "density (c5 boxes)", "ms": round(1e3 * t, 3), "charge": q[0], "electrons_expected": 2 * n,
           "algorithmic_MB": round(nbytes / 1e6, 1), "GBps": round(nbytes / t / 1e9, 1)}
    for a in (psi, S, H, dm, rho):
        a.free()
    bs.close()
    return out


def poisson(ctx):
    n, h = 256, 0.2
    g = mgx.Grid.make((n, n, n), 1, [h] * 3)
    rng = np.random.default_rng(9)
    rho = np.zeros((n + 2,) * 3)
    inner = rng.standard_normal((n, n, n))
    rho[1:-1, 1:-1, 1:-1] = 4 * np.pi * (inner - inner.mean())
    trho = mgx.DeviceArray(ctx, rho)
    tvh = mgx.DeviceArray(ctx, shape=rho.shape, dtype=np.float64)
    st = mgx.PoissonStats()

    zero = np.zeros_like(rho)

    def run():
        ctx.check(mgx.lib().mgx_memcpy_htod(ctx.h, mgx.ptr(tvh), zero.ctypes.data_as(C.c_void_p), tvh.nbytes),
                  "mgx_memcpy_htod")
        ctx.check(mgx.lib().mgx_poisson_mg(ctx.h, 0, mgx.block(trho, g, 1), mgx.block(tvh, g, 1), 2, 2, 40, 1e-10, 10,
                                           C.byref(st)), "mgx_poisson_mg")

    run()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        run()
    t = (time.perf_counter() - t0) / reps
    out = {"config": "poisson 256^3 Mehrstellen", "ms_per_solve_incl_zeroing": round(1e3 * t, 2),
           "v_cycles": st.nb_sweeps, "converged": bool(st.converged),
           "final_relative_residual": st.final_relative_residual,
           "ms_per_v_cycle": round(1e3 * t / max(st.nb_sweeps, 1), 3)}
    trho.free()
    tvh.free()
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    ctx = mgx.Ctx(0)
    for name in argv or ["c2", "c3", "c4", "density", "poisson"]:
        res = {"c2": c2, "c3": c3, "c4": c4, "density": density, "poisson": poisson}[name](ctx)
        print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
