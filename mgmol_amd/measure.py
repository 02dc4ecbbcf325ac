"""Measurements of the other workloads of the measurement plan, one JSON line
each (the converged C5 solve is bench.py's):

    python -m mgmol_amd.measure c1     # full-grid orbitals through the drop-in entry points
    python -m mgmol_amd.measure c2     # batched uneven boxes: a converged solve
    python -m mgmol_amd.measure c3     # tall-skinny S / H contractions, FP64 and split
    python -m mgmol_amd.measure c4     # the CNN: inference and training throughput

c2: 512 orbitals in boxes of edge {24..40} on a 192^3 grid, h = 0.25 bohr,
    16 wells, centres at least 2 bohr apart; ABPG (Anderson m = 1) to
    |R| < 1e-6 on the plain and mixed paths, kernel time per group.
c3: the dense 256^3 x 1024 problem (N = 1024 Gaussian orbitals with sigma ~
    U(0.8, 1.5) bohr, h = 0.2, 64 wells) streamed through 32 z-slabs of 8
    planes: S = h^3 Psi^T Psi and H = h^3 Psi^T (H Psi) in FP64 (library
    dgemm) and with the hi/lo split (hand-written FP32 MFMA kernel), summed
    over the slabs in FP64, then the FP64 Cholesky + inverse of S;
    FP64-equivalent flops K N (N+1) and 2 K N^2, FP32 MFMA utilisation of the
    split; error of the split against the FP64 result.
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
    """median wall time of reps calls (each entry point waits for its own
    kernels), after one warm-up call"""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _cpu_info():
    import os
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"cpu_model": model, "cpus_available": avail,
            "omp_threads": int(os.environ.get("OMP_NUM_THREADS") or 0) or avail}


def c1(ctx, cpu=True):
    """The plain FP64 method on full-grid orbitals through the drop-in entry
    points: 64 orbitals on a 96^3 periodic grid (h = 0.3, 16 wells), H psi
    (mgx_hloc_apply, 4th order), the FP64 3-level V-cycle (mgx_precond_mg:
    2/2 sweeps, 4 at the coarsest, omega = 2/3 / 6/7), S (mgx_local_gram);
    algorithmic HBM bytes per call (H psi: psi read, H psi written, V once;
    V-cycle: 18.5 x 8 B per orbital point for the fused minimum schedule)
    and the oracle's CPU time for the same calls on this host."""
    kw = dict(synth.CONFIGS["c1"])
    n, gd, h, nwells = kw["n"], kw["gdim"], kw["h"], kw["nwells"]
    psi, v = synth.grid_orbitals(n, gd, h, seed=1, nwells=nwells)
    K = gd ** 3
    grid = mgx.Grid.make((gd, gd, gd), 2, (h, h, h))
    dp, dv = mgx.DeviceArray(ctx, psi), mgx.DeviceArray(ctx, v)
    dh = mgx.DeviceArray(ctx, shape=(n, K), dtype=np.float64)
    dr = mgx.DeviceArray(ctx, shape=(n, K), dtype=np.float64)
    dS = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
    L = mgx.lib()
    vel = h ** 3
    gam = mgx.lib().mgx_gamma((C.c_double * 3)(h, h, h), 2, float(np.max(v)), -1.0)
    t_h = _timeit(lambda: ctx.check(L.mgx_hloc_apply(ctx.h, mgx.LAP_4TH, C.byref(grid), mgx.F64, mgx.ptr(dp), K, n,
                                                     mgx.ptr(dv), mgx.ptr(dh), K), "hloc"), 30)
    t_p = _timeit(lambda: ctx.check(L.mgx_precond_mg(ctx.h, mgx.LAP_4TH, 2, gam, mgx.F64, C.byref(grid), mgx.ptr(dh),
                                                     K, n, None, mgx.ptr(dr), K), "precond"), 15)
    t_s = _timeit(lambda: ctx.check(L.mgx_local_gram(ctx.h, mgx.PREC_F64, mgx.ptr(dp), K, K, n, 1, vel,
                                                     mgx.ptr(dS)), "gram"), 30)
    bh = 16.0 * n * K + 8.0 * K
    bp = 18.5 * 8.0 * n * K
    fs = K * n * (n + 1.0)
    out = {"config": "c1", "orbitals": n, "grid": gd, "points": K,
           "hloc_ms": round(1e3 * t_h, 3), "hloc_GBps": round(bh / t_h / 1e9, 1),
           "hloc_frac_of_8TBps": round(bh / t_h / 8e12, 4),
           "vcycle_f64_ms": round(1e3 * t_p, 3), "vcycle_GBps_algorithmic": round(bp / t_p / 1e9, 1),
           "vcycle_frac_of_8TBps": round(bp / t_p / 8e12, 4),
           "gram_ms": round(1e3 * t_s, 3), "gram_tflops": round(fs / t_s / 1e12, 2)}
    if cpu:
        import oracle as O
        og = O.Grid.make((gd, gd, gd), 2, (h, h, h))
        t0 = time.perf_counter()
        hp = O.hloc(og, 2, O.embed(og, psi.reshape((n, gd, gd, gd))), v)
        c_h = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.precond(og, 2, 2, gam, hp, dtype=np.float64)
        c_p = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.local_gram(psi, K, 1, vel)
        c_s = time.perf_counter() - t0
        out["cpu_oracle"] = dict(_cpu_info(), hloc_ms=round(1e3 * c_h, 1), vcycle_f64_ms=round(1e3 * c_p, 1),
                                 gram_ms=round(1e3 * c_s, 1),
                                 note="oracle (C++/OpenMP restatement of the reference), one call each")
        hg = dh.get()
        out["hloc_max_rel_err_vs_oracle"] = float(np.max(np.abs(hg - hp)) / np.max(np.abs(hp)))
    for a in (dp, dv, dh, dr, dS):
        a.free()
    return out


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


def _separable_factors(n, gdim, h, seed):
    """1D factors of n Gaussian orbitals (sigma ~ U(0.8, 1.5) bohr, centres
    ~ U([0, L)^3)) on a periodic gdim^3 grid, each orbital normalised to
    h^3 sum psi^2 = 1, and the factors' 1D 4th-order second differences
    (-d^2/dx^2: c0 = 30/(12 h^2), c1 = -16/(12 h^2), c2 = 1/(12 h^2))."""
    rng = np.random.default_rng(seed)
    L = gdim * h
    cen = rng.uniform(0.0, 1.0, (n, 3)) * L
    sig = rng.uniform(0.8, 1.5, n)
    x = np.arange(gdim) * h
    c0, c1, c2 = 30.0 / (12 * h * h), -16.0 / (12 * h * h), 1.0 / (12 * h * h)
    f, lf = [], []
    for d in range(3):
        dd = x[None, :] - cen[:, d:d + 1]
        dd -= L * np.round(dd / L)
        f.append(np.exp(-dd * dd / (2.0 * sig[:, None] ** 2)))
    norm = np.sqrt(h ** 3 * np.prod([np.sum(g * g, axis=1) for g in f], axis=0))
    f[0] /= norm[:, None]
    for g in f:
        lf.append(c0 * g + c1 * (np.roll(g, 1, 1) + np.roll(g, -1, 1)) + c2 * (np.roll(g, 2, 1) + np.roll(g, -2, 1)))
    return f, lf, rng


def c3(ctx, nslab=None):
    """The dense contraction workload in full: S = h^3 Psi^T Psi and
    H = h^3 Psi^T (H Psi) over the 256^3 grid, N = 1024, streamed through
    32 z-slabs of 8 planes (K = 524,288 rows each; Psi and H Psi of a slab
    generated on the device, H = -Delta_h4 + V with 64 wells), slab results
    summed in FP64; then the FP64 Cholesky + inverse of S.  FP64 (library
    dgemm) and the hi/lo split (hand-written FP32 MFMA kernel) on the same
    slabs; the split's error against FP64 for every entry."""
    n, gdim, h, nzs, nwells = 1024, 256, 0.2, 8, 64
    nslab = nslab or gdim // nzs
    f, lf, rng = _separable_factors(n, gdim, h, seed=3)
    v, _, _ = synth.wells((gdim,) * 3, h, nwells, rng)
    K = gdim * gdim * nzs
    vel = h ** 3
    dev = [mgx.DeviceArray(ctx, np.ascontiguousarray(a)) for a in f + lf]
    dv = mgx.DeviceArray(ctx, np.ascontiguousarray(v))
    del v
    dp = mgx.DeviceArray(ctx, shape=(n, K), dtype=np.float64, zero=False)
    dh = mgx.DeviceArray(ctx, shape=(n, K), dtype=np.float64, zero=False)
    tmp = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
    precs = ((mgx.PREC_F64, "f64"), (mgx.PREC_SPLIT, "split"))
    acc = {(p, m): mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64) for p, _ in precs for m in "SH"}
    gd = (C.c_int * 3)(gdim, gdim, gdim)
    t_gen, t = 0.0, {(p, m): 0.0 for p, _ in precs for m in "SH"}
    L = mgx.lib()
    for sl in range(nslab):
        t0 = time.perf_counter()
        ctx.check(L.mgx_synth_separable(ctx.h, n, gd, sl * nzs, nzs, *[mgx.ptr(d) for d in dev], mgx.ptr(dv),
                                        mgx.ptr(dp), mgx.ptr(dh)), "mgx_synth_separable")
        t_gen += time.perf_counter() - t0
        for p, _ in precs:
            t0 = time.perf_counter()
            ctx.check(L.mgx_local_gram(ctx.h, p, mgx.ptr(dp), K, K, n, 1, vel, mgx.ptr(tmp)), "gram")
            t[(p, "S")] += time.perf_counter() - t0
            ctx.check(L.mgx_axpy(ctx.h, mgx.F64, n * n, 1.0, mgx.ptr(tmp), mgx.ptr(acc[(p, "S")])), "axpy")
            t0 = time.perf_counter()
            ctx.check(L.mgx_local_product(ctx.h, p, mgx.ptr(dp), mgx.ptr(dh), K, K, n, 1, vel, mgx.ptr(tmp)),
                      "product")
            t[(p, "H")] += time.perf_counter() - t0
            ctx.check(L.mgx_axpy(ctx.h, mgx.F64, n * n, 1.0, mgx.ptr(tmp), mgx.ptr(acc[(p, "H")])), "axpy")
    # GramMatrix::updateLS / computeInverse on the summed FP64 S
    Lm = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
    Si = mgx.DeviceArray(ctx, shape=(n, n), dtype=np.float64)
    t0 = time.perf_counter()
    ctx.check(L.mgx_chol_inv(ctx.h, n, mgx.ptr(acc[(mgx.PREC_F64, "S")]), mgx.ptr(Lm), mgx.ptr(Si)), "chol_inv")
    t_chol = time.perf_counter() - t0
    Ktot = K * nslab
    fs, fh = Ktot * n * (n + 1.0), 2.0 * Ktot * n * n
    nt = n // 64
    # executed FP32 MFMA flops of the split: three products per output; the
    # Gram computes the 64 x 64 tiles on and below the diagonal
    xs, xh = 3 * 2.0 * Ktot * 64 * 64 * nt * (nt + 1) / 2, 3 * 2.0 * Ktot * n * n
    out = {"config": "c3", "K": Ktot, "N": n, "slabs": nslab, "rows_per_slab": K,
           "note": "Psi and H Psi generated on the device per 8-plane z-slab, slab products summed in FP64",
           "generate_s": round(t_gen, 3), "cholesky_inverse_ms": round(1e3 * t_chol, 3)}
    for p, name in precs:
        ts, th = t[(p, "S")], t[(p, "H")]
        out[name] = {"S_s": round(ts, 4), "H_s": round(th, 4), "S_plus_H_s": round(ts + th, 4),
                     "S_tflops_fp64_equiv": round(fs / ts / 1e12, 2), "H_tflops_fp64_equiv": round(fh / th / 1e12, 2)}
        if p == mgx.PREC_SPLIT:
            out[name]["S_fp32_mfma_tflops"] = round(xs / ts / 1e12, 2)
            out[name]["H_fp32_mfma_tflops"] = round(xh / th / 1e12, 2)
            out[name]["S_fp32_mfma_utilisation"] = round(xs / ts / 157.3e12, 4)
            out[name]["H_fp32_mfma_utilisation"] = round(xh / th / 157.3e12, 4)
        else:
            out[name]["S_fp64_mfma_utilisation_of_equiv"] = round(fs / ts / 78.6e12, 4)
            out[name]["H_fp64_mfma_utilisation"] = round(fh / th / 78.6e12, 4)
    err = {}
    for m in "SH":
        a64, a32 = acc[(mgx.PREC_F64, m)].get(), acc[(mgx.PREC_SPLIT, m)].get()
        d = np.abs(a32 - a64)
        err[m + "_max_abs"] = float(np.max(d))
        err[m + "_max_rel_to_largest"] = float(np.max(d) / np.max(np.abs(a64)))
        err[m + "_max_rel_diag"] = float(np.max(np.abs(np.diag(a32) - np.diag(a64)) / np.abs(np.diag(a64))))
    out["split_error_vs_fp64"] = err
    for a in dev + [dv, dp, dh, tmp, Lm, Si] + list(acc.values()):
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


def pb(ctx):
    """Poisson-Boltzmann solves on 256^3 (h = 0.2, periodic, 2 ghosts): the
    dielectric function of a density of 64 Gaussian blobs (values across
    rho0), a zero-mean charge; up to 10 V-cycles to 1e-10 with PBh4M, PBh4
    and PBh2, the dielectric KS potential included."""
    n, h, ng = 256, 0.2, 2
    e0, rho0, drho0 = 78.36, 0.0004, 1.3
    g = mgx.Grid.make((n, n, n), ng, [h] * 3)
    rng = np.random.default_rng(11)
    x = (np.arange(n) * h).astype(np.float64)
    L = n * h
    rhod_i = np.zeros((n, n, n))
    for c in rng.uniform(0, L, (64, 3)):
        f = [np.exp(-np.minimum(np.abs(x - c[k]), L - np.abs(x - c[k])) ** 2 / 1.2) for k in range(3)]
        rhod_i += 0.01 * (f[0][:, None, None] * f[1][None, :, None]) * f[2][None, None, :]
    sh = (n + 2 * ng,) * 3
    inner = (slice(ng, ng + n),) * 3
    rhod = np.zeros(sh)
    rhod[inner] = rhod_i
    rho = np.zeros(sh)
    q = rng.standard_normal((n, n, n))
    rho[inner] = 4 * np.pi * (q - q.mean())
    del rhod_i, q
    trho, trhod = mgx.DeviceArray(ctx, rho), mgx.DeviceArray(ctx, rhod)
    tvh = mgx.DeviceArray(ctx, shape=sh, dtype=np.float64)
    tvks = mgx.DeviceArray(ctx, shape=sh, dtype=np.float64)
    zero = np.zeros(sh)
    out = {"config": "pb 256^3 periodic, e0 = 78.36, 64 density blobs, up to 10 V-cycles to 1e-10"}
    for name, lap in (("PBh4M", 0), ("PBh4", 2), ("PBh2", 1)):
        st = mgx.PoissonStats()

        def run():
            ctx.check(mgx.lib().mgx_memcpy_htod(ctx.h, mgx.ptr(tvh), zero.ctypes.data_as(C.c_void_p), tvh.nbytes),
                      "mgx_memcpy_htod")
            ctx.check(mgx.lib().mgx_pb_mg(ctx.h, lap, mgx.block(trho, g, 1), mgx.block(trhod, g, 1),
                                          mgx.block(tvh, g, 1), mgx.block(tvks, g, 1), e0, rho0, drho0, 2, 2, 10,
                                          1e-10, 10, C.byref(st)), "mgx_pb_mg")

        run()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            run()
        t = (time.perf_counter() - t0) / reps
        out[name] = {"ms_per_solve_incl_zeroing": round(1e3 * t, 2), "v_cycles": st.nb_sweeps,
                     "converged": bool(st.converged), "final_relative_residual": st.final_relative_residual,
                     "ms_per_v_cycle": round(1e3 * t / max(st.nb_sweeps, 1), 3)}
    for a in (trho, trhod, tvh, tvks):
        a.free()
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    ctx = mgx.Ctx(0)
    for name in argv or ["c1", "c2", "c3", "c4", "density", "poisson", "pb"]:
        res = {"c1": c1, "c1_gpu": lambda c: c1(c, cpu=False), "c2": c2, "c3": c3, "c4": c4, "density": density,
               "poisson": poisson, "pb": pb}[name](ctx)
        print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
