"""Benchmark: time per converged orbital solve on one MI355X.

Workload (default c5): 2048 localized orbitals in cubic 32^3 boxes on a
256^3 periodic grid, h = 0.2 bohr, 64 Gaussian wells, seeded Gaussian
orbitals (centres at least 2 bohr apart); ABPG steps with Anderson
extrapolation (the reference's default quench solver, m = 1, beta = 1) until
the residual norm is below 1e-6 Ha.  One "step" is one such converged solve.
Three paths run on identical inputs:
  plain   : FP64 V-cycle, FP64 S/H, FP64 dense layer;
  mixed   : FP16 V-cycle (FP32 arithmetic, the residual scaled per box, the
            correction in FP32; the FP32 V-cycle where box edges are not
            multiples of 8 on every level), hi/lo-split S/H on FP32 MFMA,
            FP64 dense layer with the FP32 Newton-Schulz correction;
  learned : mixed, with the trained CNN in place of the V-cycle (BF16 MFMA),
            the V-cycle judging each step and taking over when it fails.
Hpsi, the residual and the energy are FP64 in all three.
value = seconds per converged solve of the mixed+learned path (lower is
better); plain and mixed times, iteration counts and band-energy differences
are reported beside it.

Roofline: the dominant hand-written kernel of the headline path, timed with
hipEvents on the engine's stream inside the timed solves: the S / H cell
partials (k_cell_gram, HBM-bound: psi and H psi read once).  roofline_other
lists the fine-level Jacobi sweep of the mixed path (HBM) and the CNN inference (BF16
MFMA, learned steps only); stream is the achievable HBM rate of a plain
axpy on this device.
CPU baseline: the oracle (C++/OpenMP restatement of the reference) timed on
this host for one full-size orbital step, times the iteration count.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mgmol_amd import mgx, synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFS = 2500.0
CNN_FLOP_PER_VOXEL = 2 * 27 * (1 * 16 + 16 * 16 + 16 * 16 + 16 * 1)
WEIGHTS = os.path.join(ROOT, "mgmol_amd", "weights", "cnn_e32.npz")
# HBM bytes per launch at C5, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes
# (FETCH_SIZE scaled by 2.000, WRITE_SIZE by 1.000, both calibrated on
# streaming kernels of known size; profiles/pmc_traffic_c5.md)
TRAFFIC_C5 = {"cnn_application": 17287778496, "jacobi_fp32_sweep": 895234355, "jacobi_fp16_sweep": 492663945,
              "cell_gram_split": 1540153344}

# the mixed paths take the FP16 V-cycle (FP32 arithmetic, residual scaled per
# box, FP32 correction) where every box edge is a multiple of 8 on every
# level, else the FP32 V-cycle; set_mixed_precond picks it for the workload
PATHS = {
    "plain": dict(precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64),
    "mixed": dict(precond_dtype=mgx.F16, gram_prec=mgx.PREC_SPLIT),
    "learned": dict(precond_dtype=mgx.F16, gram_prec=mgx.PREC_SPLIT, precond_kind=mgx.PRECOND_LEARNED),
    # not in the default modes: the mixed path with the reference's FP32 V-cycle
    "mixed32": dict(precond_dtype=mgx.F32, gram_prec=mgx.PREC_SPLIT),
}
LABELS = {
    "plain": "plain FP64 (FP64 V-cycle, FP64 S/H)",
    "mixed": "mixed (FP16 V-cycle, hi/lo-split S/H)",
    "learned": "mixed+learned (BF16 CNN preconditioner, FP16 V-cycle fallback)",
    "mixed32": "mixed with the FP32 V-cycle (hi/lo-split S/H)",
}


def set_mixed_precond(cfg):
    m = 8 << (cfg.levels - 1)
    pd = mgx.F16 if all(int(e) % m == 0 for e in cfg.edge) else mgx.F32
    for k in ("mixed", "learned"):
        PATHS[k]["precond_dtype"] = pd
        LABELS[k] = LABELS[k].replace("FP16", "FP32") if pd == mgx.F32 else LABELS[k]
    return pd


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_config(name, seed):
    c = dict(synth.CONFIGS[name])
    if c.pop("kind") != "box":
        raise SystemExit("bench supports box workloads (c2, c5)")
    return synth.box_config(seed=seed, **c), c


def run_mode(ctx, bs, cfg, mode, tol, maxit, steps, warmup, cnn):
    kw = dict(mg_levels=cfg.levels - 1, profile=1, precond_start=mgx.START_ZERO, anderson_m=1,
              conv_criterion=mgx.CONV_RESIDUAL)
    kw.update(PATHS[mode])
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, cnn=cnn if mode == "learned" else None, **kw)
    for _ in range(warmup):
        s.run(tol, 3)
        s.reset()
    s.profile(reset=True)
    times, stats = [], None
    for k in range(steps):
        s.reset()
        t0 = time.perf_counter()
        stats = s.run(tol, maxit)
        times.append(time.perf_counter() - t0)
        log("[%s] solve %d: %d iterations, |R|=%.3e Ha, E=%.12f Ha, %.3f s (converged=%d, learned steps %d, "
            "fallbacks %d)" % (mode, k, stats.iterations, stats.residual, stats.band_energy, times[-1],
                               stats.converged, stats.learned_steps, stats.fallbacks))
    prof = s.profile()
    s.close()
    return dict(seconds=float(np.mean(times)), iterations=stats.iterations, residual=stats.residual,
                energy=stats.band_energy, converged=bool(stats.converged), learned_steps=stats.learned_steps,
                fallbacks=stats.fallbacks, newton_schulz_steps=stats.ns_steps, cholesky_calls=stats.chol_calls,
                dots_sweeps=stats.dots_sweeps, profile=prof)


def cpu_baseline(cfg, iters, mg_levels):
    """One full-size orbital step of the oracle (FP32 V-cycle as the reference's
    MGPRECONDTYPE) on this host, scaled by the iteration count."""
    import oracle as O
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    t0 = time.perf_counter()
    O.b_solver_step(ob, cfg.psi.copy(), cfg.vtot, mg_levels, O.F32, 1.5, None, zero_start=True)
    t_step = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return dict(value=round(t_step * iters, 2), unit="s per converged solve", cores=cores, kind="port",
                sample="1 oracle orbital step on the full workload (%.2f s) x %d iterations" % (t_step, iters))


def jacobi_roofline(r, precond_dtype, npts):
    ms, launches = r["profile"]["jacobi0"]
    if not launches:
        return None
    # the 3 fine sweeps launched per V-cycle (the zero-start first sweep is
    # pointwise, fused with the cast): read x, f and write x, plus w on the
    # last pre-sweep -> (4 + 3 + 3) / 3 values per point; a last sweep that
    # also forms the Anderson dots reads the previous correction as well.
    # FP16: 2-byte values, but the last sweep writes (and the dots read) the
    # FP32 correction: (8 + 6 + 8) / 3 bytes per point, + 4 with the dots
    if precond_dtype == mgx.F16:
        per_launch = (22.0 / 3.0 * launches + 4.0 * r["dots_sweeps"]) / launches * npts
        s_bytes = 2
    else:
        s_bytes = 4 if precond_dtype == mgx.F32 else 8
        per_launch = (10.0 * launches + r["dots_sweeps"]) / (3.0 * launches) * s_bytes * npts
    avg_s = (ms / 1e3) / launches
    achieved = per_launch / avg_s / 1e9
    # measured without the Anderson dots (the first 12 steps of a solve)
    key = {4: "jacobi_fp32_sweep", 2: "jacobi_fp16_sweep"}.get(s_bytes)
    traffic = TRAFFIC_C5[key] if (key and npts == 2048 * 32 ** 3) else None
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel="k_box_march<JACOBI> fine-level sweep (%s)" % {2: "FP16", 4: "FP32", 8: "FP64"}[s_bytes],
                avg_launch_ms=round(1e3 * avg_s, 4), launches=int(launches), bytes_per_launch=int(per_launch))


def cellgram_roofline(r, npts, split):
    """k_cell_gram, the S / H cell partials: psi and H psi (FP64) read once,
    16 bytes per orbital point per launch (the partials written are < 2%)."""
    ms, launches = r["profile"]["cellgram"]
    if not launches:
        return None
    per_launch = 16.0 * npts
    avg_s = (ms / 1e3) / launches
    achieved = per_launch / avg_s / 1e9
    key = "cell_gram_split" if split else "cell_gram_f64"
    traffic = TRAFFIC_C5.get(key) if npts == 2048 * 32 ** 3 else None
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel="k_cell_gram<%s> S/H partials per 8^3 cell (MFMA)" % ("SPLIT" if split else "F64"),
                avg_launch_ms=round(1e3 * avg_s, 4), launches=int(launches), bytes_per_launch=int(per_launch))


def cnn_roofline(r, npts):
    ms, launches = r["profile"]["learned"]
    if not launches:
        return None
    per_launch = CNN_FLOP_PER_VOXEL * npts
    avg_s = (ms / 1e3) / launches
    achieved = per_launch / avg_s / 1e12
    traffic = TRAFFIC_C5["cnn_application"] if npts == 2048 * 32 ** 3 else None
    return dict(bound="mfma", achieved=round(achieved, 2), peak=BF16_PEAK_TFS, unit="TFLOP/s",
                frac=round(achieved / BF16_PEAK_TFS, 5), traffic=traffic,
                kernel="CNN inference (4 conv3d layers, BF16 MFMA) over all boxes",
                avg_launch_ms=round(1e3 * avg_s, 3), launches=int(launches), flop_per_launch=int(per_launch))


def stream_reference(ctx, n=1 << 28, reps=20):
    """Achievable HBM rate on this device: the engine's FP32 axpy (y += a x,
    STREAM-triad form: two reads and one write per element) over 1 GiB
    arrays; bytes / wall time of back-to-back launches."""
    import ctypes as C
    a = mgx.DeviceArray(ctx, shape=(n,), dtype=np.float32)
    b = mgx.DeviceArray(ctx, shape=(n,), dtype=np.float32)
    sync = mgx.DeviceArray(ctx, shape=(1,), dtype=np.float64)
    L = mgx.lib()

    def run(k):
        for _ in range(k):
            ctx.check(L.mgx_axpy(ctx.h, mgx.F32, C.c_int64(n), C.c_double(0.5), mgx.ptr(a), mgx.ptr(b)), "mgx_axpy")
        sync.get()

    run(3)
    t0 = time.perf_counter()
    run(reps)
    dt = (time.perf_counter() - t0) / reps
    for d in (a, b, sync):
        d.free()
    gbs = 3 * 4 * n / dt / 1e9
    return dict(kernel="k_axpy<float> (y += a x, 1 GiB arrays: 2 reads + 1 write per element)",
                GBps=round(gbs, 1), frac_of_peak=round(gbs / HBM_PEAK_GBS, 4))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed converged solves per path")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=2000)
    ap.add_argument("--modes", default="plain,mixed,learned")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    modes = args.modes.split(",")

    t0 = time.perf_counter()
    cfg, meta = make_config(args.config, args.seed)
    set_mixed_precond(cfg)
    log("config %s: %d boxes, %d points, grid %s, h=%.3f (generated in %.1f s)"
        % (args.config, cfg.nbox, cfg.psi.size, cfg.gdim, cfg.h, time.perf_counter() - t0))
    ctx = mgx.Ctx(0)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, cfg.levels)
    log("box pairs:", bs.npairs())
    cnn, wmeta = None, None
    if "learned" in modes:
        from mgmol_amd import learned, train_cnn
        params, wmeta = train_cnn.load(args.weights)
        cnn = learned.Cnn(ctx, wmeta["edge"], 256, params)
    res = {}
    for mode in modes:
        res[mode] = run_mode(ctx, bs, cfg, mode, args.tol, args.maxit, args.steps, args.warmup, cnn)
    headname = "learned" if "learned" in res else ("mixed" if "mixed" in res else modes[0])
    head = res[headname]
    npts = cfg.psi.size

    # the headline path's dominant hand-written kernel: the S / H cell
    # partials (one launch per step); beside it the fine FP32 Jacobi sweep
    # (three per step) and the network (learned steps only)
    vc_name = "mixed" if "mixed" in res else headname
    roof = cellgram_roofline(head, npts, PATHS[headname]["gram_prec"] == mgx.PREC_SPLIT)
    other = [x for x in (jacobi_roofline(res[vc_name], PATHS[vc_name]["precond_dtype"], npts),
                         cnn_roofline(res["learned"], npts) if "learned" in res else None) if x]
    if roof is None and other:
        roof, other = other[0], other[1:]

    stream = stream_reference(ctx)
    for r in [roof] + other:
        if r and r["unit"] == "GB/s":
            r["frac_of_stream"] = round(r["achieved"] / stream["GBps"], 4)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(cfg, head["iterations"], cfg.levels - 1)
        except Exception as e:  # reported, never fatal
            log("cpu baseline failed:", e)

    out = {
        "metric": "time per converged solve & SCF iters: mixed+learned vs plain FP64 GPU",
        "value": round(head["seconds"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * head["seconds"], 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ({"learned": "bf16+", "mixed": ""}.get(headname, "")
                  + ("f16+f32+f64" if PATHS[headname]["precond_dtype"] == mgx.F16 else
                     ("f32+f64" if headname != "plain" else "f64"))),
        "data": "synthetic (seeded Gaussian orbitals and wells)",
        "config": {"workload": "%s: %d orbitals in %s^3 boxes, %d^3 periodic grid, h=%.2f bohr, %d wells, "
                               "centres >= %s bohr apart, ABPG (Anderson m=1) to |R| < %g Ha"
                               % (args.config, cfg.nbox, "/".join(str(e) for e in meta["edges"]), cfg.gdim[0],
                                  cfg.h, meta["nwells"], meta.get("min_sep"), args.tol),
                   "path": LABELS[headname]},
        "iterations": head["iterations"],
        "converged": head["converged"],
        "paths": {m: {k: v for k, v in r.items() if k != "profile"} for m, r in res.items()},
        "kernel_ms": {m: {g: round(t[0], 2) for g, t in r["profile"].items()} for m, r in res.items()},
        "roofline": roof,
        "roofline_other": other,
        "stream": stream,
        "cpu_baseline": cpu,
    }
    if "plain" in res:
        for m in res:
            if m == "plain":
                continue
            out["paths"][m]["speedup_vs_plain"] = round(res["plain"]["seconds"] / res[m]["seconds"], 3)
            out["paths"][m]["energy_rel_diff_vs_plain"] = (abs(res[m]["energy"] - res["plain"]["energy"])
                                                           / abs(res["plain"]["energy"]))
    if wmeta:
        out["learned_training"] = {k: wmeta[k] for k in ("pairs", "epochs", "heldout_rel_error", "seconds_train",
                                                         "seconds_dataset") if k in wmeta}
    if cnn:
        cnn.close()
    bs.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
