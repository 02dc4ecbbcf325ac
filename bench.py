"""Benchmark: time per converged orbital solve on one MI355X.

Workload (default): 2048 localized orbitals in cubic 32^3 boxes on a 256^3
periodic grid, h = 0.2 bohr, 64 Gaussian wells, seeded Gaussian orbitals;
iterate until the residual norm (Ha) is below 1e-6.  One "step" is one such
converged solve.  Two paths run on identical inputs:
  plain : FP64 stencils, FP64 V-cycle, FP64 S/H, FP64 dense layer;
  mixed : FP64 stencils/residual, FP32 V-cycle, hi/lo-split S/H, FP64 dense.
value = seconds per converged solve of the mixed path (lower is better);
the plain path's time, both iteration counts and the band-energy difference
are reported beside it.

Roofline: the fine-level Jacobi sweep (the launch that moves the most bytes),
timed with hipEvents on the engine's stream during the timed solves.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_config(name, seed):
    c = dict(synth.CONFIGS[name])
    if c.pop("kind") != "box":
        raise SystemExit("bench supports box workloads (c2, c5)")
    return synth.box_config(seed=seed, **c), c


def run_mode(ctx, bs, cfg, mode, tol, maxit, steps, warmup):
    kw = dict(mg_levels=cfg.levels - 1, profile=1, precond_start=mgx.START_ZERO)
    if mode == "plain":
        kw.update(precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64)
    else:
        kw.update(precond_dtype=mgx.F32, gram_prec=mgx.PREC_SPLIT)
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, **kw)
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
        log("[%s] solve %d: %d iterations, |R|=%.3e Ha, E=%.12f Ha, %.3f s (converged=%d)"
            % (mode, k, stats.iterations, stats.residual, stats.band_energy, times[-1], stats.converged))
    prof = s.profile()
    s.close()
    return dict(seconds=float(np.mean(times)), iterations=stats.iterations, residual=stats.residual,
                energy=stats.band_energy, converged=bool(stats.converged), profile=prof)


def cpu_baseline(cfg, iters, mg_levels):
    """One full-size orbital step of the oracle (FP32 V-cycle as the reference's
    MGPRECONDTYPE) on this host, scaled by the iteration count."""
    import oracle as O
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    t0 = time.perf_counter()
    O.b_solver_step(ob, cfg.psi.copy(), cfg.vtot, mg_levels, O.F32, 1.5, None, zero_start=True)
    t_step = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return dict(value=t_step * iters, unit="s per converged solve", cores=cores, kind="port",
                sample="1 oracle orbital step on the full workload (%.1f s) x %d iterations" % (t_step, iters))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed converged solves per path")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=3000)
    ap.add_argument("--modes", default="plain,mixed")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    cfg, meta = make_config(args.config, args.seed)
    log("config %s: %d boxes, %d points, grid %s, h=%.3f (generated in %.1f s)"
        % (args.config, cfg.nbox, cfg.psi.size, cfg.gdim, cfg.h, time.perf_counter() - t0))
    ctx = mgx.Ctx(0)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, cfg.levels)
    log("box pairs:", bs.npairs())
    res = {}
    for mode in args.modes.split(","):
        res[mode] = run_mode(ctx, bs, cfg, mode, args.tol, args.maxit, args.steps, args.warmup)
    head = res.get("mixed") or res[list(res)[0]]
    npts = cfg.psi.size

    # roofline of the fine-level Jacobi sweep of the headline path
    ms, launches = head["profile"]["jacobi0"]
    s_bytes = 4 if "mixed" in res else 8
    # the 3 fine sweeps launched per V-cycle (the zero-start first sweep is
    # pointwise, fused with the cast): read x, f and write x, plus w on the
    # last pre-sweep -> (4 + 3 + 3) / 3 values per point
    per_launch_bytes = (10.0 / 3.0) * s_bytes * npts
    avg_s = (ms / 1e3) / max(launches, 1)
    achieved = per_launch_bytes / avg_s / 1e9 if launches else 0.0
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel="k_box_jacobi (fine level, %s)" % ("FP32" if s_bytes == 4 else "FP64"),
                avg_launch_ms=round(1e3 * avg_s, 4), launches=int(launches),
                bytes_per_launch=int(per_launch_bytes))

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
        "dtype": "f32+f64" if "mixed" in res else "f64",
        "data": "synthetic (seeded Gaussian orbitals and wells)",
        "config": {"workload": "%s: %d orbitals in %s boxes, %d^3 periodic grid, h=%.2f bohr, %d wells, |R|<%g Ha"
                   % (args.config, cfg.nbox, "/".join(str(e) for e in meta["edges"]), cfg.gdim[0], cfg.h,
                      meta["nwells"], args.tol),
                   "path": "mixed (FP32 V-cycle, split S/H)" if "mixed" in res else "plain FP64"},
        "paths": {m: {k: v for k, v in r.items() if k != "profile"} for m, r in res.items()},
        "kernel_ms": {m: {g: round(t[0], 2) for g, t in r["profile"].items()} for m, r in res.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "plain" in res and "mixed" in res:
        out["speedup_vs_plain"] = round(res["plain"]["seconds"] / res["mixed"]["seconds"], 3)
        out["energy_rel_diff"] = abs(res["mixed"]["energy"] - res["plain"]["energy"]) / abs(res["plain"]["energy"])
    bs.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
