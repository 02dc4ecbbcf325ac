"""Benchmark: time per converged orbital solve on one MI355X.

Workload (default c5): 2048 localized orbitals in cubic 32^3 boxes on a
256^3 periodic grid, h = 0.2 bohr, 64 Gaussian wells, seeded Gaussian
orbitals (centres at least 2 bohr apart); ABPG steps with Anderson
extrapolation (the reference's default quench solver, m = 1, beta = 1) until
the residual norm is below 1e-6 Ha.  One "step" is one such converged solve.
Paths, on identical inputs:
  plain   : FP64 V-cycle, FP64 S/H, S^-1 refreshed with FP64 Newton-Schulz
            products to 1e-15;
  mixed   : FP32 V-cycle (the reference's MGPRECONDTYPE), hi/lo-split S/H on
            FP32 MFMA with FP64 totals, S^-1 refreshed with FP64 Newton-Schulz
            products (to 1e-10, the split S's own accuracy) and Cholesky in
            FP64 where the refresh fails;
  learned : mixed, with the trained BF16 CNN tried in place of the V-cycle
            (a plain preconditioned step, compared with the V-cycle step
            from the same state; the V-cycle takes over after a rejection);
  ext16   : an extension beyond the specified precisions: FP16 V-cycle
            storage (FP32 arithmetic, residual scaled per box) and the
            Newton-Schulz correction as one BF16 hi/lo product.
H psi, the residual and the energy are FP64 in all paths.
value = seconds per converged solve of the mixed+learned path (lower is
better), timed with kernel profiling off; the other paths' times, iteration
counts and band-energy differences are reported beside it.

Roofline: every kernel group the solver brackets with hipEvents (one
profiled solve per path after the timed ones) carries its algorithmic bytes
(or flops); the headline path's groups are ranked by time per solve, the
top one is `roofline`, the others above 5% of a solve are `roofline_other`.
`stream` is the measured HBM ceiling: a STREAM triad with 16-byte lanes.
CPU baseline: the oracle (C++/OpenMP restatement of the reference) timed on
this host for a bounded number of full-size orbital steps (about 12 s),
scaled by the headline path's iteration count.
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
PEAK_TFS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}
WEIGHTS = os.path.join(ROOT, "mgmol_amd", "weights", "cnn_e32.npz")
C5_POINTS = 2048 * 32 ** 3
# HBM bytes per launch at C5 from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# passes (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000, both calibrated on
# streaming kernels of known size), copied from profiles/pmc_traffic_c5.md;
# keyed by (group, precision of the path's kernel)
TRAFFIC_C5 = {("jacobi0", "f32"): 855957812, ("jacobi0", "f16"): 428524544, ("cellgram", "split", 1): 1423482880,
              ("cellgram", "split", 2): 1308473754, ("residk", "f32"): 1738115277,
              ("learned", "bf16"): 17287778496, ("skin", "f32"): 1679062938, ("skin", "f64"): 1679062938,
              ("skin", "f16"): 1679062938, ("nsgemm", "f64"): 912851968}
TRAFFIC_SOURCE = "profiles/pmc_traffic_c5.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; calibrated)"

PATHS = {
    "plain": dict(precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64, inverse_mode=mgx.INV_NS_F64),
    "mixed": dict(precond_dtype=mgx.F32, gram_prec=mgx.PREC_SPLIT, inverse_mode=mgx.INV_NS_F64),
    "learned": dict(precond_dtype=mgx.F32, gram_prec=mgx.PREC_SPLIT, inverse_mode=mgx.INV_NS_F64,
                    precond_kind=mgx.PRECOND_LEARNED),
    "ext16": dict(precond_dtype=mgx.F16, gram_prec=mgx.PREC_SPLIT, inverse_mode=mgx.INV_NS_BF16),
}
LABELS = {
    "plain": "plain FP64 (FP64 V-cycle, FP64 S/H, FP64 Newton-Schulz S^-1)",
    "mixed": "mixed (FP32 V-cycle, hi/lo-split S/H, FP64 Newton-Schulz / Cholesky S^-1)",
    "learned": "mixed+learned (BF16 CNN preconditioner with the FP32 V-cycle as judge and fallback, "
               "hi/lo-split S/H, FP64 S^-1)",
    "ext16": "extension, not the specified precisions: FP16-storage V-cycle, hi/lo-split S/H, "
             "S^-1 Newton-Schulz correction as one BF16 hi/lo product",
}
DTYPES = {"plain": "f64", "mixed": "f32+f64", "learned": "bf16+f32+f64", "ext16": "f16+bf16+f32+f64"}

# kernel groups timed by the solver: rocprof kernel name (C5 instantiation)
# and what bounds it; unit of the accumulated work: bytes ("hbm") or flops
GROUPS = {
    "hloc": ("k_box_march4<double, 2, 1, 512, ...> H psi (FP64, V gathered per box)", "hbm"),
    "cellgram": ("k_cell_gram<PREC, true, 1> S/H partials per 8^3 cell (MFMA)", "hbm"),
    "skin": ("k_slab_skin Galerkin H outside each box (the reach through each face of box j into boxes i < j)", "hbm"),
    "residk": ("k_cell_residual<true, OM> R = psi theta - H psi per 8^3 cell (FP64 MFMA), written as the "
               "V-cycle's zero start f, x (or as R on network-trial steps)", "hbm"),
    "update": ("k_box_and_update_nrm<T, 32> Anderson update + normalisation", "hbm"),
    "jacobi0": ("k_box_march4<T, 2, 2, ..., XT> fine-level Jacobi sweep: the pre-sweep with the restriction "
                "inside (XT=2), the first post-sweep with the prolongation inside (XT=3)", "hbm"),
    "jacobi0d": ("k_box_march4<T, 2, 2, ..., XT=1> last fine sweep with the Anderson dots", "hbm"),
    "first": ("k_first_sweep<T> zero-start first fine sweep (pointwise)", "hbm"),
    "transfer": ("k_box_restrict2p<T> + k_box_prolong8<T, true> fine-level transfers where the sweeps do not "
                 "hold them (2 kernels)", "hbm"),
    "spmm": ("k_spmm_tiles pair-sparse X S and S^-1 H", "hbm"),
    "nsgemm": ("rocBLAS dgemm (Cijk_..._DB_MT64x64x16) Newton-Schulz X <- 2X - (XS)X", "mfma:f64"),
    "nsgemm16": ("rocBLAS gemm_ex BF16 (Cijk_..._BSS_BH_...) Newton-Schulz correction, K = 3n", "mfma:bf16"),
    "learned": ("CNN inference (4 conv3d layers, BF16 MFMA) over all boxes", "mfma:bf16"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_config(name, seed):
    c = dict(synth.CONFIGS[name])
    if c.pop("kind") != "box":
        raise SystemExit("bench supports box workloads (c2, c5)")
    return synth.box_config(seed=seed, **c), c


def fit_precond(cfg):
    """FP16 storage needs box edges that are multiples of 8 on every level."""
    m = 8 << (cfg.levels - 1)
    if not all(int(e) % m == 0 for e in cfg.edge):
        PATHS["ext16"]["precond_dtype"] = mgx.F32
        LABELS["ext16"] = LABELS["ext16"].replace("FP16-storage V-cycle", "FP32 V-cycle (edges not multiples of 8)")


def run_mode(ctx, bs, cfg, mode, tol, maxit, steps, warmup, cnn, graph=True, nograph_steps=0):
    kw = dict(mg_levels=cfg.levels - 1, profile=0, precond_start=mgx.START_ZERO, anderson_m=1,
              conv_criterion=mgx.CONV_RESIDUAL)
    kw.update(PATHS[mode])
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, cnn=cnn if mode == "learned" else None, **kw)
    # the step's launch sequences replayed as HIP graphs (captured during the
    # warm-up); the same solves with eager launches are timed for comparison
    s.set_graph(graph)
    for _ in range(warmup):
        s.reset()
        s.run(tol, maxit)
    nograph = []
    if nograph_steps:
        s.set_graph(False)
        for k in range(nograph_steps):
            s.reset()
            t0 = time.perf_counter()
            st = s.run(tol, maxit)
            nograph.append(time.perf_counter() - t0)
            log("[%s] solve without graphs %d: %d iterations, %.3f s" % (mode, k, st.iterations, nograph[-1]))
        s.set_graph(graph)
    times, stats = [], None
    g0 = s.graph_stats()
    for k in range(steps):
        s.reset()
        t0 = time.perf_counter()
        stats = s.run(tol, maxit)
        times.append(time.perf_counter() - t0)
        log("[%s] solve %d: %d iterations, |R|=%.3e Ha, E=%.12f Ha, %.3f s (converged=%d, learned steps %d, "
            "fallbacks %d)" % (mode, k, stats.iterations, stats.residual, stats.band_energy, times[-1],
                               stats.converged, stats.learned_steps, stats.fallbacks))
    g1 = s.graph_stats()
    # one more solve with the kernel groups timed (hipEvents on the engine's stream)
    s.reset()
    s.set_profile(True)
    s.profile(reset=True)
    t0 = time.perf_counter()
    pst = s.run(tol, maxit)
    t_prof = time.perf_counter() - t0
    prof = s.profile()
    log("[%s] profiled solve: %d iterations, %.3f s" % (mode, pst.iterations, t_prof))
    s.close()
    return dict(seconds=float(np.mean(times)), seconds_min=float(np.min(times)), iterations=stats.iterations,
                residual=stats.residual, energy=stats.band_energy, converged=bool(stats.converged),
                learned_steps=stats.learned_steps, fallbacks=stats.fallbacks, newton_schulz_steps=stats.ns_steps,
                cholesky_calls=stats.chol_calls, ns_continued=stats.ns_continued, ns_restarts=stats.ns_restarts,
                dots_sweeps=stats.dots_sweeps, profiled_seconds=t_prof, profiled_iterations=pst.iterations,
                profile=prof, graph=bool(graph),
                graph_launches_per_solve=(g1[0] - g0[0]) / max(steps, 1),
                graph_captures_in_timed_solves=g1[1] - g0[1], graphs_held=g1[2],
                seconds_without_graphs=float(np.mean(nograph)) if nograph else None)


def group_rates(r, mode, npts, gram_waves=None):
    """Per kernel group of one profiled solve: ms per solve, launches, average
    launch, algorithmic work per launch and the rate against its peak; the
    PMC traffic of the same kernel where a pass measured it (the cell Gram's
    for the waves per cell the autotuner chose)."""
    out = {}
    prec = {"plain": "f64", "mixed": "f32", "learned": "f32", "ext16": "f16"}[mode]
    gprec = "split" if PATHS[mode]["gram_prec"] == mgx.PREC_SPLIT else "f64"
    for g, (name, bound) in GROUPS.items():
        ms, launches, work = r["profile"][g]
        if not launches:
            continue
        avg_s = ms / 1e3 / launches
        e = dict(kernel=name.replace("<T", "<" + {"f64": "double", "f32": "float", "f16": "_Float16"}[prec])
                 .replace("<PREC", "<" + ("1" if gprec == "split" else "0")),
                 group=g, ms_per_solve=round(ms, 3), launches=int(launches), avg_launch_ms=round(1e3 * avg_s, 4))
        if work > 0 and bound == "hbm":
            per = work / launches
            ach = per / avg_s / 1e9
            key = (g, {"cellgram": gprec, "learned": "bf16"}.get(g, prec))
            if g == "cellgram":
                key = (g, gprec, gram_waves)
            e.update(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=round(ach / HBM_PEAK_GBS, 4), bytes_per_launch=int(per),
                     traffic=TRAFFIC_C5.get(key) if npts == C5_POINTS else None)
        elif work > 0 and bound.startswith("mfma"):
            dt = bound.split(":")[1]
            per = work / launches
            ach = per / avg_s / 1e12
            e.update(bound="mfma", achieved=round(ach, 2), peak=PEAK_TFS[dt], unit="TFLOP/s",
                     frac=round(ach / PEAK_TFS[dt], 4), flop_per_launch=int(per), mfma_dtype=dt,
                     traffic=TRAFFIC_C5.get((g, dt)) if npts == C5_POINTS else None)
        else:
            e.update(bound=bound, achieved=None, frac=None, traffic=None)
        if e.get("traffic"):
            e["traffic_source"] = TRAFFIC_SOURCE
        out[g] = e
    return out


def stream_reference(ctx, n=1 << 28, reps=20):
    """Measured HBM ceiling: STREAM triad a = b + s c over three 1 GiB FP32
    arrays, 16-byte lanes, 2/4/8 loads in flight per lane (best depth)."""
    import ctypes as C
    arrs = [mgx.DeviceArray(ctx, shape=(n,), dtype=np.float32) for _ in range(3)]
    gbps, unroll = C.c_double(), C.c_int()
    ctx.check(mgx.lib().mgx_stream_triad(ctx.h, C.c_int64(n), mgx.ptr(arrs[0]), mgx.ptr(arrs[1]), mgx.ptr(arrs[2]),
                                         reps, C.byref(gbps), C.byref(unroll)), "mgx_stream_triad")
    for d in arrs:
        d.free()
    return dict(kernel="k_triad4<%d> (a = b + s c, 3 x 1 GiB FP32, 16-byte lanes, %d loads per operand in flight)"
                % (unroll.value, unroll.value), GBps=round(gbps.value, 1),
                frac_of_peak=round(gbps.value / HBM_PEAK_GBS, 4))


def host_cpu():
    """CPU model, logical CPUs, physical cores (from /proc/cpuinfo) and the
    CPUs this process may run on."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return dict(model=model, logical_cpus=os.cpu_count(), physical_cores=len(cores) or None, cpus_available=avail)


def cpu_baseline(cfg, iters, mg_levels, budget_s=12.0):
    """Full-size orbital steps of the oracle (FP32 V-cycle, the reference's
    MGPRECONDTYPE; C++/OpenMP) on this host until about budget_s seconds
    have passed (at least 2 steps), scaled by the iteration count."""
    import oracle as O
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    psi = cfg.psi.copy()
    ts = []
    while len(ts) < 2 or sum(ts) < budget_s:
        t0 = time.perf_counter()
        psi, _, _, _ = O.b_solver_step(ob, psi, cfg.vtot, mg_levels, O.F32, 1.5, None, zero_start=True)
        ts.append(time.perf_counter() - t0)
    t_step = float(np.mean(ts))
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or host_cpu()["cpus_available"]
    cpu = host_cpu()
    return dict(value=round(t_step * iters, 2), unit="s per converged solve", cores=threads, kind="port",
                sample="%d oracle orbital steps on the full workload (%.2f s each, %d OpenMP threads), "
                       "scaled by %d iterations" % (len(ts), t_step, threads, iters),
                cpu_model=cpu["model"], physical_cores=cpu["physical_cores"], logical_cpus=cpu["logical_cpus"],
                cpus_available=cpu["cpus_available"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed converged solves per path")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=2000)
    ap.add_argument("--modes", default="plain,mixed,learned,ext16")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch every kernel eagerly")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    modes = args.modes.split(",")

    t0 = time.perf_counter()
    cfg, meta = make_config(args.config, args.seed)
    fit_precond(cfg)
    log("config %s: %d boxes, %d points, grid %s, h=%.3f (generated in %.1f s)"
        % (args.config, cfg.nbox, cfg.psi.size, cfg.gdim, cfg.h, time.perf_counter() - t0))
    ctx = mgx.Ctx(0)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, cfg.levels)
    log("box pairs:", bs.npairs())
    # launch shapes of the fine-level kernels for this box set (the solvers use them)
    tune = bs.autotune()
    log("autotune:", tune)
    cnn, wmeta = None, None
    if "learned" in modes:
        from mgmol_amd import learned, train_cnn
        params, wmeta = train_cnn.load(args.weights)
        cnn = learned.Cnn(ctx, wmeta["edge"], 256, params)
    res = {}
    headname = next(m for m in ("learned", "mixed", "ext16", "plain") if m in modes)
    for mode in modes:
        res[mode] = run_mode(ctx, bs, cfg, mode, args.tol, args.maxit, args.steps, args.warmup, cnn,
                             graph=not args.no_graph, nograph_steps=args.steps if mode == headname else 0)
    head = res[headname]
    npts = cfg.psi.size

    stream = stream_reference(ctx)
    kernels = {m: group_rates(r, m, npts, tune["gram_split" if PATHS[m]["gram_prec"] == mgx.PREC_SPLIT
                                                else "gram_f64"][0]) for m, r in res.items()}
    for ks in kernels.values():
        for e in ks.values():
            if e.get("unit") == "GB/s":
                e["frac_of_stream"] = round(e["achieved"] / stream["GBps"], 4)
    # the headline path's kernel groups by time per solve; the top one with a
    # roofline is `roofline`, the others above 5% of a solve follow it
    ranked = sorted(kernels[headname].values(), key=lambda e: -e["ms_per_solve"])
    solve_ms = 1e3 * head["profiled_seconds"]
    for e in ranked:
        e["share_of_solve"] = round(e["ms_per_solve"] / solve_ms, 4)
    with_roof = [e for e in ranked if e.get("frac") is not None]
    roof = with_roof[0] if with_roof else None
    other = [e for e in ranked if e is not roof and e["share_of_solve"] > 0.05]

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
        "dtype": DTYPES[headname],
        "data": "synthetic (seeded Gaussian orbitals and wells)",
        "config": {"workload": "%s: %d orbitals in %s^3 boxes, %d^3 periodic grid, h=%.2f bohr, %d wells, "
                               "centres >= %s bohr apart, ABPG (Anderson m=1) to |R| < %g Ha"
                               % (args.config, cfg.nbox, "/".join(str(e) for e in meta["edges"]), cfg.gdim[0],
                                  cfg.h, meta["nwells"], meta.get("min_sep"), args.tol),
                   "path": LABELS[headname]},
        "iterations": head["iterations"],
        "converged": head["converged"],
        "paths": {m: dict({k: v for k, v in r.items() if k != "profile"}, label=LABELS[m], dtype=DTYPES[m])
                  for m, r in res.items()},
        "roofline": roof,
        "roofline_other": other,
        "kernels": {m: sorted(ks.values(), key=lambda e: -e["ms_per_solve"]) for m, ks in kernels.items()},
        "phase_ms_per_solve": {m: {g: round(r["profile"][g][0], 2) for g in ("hloc", "gram", "dense", "resid",
                                                                             "precond", "normalize", "learned")}
                               for m, r in res.items()},
        "stream": stream,
        "autotune": {k: {"choice": c, "ms_per_launch": [round(t1, 4), round(t2, 4)],
                         "candidates": ("waves per cell 1, 2" if k.startswith("gram")
                                        else "restriction inside the FP32 fine pre-sweep / separate kernel"
                                        if k == "restrict"
                                        else "prolongation inside the FP32 fine post-sweep / separate kernel"
                                        if k == "prolong" else "lane groups per thread 1, 2")}
                     for k, (c, t1, t2) in tune.items()},
        "cpu_baseline": cpu,
    }
    out["hipgraph"] = {"path": headname, "seconds_graph_on": round(head["seconds"], 4),
                       "seconds_graph_off": (round(head["seconds_without_graphs"], 4)
                                             if head["seconds_without_graphs"] else None),
                       "graph_launches_per_solve": head["graph_launches_per_solve"],
                       "graphs_held": head["graphs_held"],
                       "captures_in_timed_solves": head["graph_captures_in_timed_solves"]}
    if roof and roof.get("unit") == "GB/s":
        roof["frac_of_stream"] = round(roof["achieved"] / stream["GBps"], 4)
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
