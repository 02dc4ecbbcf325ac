"""The box set's autotuner (mgx_boxset_autotune) and every launch shape it
can choose.

* The tuner times both candidates of every tunable on the box set's layout
  and keeps the faster one (times reported per candidate).
* Each shape gives the oracle's results: the fine Jacobi sweep (FP32, FP64),
  H psi (FP64) and the FP16 V-cycle bit-identical between one and two lane
  groups per thread (the arithmetic per point does not depend on the lane
  mapping); S / H with one or two waves per cell within the oracle's
  tolerances and bit-identical to each other (the cell's rows are summed in
  two fixed halves whatever the number of waves).
* The V-cycle's transfers inside the sweeps against the separate kernels:
  the restriction inside the last pre-sweep (FP32 and FP16 storage
  bit-identical, FP64 to the last bits: the 27 FP64 terms are summed in
  another order) and the prolongation inside the first post-sweep
  (bit-identical in every precision), all within the oracle's tolerances,
  on levels where the fused forms apply and where they do not.
* A solver on a box set that was not tuned tunes it at creation, and its
  steps do not depend on the lane mappings (the same energies bit for bit).
"""
import numpy as np
import pytest

import oracle as O
from mgmol_amd import mgx, synth
from tests.gpu_util import relmax, say, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


def _cfg():
    # 32^3 boxes (all tunables apply: 256 / 512 lane groups of the fine sweeps)
    return synth.box_config(nbox=12, gdim=72, h=0.3, edges=(32,), nwells=4, seed=17, levels=3)


def test_autotune_reports_and_chooses(ctx):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        assert bs.get_tune()[1] is False
        rep = bs.autotune()
        say("autotune", rep)
        for k, (choice, t1, t2) in rep.items():
            assert choice in (1, 2)
            assert t1 > 0 and t2 > 0
            assert (choice == 1) == (t1 <= t2)
        assert bs.get_tune()[1] is True
    finally:
        bs.close()


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_march_shapes_match_oracle(ctx, dt):
    cfg = _cfg()
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    DT = np.float32 if dt == "f32" else np.float64
    code = mgx.F32 if dt == "f32" else mgx.F64
    key = "march_" + dt
    try:
        rng = np.random.default_rng(3)
        f = rng.standard_normal(ob.size(0)).astype(DT)
        x = rng.standard_normal(ob.size(0)).astype(DT)
        jf = O.jacobi_factor(2, list(ob.h))
        xr, wr = x.copy(), np.zeros_like(x)
        O.b_jacobi(ob, 0, 2, f, xr, wr, jf)
        outs = []
        for p in (1, 2):
            bs.set_tune({key: p, "hloc_f64": p})
            tf, tx, tw = to_dev(f), to_dev(x), to_dev(np.zeros_like(x))
            ctx.call("mgx_box_jacobi", bs.h, 0, 2, code, jf, mgx.ptr(tf), mgx.ptr(tx), mgx.ptr(tw))
            outs.append((to_np(tx), to_np(tw)))
            say("jacobi", dt, "groups per thread", p, relmax(outs[-1][0], xr))
            assert relmax(outs[-1][0], xr) <= (1e-6 if dt == "f32" else 1e-13)
        assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
        if dt == "f64":
            ref = O.b_hloc(ob, cfg.vtot, cfg.psi)
            hs = []
            for p in (1, 2):
                bs.set_tune({"hloc_f64": p})
                tp, tv, th = to_dev(cfg.psi), to_dev(cfg.vtot), to_dev(np.zeros_like(cfg.psi))
                ctx.call("mgx_box_hloc", bs.h, mgx.F64, mgx.ptr(tp), mgx.ptr(tv), mgx.ptr(th))
                hs.append(to_np(th))
                assert relmax(hs[-1], ref) <= 1e-13
            assert np.array_equal(hs[0], hs[1])
    finally:
        bs.close()


def test_f16_vcycle_shapes_identical(ctx):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        r = np.random.default_rng(4).standard_normal(int(cfg.psi.size))
        outs = []
        for p in (1, 2):
            bs.set_tune({"march_f16": p})
            tr, to = to_dev(r), to_dev(np.zeros_like(r))
            ctx.call("mgx_box_precond_mg", bs.h, 2, 2, 0.0, mgx.F16, mgx.ptr(tr), mgx.ptr(to))
            outs.append(to_np(to))
        assert np.array_equal(outs[0], outs[1])
    finally:
        bs.close()


@pytest.mark.parametrize("edges,gdim", [((32,), 72), ((8, 12, 16, 20, 24), 56)])
@pytest.mark.parametrize("start", ["zero", "gamma"])
def test_fused_transfers_match_separate(ctx, edges, gdim, start):
    # (8 .. 24): at level 1 the edges 6, 10 are not multiples of 4, so that
    # level keeps the separate restriction under either choice
    cfg = synth.box_config(nbox=12, gdim=gdim, h=0.3, edges=edges, nwells=4, seed=23, levels=3)
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        r = np.random.default_rng(6).standard_normal(ob.size(0))
        gam = 0.0 if start == "zero" else O.gamma(list(ob.h), 2, 0.0, -1.2)
        # the FP16 V-cycle needs edges that are multiples of 8 at every level
        kinds = ((mgx.F16, "f16"),) if edges == (32,) else ()
        for code, name in kinds + ((mgx.F32, "f32"), (mgx.F64, "f64")):
            outs = {}
            for rs in (1, 2):
                for pr in (1, 2):
                    bs.set_tune({"restrict": rs, "prolong": pr})
                    tr, to = to_dev(r), to_dev(np.zeros_like(r))
                    ctx.call("mgx_box_precond_mg", bs.h, 2, 2, gam, code, mgx.ptr(tr), mgx.ptr(to))
                    outs[rs, pr] = to_np(to)
            sep = outs[2, 2]
            for rs in (1, 2):
                # the prolongation's choice changes nothing
                assert np.array_equal(outs[rs, 1], outs[rs, 2])
            d = relmax(outs[1, 2], sep)
            say("transfers fused / separate", edges, start, name, d)
            if name == "f64":
                assert d <= 1e-14
            else:
                assert np.array_equal(outs[1, 2], sep)
            if name != "f16":
                ref = O.b_precond(ob, 2, 2, gam, r, dtype=np.float64 if name == "f64" else np.float32)
                for o in outs.values():
                    assert relmax(o, ref) <= (1e-12 if name == "f64" else 1e-5)
    finally:
        bs.close()


@pytest.mark.parametrize("prec", [mgx.PREC_F64, mgx.PREC_SPLIT])
def test_gram_waves_match_oracle(ctx, prec):
    cfg = _cfg()
    ob = O.Boxes.make(cfg.origin, cfg.edge, cfg.gdim, cfg.h)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    key = "gram_f64" if prec == mgx.PREC_F64 else "gram_split"
    try:
        hp = O.b_hloc(ob, cfg.vtot, cfg.psi)
        S, H = O.b_gram_h(ob, cfg.psi, hp)
        n = ob.nbox
        res = []
        for nw in (1, 2):
            bs.set_tune({key: nw})
            tp, th = to_dev(cfg.psi), to_dev(hp)
            tS, tH = to_dev(np.zeros((n, n))), to_dev(np.zeros((n, n)))
            ctx.call("mgx_box_gram_h", bs.h, prec, mgx.ptr(tp), mgx.ptr(th), mgx.ptr(tS), mgx.ptr(tH))
            es, eh = relmax(to_np(tS), S), relmax(to_np(tH), H)
            say("gram", prec, "waves", nw, es, eh)
            tol = 1e-13 if prec == mgx.PREC_F64 else 1e-7
            assert es <= tol and eh <= tol
            res.append((to_np(tS), to_np(tH)))
        assert np.array_equal(res[1][0], res[0][0]) and np.array_equal(res[1][1], res[0][1])
    finally:
        bs.close()


def test_solver_tunes_its_box_set(ctx):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, anderson_m=1)
        choice, tuned = bs.get_tune()
        say("solver's box set tuned:", choice)
        assert tuned
        e = [s.step().band_energy for _ in range(6)]
        s.close()
        # the other lane mappings and prolongations give the same steps bit
        # for bit (the restriction's choice, which moves FP64 sums by
        # rounding, is kept)
        other = {k: (v if k == "restrict" else 3 - v) for k, v in choice.items()}
        bs.set_tune(other)
        s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, anderson_m=1)
        assert bs.get_tune()[0] == other
        e2 = [s.step().band_energy for _ in range(6)]
        s.close()
        assert e == e2
    finally:
        bs.close()
