"""The solver's steps replayed as HIP graphs (mgx_solver_set_graph) against
the same steps launched eagerly, from the same orbitals.

Every step's band energy and residual norm must be bit-identical: a graph
replays the kernels its capture recorded, with the same arguments, in the
same order.  Each graph is replayed many times, with the host's own work and
a stream wait (the energy and norm read) between replays, and each replay is
checked.  Paths: plain FP64, the mixed path (FP32 V-cycle, split S/H), the
FP16/BF16 extension, and the learned path (whose network trials run
eagerly between replays).  A workspace that moves between replays (a larger
call on the same context) must make the solver capture again.
"""
import numpy as np
import pytest

from mgmol_amd import learned, mgx, synth
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tests.gpu_util import ctx as shared
    return shared()


PATHS = {
    "plain": dict(precond_dtype=mgx.F64, gram_prec=mgx.PREC_F64, inverse_mode=mgx.INV_NS_F64),
    "mixed": dict(precond_dtype=mgx.F32, gram_prec=mgx.PREC_SPLIT, inverse_mode=mgx.INV_NS_F64),
    "ext16": dict(precond_dtype=mgx.F16, gram_prec=mgx.PREC_SPLIT, inverse_mode=mgx.INV_NS_BF16),
}


def _cfg():
    # 32^3 boxes: edges that stay multiples of 8 on the three levels (FP16 V-cycle)
    return synth.box_config(nbox=12, gdim=72, h=0.3, edges=(32,), nwells=4, seed=13, levels=3)


def _run(ctx, bs, cfg, graph, nsteps, **kw):
    s = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, anderson_m=1, **kw)
    s.set_graph(graph)
    out = []
    for _ in range(nsteps):
        st = s.step()
        out.append((st.band_energy, st.residual, st.ns_steps, st.chol_calls, st.dots_sweeps))
    return s, out


@pytest.mark.parametrize("path", list(PATHS))
def test_graph_steps_equal_eager(ctx, path):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        s0, eager = _run(ctx, bs, cfg, False, 40, **PATHS[path])
        s1, graph = _run(ctx, bs, cfg, True, 40, **PATHS[path])
        launches, captures, held = s1.graph_stats()
        say(path, "graph launches %d captures %d held %d" % (launches, captures, held), eager[-1], graph[-1])
        assert s0.graph_stats()[0] == 0
        assert captures >= 2 and launches >= 40
        for k, (a, b) in enumerate(zip(eager, graph)):
            assert a == b, "step %d: eager %r graph %r" % (k, a, b)
        assert np.array_equal(s0.psi(), s1.psi())
        s0.close()
        s1.close()
    finally:
        bs.close()


def test_graph_learned_path(ctx):
    """Network trials (eager, with their saves and restores) between graph
    replays of the V-cycle steps."""
    cfg = synth.box_config(nbox=16, gdim=48, h=0.3, edges=(16,), nwells=4, seed=11, levels=3)
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    net = learned.Cnn(ctx, 16, 8, learned.init_params(seed=0, skip=0.5))
    try:
        kw = dict(precond_kind=mgx.PRECOND_LEARNED, cnn=net, **PATHS["mixed"])
        s0, eager = _run(ctx, bs, cfg, False, 30, **kw)
        s1, graph = _run(ctx, bs, cfg, True, 30, **kw)
        st = s1.step()
        say("learned: trials %d, graph stats" % st.learned_steps, s1.graph_stats())
        assert st.learned_steps >= 2
        assert s1.graph_stats()[0] >= 10
        for k, (a, b) in enumerate(zip(eager, graph)):
            assert a == b, "step %d: eager %r graph %r" % (k, a, b)
        s0.close()
        s1.close()
    finally:
        net.close()
        bs.close()


def test_graph_recaptured_after_workspace_moves(ctx):
    cfg = _cfg()
    bs = mgx.BoxSet(ctx, cfg.origin, cfg.edge, cfg.gdim, cfg.h, 3)
    try:
        s0, eager = _run(ctx, bs, cfg, False, 30, **PATHS["ext16"])
        s1 = mgx.Solver(ctx, bs, cfg.vtot, cfg.psi, mg_levels=2, anderson_m=1, **PATHS["ext16"])
        s1.set_graph(True)
        out = []
        caps = None
        for k in range(30):
            if k == 15:
                caps = s1.graph_stats()[1]
                # a larger product on the same context moves its workspaces
                n, ld = 200, 1 << 17
                tp = to_dev(np.random.default_rng(1).standard_normal((n, ld)))
                ts = to_dev(np.zeros((1, n, n)))
                ctx.call("mgx_local_gram", mgx.PREC_F64, mgx.ptr(tp), ld, ld, n, 1, 1.0, mgx.ptr(ts))
            st = s1.step()
            out.append((st.band_energy, st.residual, st.ns_steps, st.chol_calls, st.dots_sweeps))
        say("recapture: captures before %d, after %d" % (caps, s1.graph_stats()[1]))
        assert s1.graph_stats()[1] > caps
        for k, (a, b) in enumerate(zip(eager, out)):
            assert a == b, "step %d: eager %r graph %r" % (k, a, b)
        s0.close()
        s1.close()
    finally:
        bs.close()
