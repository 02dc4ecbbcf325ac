"""ctypes binding of libmgx.so (include/mgx.h).

Device memory is plain pointers, allocated and copied through the C ABI
itself (``DeviceArray``: mgx_malloc / mgx_memcpy_*), so a process loads one
HIP runtime only.  There is no CPU path: if ``libmgx.so`` is missing or
fails to load, ``lib()`` raises.
"""
import ctypes as C
import os

from . import LIB_PATH

MGX_OK, MGX_EINVAL, MGX_EHIP, MGX_ENOTSPD, MGX_ENUMERIC, MGX_ENOMEM = 0, -1, -2, -3, -4, -5
F32, F64 = 0, 1
F16 = 2  # preconditioner storage only: FP16 V-cycle, FP32 correction
LAP_MEHRSTELLEN, LAP_2ND, LAP_4TH = 0, 1, 2
PREC_F64, PREC_SPLIT = 0, 1
START_GAMMA, START_ZERO = 0, 1
CONV_DELTAE, CONV_RESIDUAL = 0, 1
PRECOND_VCYCLE, PRECOND_LEARNED = 0, 1
INV_NS_F64, INV_NS_BF16, INV_CHOL = 0, 1, 2
PROF_GROUPS = ["hloc", "gram", "dense", "resid", "jacobi0", "precond", "normalize", "learned", "cellgram",
               "skin", "residk", "update", "jacobi0d", "first", "nsgemm", "spmm", "transfer", "nsgemm16"]

_LIB = None

# every symbol include/mgx.h declares
EXPORTS = [
    "mgx_ctx_create", "mgx_ctx_destroy", "mgx_last_error", "mgx_ctx_stream",
    "mgx_malloc", "mgx_free", "mgx_memcpy_htod", "mgx_memcpy_dtoh",
    "mgx_halo", "mgx_lap_apply", "mgx_rhs_apply", "mgx_pointwise_product", "mgx_axpy", "mgx_stream_triad", "mgx_synth_separable",
    "mgx_jacobi",
    "mgx_restrict", "mgx_prolong", "mgx_mask", "mgx_mg",
    "mgx_hloc_apply", "mgx_precond_mg", "mgx_gamma", "mgx_poisson_mg", "mgx_pb_mg",
    "mgx_local_gram", "mgx_local_gram_noalloc", "mgx_local_product", "mgx_local_to_global", "mgx_chol_inv", "mgx_theta", "mgx_eigsum",
    "mgx_ritz", "mgx_residual", "mgx_residual_norm", "mgx_mehr_bmatrix", "mgx_mehr_projected", "mgx_mehr_residual",
    "mgx_boxset_create", "mgx_boxset_destroy", "mgx_boxset_autotune", "mgx_boxset_set_tune", "mgx_boxset_get_tune", "mgx_boxset_size", "mgx_boxset_offset", "mgx_boxset_npairs",
    "mgx_box_lap", "mgx_box_hloc", "mgx_box_jacobi", "mgx_box_restrict", "mgx_box_prolong", "mgx_box_precond_mg",
    "mgx_box_gram_h", "mgx_box_residual", "mgx_box_residual_norm", "mgx_box_density",
    "mgx_box_hloc_mehr", "mgx_box_gram_hb_mehr", "mgx_box_residual_mehr",
    "mgx_solver_create", "mgx_solver_destroy", "mgx_solver_step", "mgx_solver_run", "mgx_solver_reset",
    "mgx_solver_profile", "mgx_solver_profile_work", "mgx_solver_set_profile",
    "mgx_solver_get_psi", "mgx_solver_buffers", "mgx_solver_set_keep_residual", "mgx_solver_inverse", "mgx_solver_set_psi", "mgx_solver_set_cnn", "mgx_solver_density",
    "mgx_solver_set_graph", "mgx_solver_graph_stats",
    "mgx_cnn_nparams", "mgx_cnn_create", "mgx_cnn_destroy", "mgx_cnn_set_params", "mgx_cnn_get_params",
    "mgx_cnn_forward", "mgx_cnn_grad", "mgx_cnn_train_step", "mgx_cnn_apply", "mgx_cnn_make_pairs",
]


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int * 3), ("ghosts", C.c_int), ("h", C.c_double * 3), ("bc", C.c_int * 3)]

    @staticmethod
    def make(dim, ghosts, h, bc=(1, 1, 1)):
        g = Grid()
        for i in range(3):
            g.dim[i] = int(dim[i])
            g.h[i] = float(h[i])
            g.bc[i] = int(bc[i])
        g.ghosts = int(ghosts)
        return g


class Block(C.Structure):
    _fields_ = [("data", C.c_void_p), ("dtype", C.c_int), ("nfunc", C.c_int), ("grid", Grid)]


class Boxes(C.Structure):
    _fields_ = [("nbox", C.c_int), ("gdim", C.c_int * 3), ("h", C.c_double * 3),
                ("origin", C.POINTER(C.c_int)), ("edge", C.POINTER(C.c_int))]


class SolverCfg(C.Structure):
    _fields_ = [("lap_type", C.c_int), ("mg_levels", C.c_int), ("precond_dtype", C.c_int),
                ("gram_prec", C.c_int), ("step_length", C.c_double), ("ritz_every", C.c_int),
                ("precond_start", C.c_int), ("profile", C.c_int), ("anderson_m", C.c_int),
                ("anderson_beta", C.c_double), ("conv_criterion", C.c_int), ("precond_kind", C.c_int),
                ("inverse_mode", C.c_int), ("ns_tol", C.c_double), ("ns_max_spec", C.c_int)]


class SolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("residual", C.c_double),
                ("band_energy", C.c_double), ("seconds", C.c_double), ("lowest_eig", C.c_double),
                ("learned_steps", C.c_int), ("fallbacks", C.c_int), ("ns_steps", C.c_int), ("chol_calls", C.c_int),
                ("dots_sweeps", C.c_int), ("ns_continued", C.c_int), ("ns_restarts", C.c_int)]


class PoissonStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("nb_sweeps", C.c_int), ("final_residual", C.c_double),
                ("final_relative_residual", C.c_double), ("residual_reduction", C.c_double)]


class MgxError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MgxError("libmgx.so not built (%s): run `python mgmol_amd/build.py`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P, I, I64, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
        PP = C.POINTER(C.c_void_p)
        BP = C.POINTER(Block)
        GP = C.POINTER(Grid)
        sig = {
            "mgx_ctx_create": (I, C.POINTER(P)), "mgx_ctx_destroy": (P,), "mgx_ctx_stream": (P,),
            "mgx_malloc": (P, C.c_size_t, C.POINTER(P)), "mgx_free": (P, P),
            "mgx_memcpy_htod": (P, P, P, C.c_size_t), "mgx_memcpy_dtoh": (P, P, P, C.c_size_t),
            "mgx_halo": (P, BP), "mgx_lap_apply": (P, I, BP, BP), "mgx_rhs_apply": (P, BP, BP),
            "mgx_pointwise_product": (P, BP, P, BP), "mgx_axpy": (P, I, I64, D, P, P),
            "mgx_synth_separable": (P, I, C.POINTER(I), I, I, P, P, P, P, P, P, P, P, P),
            "mgx_stream_triad": (P, I64, P, P, P, I, C.POINTER(D), C.POINTER(I)),
            "mgx_jacobi": (P, I, D, BP, BP, BP), "mgx_restrict": (P, BP, BP), "mgx_prolong": (P, BP, BP),
            "mgx_mask": (P, BP, P), "mgx_mg": (P, I, I, BP, BP, PP),
            "mgx_poisson_mg": (P, I, BP, BP, I, I, I, D, I, C.POINTER(PoissonStats)),
            "mgx_pb_mg": (P, I, BP, BP, BP, BP, D, D, D, I, I, I, D, I, C.POINTER(PoissonStats)),
            "mgx_hloc_apply": (P, I, GP, I, P, I64, I, P, P, I64),
            "mgx_precond_mg": (P, I, I, D, I, GP, P, I64, I, PP, P, I64),
            "mgx_local_gram": (P, I, P, I64, I64, I, I, D, P),
            "mgx_local_gram_noalloc": (P, I, P, I64, I64, I, I, D, P),
            "mgx_local_product": (P, I, P, P, I64, I64, I, I, D, P),
            "mgx_local_to_global": (P, P, P, I, I, I, D, P),
            "mgx_chol_inv": (P, I, P, P, P), "mgx_theta": (P, I, P, P, P),
            "mgx_eigsum": (P, I, P, P, C.POINTER(D)), "mgx_ritz": (P, I, P, P, P, P),
            "mgx_residual": (P, P, P, P, I64, I64, I, P),
            "mgx_residual_norm": (P, P, I64, I64, I, D, P, C.POINTER(D)),
            "mgx_mehr_bmatrix": (P, GP, P, I64, I, P),
            "mgx_mehr_projected": (P, I, P, P, P, P, P, P, P, C.POINTER(D)),
            "mgx_mehr_residual": (P, GP, P, P, P, I64, I, P),
            "mgx_boxset_create": (P, C.POINTER(Boxes), I, C.POINTER(P)), "mgx_boxset_destroy": (P,),
            "mgx_boxset_size": (P, I, C.POINTER(I64)), "mgx_boxset_offset": (P, I, I, C.POINTER(I64)),
            "mgx_boxset_npairs": (P, C.POINTER(I64)),
            "mgx_boxset_autotune": (P, P, C.POINTER(I), C.POINTER(D)),
            "mgx_boxset_set_tune": (P, C.POINTER(I)), "mgx_boxset_get_tune": (P, C.POINTER(I), C.POINTER(I)),
            "mgx_box_lap": (P, P, I, I, I, P, P), "mgx_box_hloc": (P, P, I, P, P, P),
            "mgx_box_jacobi": (P, P, I, I, I, D, P, P, P), "mgx_box_restrict": (P, P, I, I, P, P),
            "mgx_box_prolong": (P, P, I, I, P, P), "mgx_box_precond_mg": (P, P, I, I, D, I, P, P),
            "mgx_box_gram_h": (P, P, I, P, P, P, P), "mgx_box_residual": (P, P, P, P, P, P),
            "mgx_box_hloc_mehr": (P, P, P, P, P), "mgx_box_gram_hb_mehr": (P, P, P, P, P, P, P, P),
            "mgx_box_residual_mehr": (P, P, P, P, P, P),
            "mgx_box_residual_norm": (P, P, P, P, C.POINTER(D)),
            "mgx_box_density": (P, P, P, P, I, D, P, C.POINTER(D)),
            "mgx_solver_create": (P, P, C.POINTER(SolverCfg), P, P, C.POINTER(P)),
            "mgx_solver_destroy": (P,), "mgx_solver_step": (P, C.POINTER(SolverStats)),
            "mgx_solver_run": (P, D, I, C.POINTER(SolverStats)), "mgx_solver_get_psi": (P, P),
            "mgx_solver_reset": (P,), "mgx_solver_profile": (P, C.POINTER(D), C.POINTER(I64), I),
            "mgx_solver_profile_work": (P, C.POINTER(D)), "mgx_solver_set_profile": (P, I),
            "mgx_solver_buffers": (P, PP, PP, PP, PP, PP), "mgx_solver_inverse": (P, PP),
            "mgx_solver_set_keep_residual": (P, I),
            "mgx_solver_set_psi": (P, P), "mgx_solver_set_cnn": (P, P),
            "mgx_solver_set_graph": (P, I),
            "mgx_solver_graph_stats": (P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64)),
            "mgx_solver_density": (P, D, P, C.POINTER(D)),
            "mgx_cnn_create": (P, I, I, C.POINTER(P)), "mgx_cnn_destroy": (P,),
            "mgx_cnn_set_params": (P, P), "mgx_cnn_get_params": (P, P), "mgx_cnn_forward": (P, P, I, P),
            "mgx_cnn_grad": (P, P, P, I, C.POINTER(D), P), "mgx_cnn_train_step": (P, P, P, I, C.c_float, C.POINTER(D)),
            "mgx_cnn_apply": (P, I, P, P), "mgx_cnn_make_pairs": (P, I, P, P, P, P),
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = list(args)
            fn.restype = I
        L.mgx_ctx_stream.restype = P
        L.mgx_last_error.argtypes = [P]
        L.mgx_last_error.restype = C.c_char_p
        L.mgx_gamma.argtypes = [C.POINTER(D), I, D, D]
        L.mgx_gamma.restype = D
        L.mgx_cnn_nparams.argtypes = []
        L.mgx_cnn_nparams.restype = I
        _LIB = L
    return _LIB


def ptr(t):
    """device pointer of a DeviceArray (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


class Ctx:
    """One mgx context (stream + rocBLAS handle) on a device."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().mgx_ctx_create(device, C.byref(self.h))
        if rc != MGX_OK:
            raise MgxError("mgx_ctx_create failed (%d)" % rc)

    def check(self, rc, what=""):
        if rc != MGX_OK:
            msg = lib().mgx_last_error(self.h)
            raise MgxError("%s failed (rc=%d): %s" % (what, rc, msg.decode() if msg else ""))
        return rc

    def call(self, name, *args):
        return self.check(getattr(lib(), name)(self.h, *args), name)

    def close(self):
        if self.h:
            lib().mgx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dtype_code(t):
    import numpy as np
    if t.dtype == np.float32:
        return F32
    if t.dtype == np.float64:
        return F64
    raise TypeError(t.dtype)


class DeviceArray:
    """A device buffer owned through the C ABI (mgx_malloc / mgx_free)."""

    def __init__(self, ctx, src=None, shape=None, dtype=None, zero=True):
        import numpy as np
        self.ctx = ctx
        if src is not None:
            src = np.ascontiguousarray(src)
            shape, dtype = src.shape, src.dtype
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._p = C.c_void_p()
        ctx.check(lib().mgx_malloc(ctx.h, self.nbytes, C.byref(self._p)), "mgx_malloc")
        if src is not None:
            self.set(src)
        elif self.nbytes and zero:
            self.set(np.zeros(self.shape, self.dtype))

    def data_ptr(self):
        return self._p.value

    def set(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            self.ctx.check(lib().mgx_memcpy_htod(self.ctx.h, self._p, a.ctypes.data_as(C.c_void_p), self.nbytes),
                           "mgx_memcpy_htod")

    def get(self):
        import numpy as np
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            self.ctx.check(lib().mgx_memcpy_dtoh(self.ctx.h, out.ctypes.data_as(C.c_void_p), self._p, self.nbytes),
                           "mgx_memcpy_dtoh")
        return out

    def free(self):
        if self._p and self._p.value and self.ctx.h:
            lib().mgx_free(self.ctx.h, self._p)
        self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def block(t, grid, nfunc=None):
    b = Block()
    b.data = t.data_ptr()
    b.dtype = dtype_code(t)
    b.nfunc = int(nfunc if nfunc is not None else t.shape[0])
    b.grid = grid
    return b


# the autotuner's tunables, in MGX_TUNE_* order
TUNE_NAMES = ["march_f16", "march_f32", "march_f64", "hloc_f64", "gram_f64", "gram_split", "restrict", "prolong"]


class BoxSet:
    def __init__(self, ctx, origin, edge, gdim, h, levels):
        import numpy as np
        self.ctx = ctx
        self._origin = np.ascontiguousarray(np.asarray(origin, dtype=np.int32).reshape(-1))
        self._edge = np.ascontiguousarray(np.asarray(edge, dtype=np.int32).reshape(-1))
        b = Boxes()
        b.nbox = len(self._edge)
        for i in range(3):
            b.gdim[i] = int(gdim[i] if hasattr(gdim, "__len__") else gdim)
            b.h[i] = float(h[i] if hasattr(h, "__len__") else h)
        b.origin = self._origin.ctypes.data_as(C.POINTER(C.c_int))
        b.edge = self._edge.ctypes.data_as(C.POINTER(C.c_int))
        self.desc = b
        self.h = C.c_void_p()
        ctx.check(lib().mgx_boxset_create(ctx.h, C.byref(b), levels, C.byref(self.h)), "mgx_boxset_create")
        self.nbox = b.nbox
        self.levels = levels

    def size(self, level=0):
        n = C.c_int64()
        self.ctx.check(lib().mgx_boxset_size(self.h, level, C.byref(n)), "mgx_boxset_size")
        return n.value

    def npairs(self):
        n = C.c_int64()
        self.ctx.check(lib().mgx_boxset_npairs(self.h, C.byref(n)), "mgx_boxset_npairs")
        return n.value

    def autotune(self):
        """run the autotuner; returns {tunable: (choice, ms of candidate 1, ms of candidate 2)}"""
        ch = (C.c_int * len(TUNE_NAMES))()
        ms = (C.c_double * (2 * len(TUNE_NAMES)))()
        self.ctx.check(lib().mgx_boxset_autotune(self.ctx.h, self.h, ch, ms), "mgx_boxset_autotune")
        return {k: (ch[i], ms[2 * i], ms[2 * i + 1]) for i, k in enumerate(TUNE_NAMES)}

    def set_tune(self, choice):
        """force the launch shapes: {tunable: 1 or 2} (missing ones stay)"""
        cur = self.get_tune()[0]
        cur.update(choice)
        ch = (C.c_int * len(TUNE_NAMES))(*[cur[k] for k in TUNE_NAMES])
        self.ctx.check(lib().mgx_boxset_set_tune(self.h, ch), "mgx_boxset_set_tune")

    def get_tune(self):
        """({tunable: choice}, tuned)"""
        ch = (C.c_int * len(TUNE_NAMES))()
        t = C.c_int()
        self.ctx.check(lib().mgx_boxset_get_tune(self.h, ch, C.byref(t)), "mgx_boxset_get_tune")
        return {k: ch[i] for i, k in enumerate(TUNE_NAMES)}, bool(t.value)

    def density(self, psi, dm, rho, nel=0.0):
        """rho (device, gdim^3) = sum_ij dm_ij psi_i psi_j (mgx_box_density);
        psi, dm, rho: DeviceArray or raw device pointers.  Returns the charge
        vel * sum rho before any rescaling to nel electrons."""
        q = C.c_double()
        self.ctx.check(lib().mgx_box_density(self.ctx.h, self.h, ptr(psi), ptr(dm), self.nbox, nel, ptr(rho),
                                             C.byref(q)), "mgx_box_density")
        return q.value

    def close(self):
        if self.h:
            lib().mgx_boxset_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    def __init__(self, ctx, boxset, vtot, psi0, mg_levels=2, precond_dtype=F64, gram_prec=PREC_F64,
                 step_length=1.5, ritz_every=0, precond_start=START_ZERO, profile=0, anderson_m=0,
                 anderson_beta=1.0, conv_criterion=CONV_RESIDUAL, precond_kind=PRECOND_VCYCLE, cnn=None,
                 inverse_mode=INV_NS_F64, ns_tol=0.0, ns_max_spec=0, lap_type=LAP_4TH):
        import numpy as np
        self.ctx = ctx
        self.boxset = boxset
        cfg = SolverCfg(lap_type, mg_levels, precond_dtype, gram_prec, step_length, ritz_every, precond_start,
                        profile, anderson_m, anderson_beta, conv_criterion, precond_kind, inverse_mode, ns_tol,
                        ns_max_spec)
        v = np.ascontiguousarray(vtot, dtype=np.float64)
        p = np.ascontiguousarray(psi0, dtype=np.float64)
        self.npts = p.size
        self.h = C.c_void_p()
        ctx.check(lib().mgx_solver_create(ctx.h, boxset.h, C.byref(cfg), v.ctypes.data_as(C.c_void_p),
                                          p.ctypes.data_as(C.c_void_p), C.byref(self.h)), "mgx_solver_create")
        if cnn is not None:
            self.set_cnn(cnn)

    def density(self, rho, nel=0.0):
        """rho (device, gdim^3) of the current orbitals with dm = 2 S^-1
        (mgx_solver_density); returns the charge before rescaling."""
        q = C.c_double()
        self.ctx.check(lib().mgx_solver_density(self.h, nel, ptr(rho), C.byref(q)), "mgx_solver_density")
        return q.value

    def set_cnn(self, cnn):
        self.cnn = cnn  # keep it alive while attached
        self.ctx.check(lib().mgx_solver_set_cnn(self.h, cnn.h), "mgx_solver_set_cnn")

    def step(self):
        st = SolverStats()
        self.ctx.check(lib().mgx_solver_step(self.h, C.byref(st)), "mgx_solver_step")
        return st

    def run(self, tol, maxit):
        st = SolverStats()
        self.ctx.check(lib().mgx_solver_run(self.h, tol, maxit, C.byref(st)), "mgx_solver_run")
        return st

    def reset(self):
        self.ctx.check(lib().mgx_solver_reset(self.h), "mgx_solver_reset")

    def profile(self, reset=False):
        """{group: (ms, launches, work)} accumulated since the last reset;
        work = algorithmic bytes (flops for learned, nsgemm, nsgemm16)."""
        ms = (C.c_double * len(PROF_GROUPS))()
        cnt = (C.c_int64 * len(PROF_GROUPS))()
        work = (C.c_double * len(PROF_GROUPS))()
        self.ctx.check(lib().mgx_solver_profile_work(self.h, work), "mgx_solver_profile_work")
        self.ctx.check(lib().mgx_solver_profile(self.h, ms, cnt, 1 if reset else 0), "mgx_solver_profile")
        return {g: (ms[i], cnt[i], work[i]) for i, g in enumerate(PROF_GROUPS)}

    def set_psi(self, psi):
        import numpy as np
        p = np.ascontiguousarray(psi, dtype=np.float64)
        assert p.size == self.npts
        self.ctx.check(lib().mgx_solver_set_psi(self.h, p.ctypes.data_as(C.c_void_p)), "mgx_solver_set_psi")

    def matrices(self):
        """host copies of (S, S^-1) of the last evaluated step (n x n)."""
        import numpy as np
        n = self.boxset.nbox
        ps, pi = C.c_void_p(), C.c_void_p()
        self.ctx.check(lib().mgx_solver_buffers(self.h, None, None, C.byref(ps), None, None), "mgx_solver_buffers")
        self.ctx.check(lib().mgx_solver_inverse(self.h, C.byref(pi)), "mgx_solver_inverse")
        out = []
        for p in (ps, pi):
            a = np.empty((n, n), dtype=np.float64)
            self.ctx.check(lib().mgx_memcpy_dtoh(self.ctx.h, a.ctypes.data_as(C.c_void_p), p, a.nbytes), "dtoh")
            out.append(a.T.copy())  # column-major on the device
        return out[0], out[1]

    def set_profile(self, on):
        self.ctx.check(lib().mgx_solver_set_profile(self.h, 1 if on else 0), "mgx_solver_set_profile")

    def set_keep_residual(self, on):
        """keep R in the residual buffer every step (mgx_solver_set_keep_residual)"""
        self.ctx.check(lib().mgx_solver_set_keep_residual(self.h, 1 if on else 0), "mgx_solver_set_keep_residual")

    def set_graph(self, on):
        """replay each step's launch sequences as HIP graphs (mgx_solver_set_graph)"""
        self.ctx.check(lib().mgx_solver_set_graph(self.h, 1 if on else 0), "mgx_solver_set_graph")

    def graph_stats(self):
        """(graph launches, captures, graphs held)"""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.ctx.check(lib().mgx_solver_graph_stats(self.h, C.byref(a), C.byref(b), C.byref(c)),
                       "mgx_solver_graph_stats")
        return a.value, b.value, c.value

    def psi(self):
        import numpy as np
        out = np.zeros(self.npts)
        self.ctx.check(lib().mgx_solver_get_psi(self.h, out.ctypes.data_as(C.c_void_p)), "mgx_solver_get_psi")
        return out

    def close(self):
        if self.h:
            lib().mgx_solver_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
