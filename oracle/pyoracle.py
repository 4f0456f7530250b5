"""ctypes binding of the CPU oracle (oracle/oracle.h) and of the reference build oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline, never as the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(ROOT, "build", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libsimple_ref.so")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run __graft_entry__.build() first")
        L = C.CDLL(LIB)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.c_void_p]
        L.or_destroy.argtypes = [vp]
        L.or_set_params.argtypes = [vp, C.c_void_p]
        L.or_field.restype = dp
        L.or_field.argtypes = [vp, C.c_int]
        for n in ["or_assemble_u", "or_bc_u", "or_assemble_v", "or_bc_v", "or_solve_p",
                  "or_correct", "or_mg_build", "or_mg_vcycle"]:
            getattr(L, n).argtypes = [vp]
            getattr(L, n).restype = None
        for n in ["or_residual_u", "or_residual_v", "or_assemble_p"]:
            getattr(L, n).argtypes = [vp]
            getattr(L, n).restype = C.c_double
        for n in ["or_sweep_u", "or_sweep_v"]:
            getattr(L, n).argtypes = [vp, C.c_int]
            getattr(L, n).restype = None
        for n in ["or_sweep_u_dc", "or_sweep_v_dc"]:
            getattr(L, n).argtypes = [vp, C.c_int]
            getattr(L, n).restype = None
        for n in ["or_sweep_u_kind", "or_sweep_v_kind"]:
            getattr(L, n).argtypes = [vp, C.c_int, C.c_int]
            getattr(L, n).restype = None
        L.or_solve_p_kind.argtypes = [vp, C.c_int]
        L.or_solve_p_kind.restype = None
        L.or_outer_iteration.argtypes = [vp, dp]
        L.or_outer_iteration.restype = None
        L.or_run.argtypes = [vp, dp, C.POINTER(C.c_int)]
        L.or_run.restype = C.c_int
        L.or_run_fixed.argtypes = [vp, C.c_int, dp]
        L.or_run_fixed.restype = C.c_double
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        L.or_cg_last_iterations.argtypes = [vp]
        L.or_cg_last_rel.argtypes = [vp]
        L.or_cg_last_rel.restype = C.c_double
        L.or_cg_total_iterations.argtypes = [vp]
        L.or_cg_total_iterations.restype = C.c_longlong
        L.or_mg_nlevels.argtypes = [vp]
        L.or_mg_dims.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_mg_get.argtypes = [vp, C.c_int, C.c_int, dp]
        L.or_mg_get.restype = None
        L.or_mg_last_coarse_sweeps.argtypes = [vp]
        L.or_apply_A_p.argtypes = [vp, dp, dp]
        L.or_apply_A_p.restype = None
        L.or_coefficients.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int, dp]
        L.or_coefficients.restype = None
        L.or_in_band.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib = L
    return _lib


def coefficients(params, ue, uw, vn, vs, i, j, which):
    """Coefficients [AE, AW, AN, AS, AEE, AWW, ANN, ASS, AP]: which 0 convective, 1 diffusive,
    2 sum."""
    out = np.zeros(9)
    lib().or_coefficients(C.byref(params), ue, uw, vn, vs, i, j, which, _dp(out))
    return out


def in_band(params, i, j):
    return bool(lib().or_in_band(C.byref(params), i, j))


FIELDS = ["u", "v", "p",
          "cu_AE", "cu_AW", "cu_AN", "cu_AS", "cu_AEE", "cu_AWW", "cu_ANN", "cu_ASS", "cu_AP",
          "cv_AE", "cv_AW", "cv_AN", "cv_AS", "cv_AEE", "cv_AWW", "cv_ANN", "cv_ASS", "cv_AP",
          "AEP", "AWP", "ANP", "ASP", "APP", "b", "pc"]
FID = {n: k for k, n in enumerate(FIELDS)}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """CPU oracle state for one cavity; params is an SgParams (simple_gpu.SgParams)."""

    def __init__(self, params):
        self.L = lib()
        self.p = params.copy()
        self.h = self.L.or_create(C.byref(self.p))
        self.shape = (self.p.nrows, self.p.ncols)

    def __del__(self):
        try:
            if self.h:
                self.L.or_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_params(self, params):
        self.p = params.copy()
        assert self.L.or_set_params(self.h, C.byref(self.p)) == 0

    def view(self, name):
        """Writable numpy view of a field (no copy)."""
        ptr = self.L.or_field(self.h, FID[name])
        return np.ctypeslib.as_array(ptr, shape=self.shape)

    def get(self, name):
        return self.view(name).copy()

    def set(self, name, a):
        self.view(name)[...] = a

    def assemble_u(self): self.L.or_assemble_u(self.h)
    def bc_u(self): self.L.or_bc_u(self.h)
    def residual_u(self): return self.L.or_residual_u(self.h)
    def sweep_u(self, n): self.L.or_sweep_u(self.h, n)
    def assemble_v(self): self.L.or_assemble_v(self.h)
    def bc_v(self): self.L.or_bc_v(self.h)
    def residual_v(self): return self.L.or_residual_v(self.h)
    def sweep_v(self, n): self.L.or_sweep_v(self.h, n)
    def assemble_p(self): return self.L.or_assemble_p(self.h)
    def solve_p(self): self.L.or_solve_p(self.h)
    def correct(self): self.L.or_correct(self.h)
    def sweep_u_dc(self, n): self.L.or_sweep_u_dc(self.h, n)
    def sweep_v_dc(self, n): self.L.or_sweep_v_dc(self.h, n)
    def sweep_u_kind(self, kind, n): self.L.or_sweep_u_kind(self.h, kind, n)
    def sweep_v_kind(self, kind, n): self.L.or_sweep_v_kind(self.h, kind, n)
    def solve_p_kind(self, kind): self.L.or_solve_p_kind(self.h, kind)

    def outer_iteration(self):
        r = np.zeros(3)
        self.L.or_outer_iteration(self.h, _dp(r))
        return r

    def run(self):
        log = np.zeros((self.p.maxit, 4))
        conv = C.c_int()
        n = self.L.or_run(self.h, _dp(log), C.byref(conv))
        return n, log[:n].copy(), bool(conv.value)

    def run_fixed(self, n):
        last = np.zeros(3)
        t = self.L.or_run_fixed(self.h, n, _dp(last))
        return t, last

    def mg_nlevels(self):
        return self.L.or_mg_nlevels(self.h)

    def mg_dims(self, l):
        a, b = C.c_int(), C.c_int()
        self.L.or_mg_dims(self.h, l, C.byref(a), C.byref(b))
        return a.value, b.value

    def mg_get(self, l, k):
        nx, ny = self.mg_dims(l)
        out = np.zeros((nx, ny))
        self.L.or_mg_get(self.h, l, k, _dp(out))
        return out

    def mg_build(self): self.L.or_mg_build(self.h)
    def mg_vcycle(self): self.L.or_mg_vcycle(self.h)

    def apply_A(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.shape)
        self.L.or_apply_A_p(self.h, _dp(x), _dp(out))
        return out

    def cg_last_iterations(self): return self.L.or_cg_last_iterations(self.h)
    def cg_last_rel(self): return self.L.or_cg_last_rel(self.h)
    def cg_total_iterations(self): return self.L.or_cg_total_iterations(self.h)
    def mg_last_coarse_sweeps(self): return self.L.or_mg_last_coarse_sweeps(self.h)


def set_threads(n):
    lib().or_set_threads(n)


def ref_available():
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        R.ref_create.restype = vp
        R.ref_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        R.ref_destroy.argtypes = [vp]
        R.ref_field.restype = dp
        R.ref_field.argtypes = [vp, C.c_int]
        R.ref_derived.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), dp, dp, dp]
        R.ref_outer_iteration.argtypes = [vp, dp]
        R.ref_run.argtypes = [vp, dp, C.POINTER(C.c_int)]
        for n in ["ref_assemble_u", "ref_bc_u", "ref_assemble_v", "ref_bc_v", "ref_solve_p",
                  "ref_correct"]:
            getattr(R, n).argtypes = [vp]
        for n in ["ref_residual_u", "ref_residual_v", "ref_assemble_p"]:
            getattr(R, n).argtypes = [vp]
            getattr(R, n).restype = C.c_double
        for n in ["ref_sweep_u", "ref_sweep_v"]:
            getattr(R, n).argtypes = [vp, C.c_int]
        R.ref_cg_last_iterations.argtypes = [vp]
        R.ref_mg_nlevels.argtypes = [vp]
        R.ref_set_threads.argtypes = [C.c_int]
        _ref = R
    return _ref


class Reference:
    """The reference's own CPU code (oracle/_ref), driven from a config text."""

    def __init__(self, cfg_text, maxit=None):
        self.R = ref_lib()
        err = C.create_string_buffer(1024)
        self.h = self.R.ref_create(cfg_text.encode(), err, 1024)
        if not self.h:
            raise ValueError(err.value.decode())
        nr, nc = C.c_int(), C.c_int()
        dx, dy, nu = C.c_double(), C.c_double(), C.c_double()
        self.R.ref_derived(self.h, C.byref(nr), C.byref(nc), C.byref(dx), C.byref(dy),
                           C.byref(nu))
        self.shape = (nr.value, nc.value)
        self.derived = (nr.value, nc.value, dx.value, dy.value, nu.value)
        self.maxit = maxit

    def __del__(self):
        try:
            if self.h:
                self.R.ref_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def view(self, name):
        ptr = self.R.ref_field(self.h, FID[name])
        return np.ctypeslib.as_array(ptr, shape=self.shape)

    def get(self, name):
        return self.view(name).copy()

    def set(self, name, a):
        self.view(name)[...] = a

    def outer_iteration(self):
        r = np.zeros(3)
        self.R.ref_outer_iteration(self.h, _dp(r))
        return r

    def run(self, maxit):
        log = np.zeros((maxit, 4))
        conv = C.c_int()
        n = self.R.ref_run(self.h, _dp(log), C.byref(conv))
        return n, log[:n].copy(), bool(conv.value)

    def assemble_u(self): self.R.ref_assemble_u(self.h)
    def bc_u(self): self.R.ref_bc_u(self.h)
    def residual_u(self): return self.R.ref_residual_u(self.h)
    def sweep_u(self, n): self.R.ref_sweep_u(self.h, n)
    def assemble_v(self): self.R.ref_assemble_v(self.h)
    def bc_v(self): self.R.ref_bc_v(self.h)
    def residual_v(self): return self.R.ref_residual_v(self.h)
    def sweep_v(self, n): self.R.ref_sweep_v(self.h, n)
    def assemble_p(self): return self.R.ref_assemble_p(self.h)
    def solve_p(self): self.R.ref_solve_p(self.h)
    def correct(self): self.R.ref_correct(self.h)
    def cg_last_iterations(self): return self.R.ref_cg_last_iterations(self.h)
