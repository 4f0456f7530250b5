"""ctypes binding of the C ABI in include/simple_gpu.h (libsimple_gpu.so) and of the host library
(libsimple_host.so: config parsing, the C++ outer loop and output writers).

The binding is plumbing for tests and bench.py; the product is the C ABI itself. Loading fails
loudly if the built library is missing: there is no fallback path.
"""
import ctypes as C
import os

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BUILD = os.path.join(ROOT, "build")

UDS, QUICK = 0, 1
NORM_ABSOLUTE, NORM_RUNNING_MAX = 0, 1
BACKEND_SERIAL, BACKEND_OPENMP, BACKEND_GPU = 0, 1, 2
MOM_GS_LEX, MOM_JACOBI, MOM_GS_MULTICOLOUR = 0, 1, 2
P_GS_LEX, P_JACOBI, P_GS_REDBLACK, P_CG, P_MG = 0, 1, 2, 3, 4
MG_FP64, MG_MIXED = 0, 1
MOMP_FP64, MOMP_MIXED = 0, 1
WARM_NONE, WARM_LEARNED = 0, 1

FIELDS = ["u", "v", "p",
          "cu_AE", "cu_AW", "cu_AN", "cu_AS", "cu_AEE", "cu_AWW", "cu_ANN", "cu_ASS", "cu_AP",
          "cv_AE", "cv_AW", "cv_AN", "cv_AS", "cv_AEE", "cv_AWW", "cv_ANN", "cv_ASS", "cv_AP",
          "AEP", "AWP", "ANP", "ASP", "APP", "b", "pc"]
FIELD_ID = {n: k for k, n in enumerate(FIELDS)}
NPHASES = 11
PHASES = ["assemble_u", "bc_u", "residual_u", "sweep_u", "assemble_v", "bc_v", "residual_v",
          "sweep_v", "assemble_p", "solve_p", "correct"]
(K_SWEEP_U, K_SWEEP_V, K_COPY, K_ASSEMBLE_U, K_PSOLVE, K_SWEEP_U_FIRST, K_ASSEMBLE_V,
 K_SWEEP_V_FIRST, K_SWEEP_U_DC, K_ASSEMBLE_U_DC) = range(10)
TUNE_TAIL_THREADS, TUNE_MID_CELLS, TUNE_LDS_CELLS, TUNE_MOM_TILE, TUNE_MG_STAGED = 0, 1, 2, 3, 4
TUNE_MOM_TILE_FIRST = 5
TUNE_DC_TILE = 6
# Values of TUNE_DC_TILE: FP32 correction sweep tiles as (rows, columns, threads).
DC_TILES = ((32, 64, 256), (32, 64, 512), (64, 64, 512), (32, 128, 512), (48, 64, 512),
            (16, 128, 256), (24, 96, 512), (40, 56, 512))
# Values of TUNE_MOM_TILE / TUNE_MOM_TILE_FIRST: sweep tile shapes as (rows, columns).
MOM_TILES = ((32, 64), (64, 64), (32, 128), (16, 128), (24, 96), (48, 96), (24, 64), (40, 128),
             (32, 72), (28, 80), (36, 64), (40, 56))
# Launch-shape knobs in the order of bench.py's --knobs "tile,threads,mid,lds[,staged[,first]]".
TUNE_KEYS = (TUNE_MOM_TILE, TUNE_TAIL_THREADS, TUNE_MID_CELLS, TUNE_LDS_CELLS, TUNE_MG_STAGED,
             TUNE_MOM_TILE_FIRST, TUNE_DC_TILE)


def tuning_of(obj):
    """The launch-shape knobs of a Solver or Batch, in TUNE_KEYS order."""
    return tuple(obj.get_tuning(k) for k in TUNE_KEYS)


def set_tunings(obj, values):
    """Set the first len(values) knobs of TUNE_KEYS on a Solver or Batch."""
    for k, v in zip(TUNE_KEYS, values):
        obj.set_tuning(k, int(v))


class SgParams(C.Structure):
    """Mirror of struct sg_params (include/simple_gpu.h)."""
    _fields_ = [
        ("NX", C.c_int), ("NY", C.c_int), ("H", C.c_int),
        ("N_halo_x", C.c_int), ("N_halo_y", C.c_int),
        ("Re", C.c_double),
        ("maxit_inner_u", C.c_int), ("maxit_inner_v", C.c_int), ("maxit_inner_p", C.c_int),
        ("error_tolerance", C.c_double),
        ("maxit", C.c_int),
        ("relaxation_factor_u", C.c_double), ("relaxation_factor_v", C.c_double),
        ("relaxation_factor_p", C.c_double),
        ("convection_scheme", C.c_int), ("diffusion_order", C.c_int),
        ("Hybrid_switch", C.c_int), ("residual_normalisation", C.c_int),
        ("backend", C.c_int), ("momentum_solver", C.c_int), ("pressure_solver", C.c_int),
        ("pressure_tolerance", C.c_double), ("pressure_maxit", C.c_int),
        ("mg_cycles", C.c_int), ("mg_pre", C.c_int), ("mg_post", C.c_int),
        ("mg_omega", C.c_double),
        ("write_fields", C.c_int),
        ("nrows", C.c_int), ("ncols", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("nu", C.c_double),
        ("mg_precision", C.c_int), ("mg_coarsest", C.c_int), ("mg_dc_maxit", C.c_int),
        ("pressure_warm_start", C.c_int), ("check_every", C.c_int),
        ("momentum_precision", C.c_int),
    ]

    def copy(self):
        q = SgParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(SgParams))
        return q


_host = None
_gpu = None


def _load(name):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() first")
    return C.CDLL(path)


def host():
    """libsimple_host.so: config -> params, outputs (pure C++, loads without a GPU)."""
    global _host
    if _host is None:
        h = _load("libsimple_host.so")
        h.sh_params_from_config.argtypes = [C.c_char_p, C.POINTER(SgParams), C.c_char_p,
                                            C.c_int, C.c_char_p, C.c_int]
        h.sh_params_from_config.restype = C.c_int
        _host = h
    return _host


def params_from_config(text):
    """Parses config text with the host's parser (same keys and errors as the reference's
    make_params). Returns (SgParams, output_dir); raises ValueError with the parser message."""
    h = host()
    p = SgParams()
    out = C.create_string_buffer(4096)
    err = C.create_string_buffer(4096)
    rc = h.sh_params_from_config(text.encode(), C.byref(p), out, 4096, err, 4096)
    if rc != 0:
        raise ValueError(err.value.decode())
    return p, out.value.decode()


def params(**kv):
    """Params from key = value overrides of the defaults (values as config strings)."""
    text = "".join(f"{k} = {v}\n" for k, v in kv.items())
    return params_from_config(text)[0]


def gpu():
    """libsimple_gpu.so, the C ABI. Raises if it is not built."""
    global _gpu
    if _gpu is None:
        g = _load("libsimple_gpu.so")
        P = C.POINTER(SgParams)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        g.sg_last_error.restype = C.c_char_p
        g.sg_abi_version.restype = C.c_int
        g.sg_device_count.argtypes = [C.POINTER(C.c_int)]
        g.sg_device_name.argtypes = [C.c_char_p, C.c_size_t]
        g.sg_params_finalize.argtypes = [P]
        g.sg_params_default.argtypes = [P]
        g.sg_params_default.restype = None
        g.sg_create.argtypes = [P, dp, dp, dp, C.POINTER(vp)]
        g.sg_destroy.argtypes = [vp]
        g.sg_upload.argtypes = [vp, C.c_int, dp]
        g.sg_download.argtypes = [vp, C.c_int, dp]
        for n in ["sg_assemble_u", "sg_bc_u", "sg_assemble_v", "sg_bc_v", "sg_solve_p",
                  "sg_correct"]:
            getattr(g, n).argtypes = [vp]
        for n in ["sg_residual_u", "sg_residual_v", "sg_assemble_p"]:
            getattr(g, n).argtypes = [vp, dp]
        for n in ["sg_sweep_u", "sg_sweep_v"]:
            getattr(g, n).argtypes = [vp, C.c_int]
        g.sg_outer_iteration.argtypes = [vp, dp, dp]
        g.sg_advance.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp, C.POINTER(C.c_int)]
        g.sg_iterations.argtypes = [vp, C.POINTER(C.c_int)]
        g.sg_kernel_probe.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.POINTER(C.c_int)]
        g.sg_mg_levels.argtypes = [vp, C.POINTER(C.c_int)]
        g.sg_mg_coarse_sweeps.argtypes = [vp, C.POINTER(C.c_int)]
        g.sg_set_tuning.argtypes = [vp, C.c_int, C.c_long]
        g.sg_get_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_long)]
        g.sg_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.c_size_t]
        ip = C.POINTER(C.c_int)
        g.sg_batch_create.argtypes = [P, C.c_int, C.POINTER(vp)]
        g.sg_batch_destroy.argtypes = [vp]
        g.sg_batch_size.argtypes = [vp, ip]
        g.sg_batch_solver.argtypes = [vp, C.c_int, C.POINTER(vp)]
        g.sg_batch_advance.argtypes = [vp, C.c_int, ip, ip, dp]
        g.sg_batch_launches.argtypes = [vp, ip]
        g.sg_batch_set_tuning.argtypes = [vp, C.c_int, C.c_long]
        g.sg_batch_get_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_long)]
        g.sg_batch_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.c_size_t]
        fp = C.POINTER(C.c_float)
        g.sg_unet_nparams.argtypes = [ip]
        g.sg_unet_create.argtypes = [C.POINTER(vp)]
        g.sg_unet_destroy.argtypes = [vp]
        g.sg_unet_set_params.argtypes = [vp, fp, C.c_int]
        g.sg_unet_get_params.argtypes = [vp, fp, C.c_int]
        g.sg_unet_forward.argtypes = [vp, C.c_int, C.c_int, C.c_int, fp, fp]
        g.sg_unet_loss_grad.argtypes = [vp, C.c_int, C.c_int, C.c_int, fp, fp, fp, fp]
        g.sg_unet_train_step.argtypes = [vp, C.c_int, C.c_int, C.c_int, fp, fp, C.c_float, fp]
        g.sg_unet_mfma_probe.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        g.sg_attach_warm_start.argtypes = [vp, vp, C.c_int]
        g.sg_unet_forward_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]
        g.sg_set_phase_timing.argtypes = [vp, C.c_int]
        g.sg_phase_times.argtypes = [vp, dp, ip]
        g.sg_warm_features.argtypes = [vp, C.c_int, fp, dp]
        g.sg_warm_stats.argtypes = [vp, ip, ip]
        g.sg_warm_level.argtypes = [vp, ip]
        _gpu = g
    return _gpu


class GpuError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise GpuError(gpu().sg_last_error().decode())


def _dptr(a):
    import numpy as np
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """Thin owner of an sg_solver handle."""

    def __init__(self, p, u=None, v=None, pf=None):
        g = gpu()
        self.p = p.copy()
        h = C.c_void_p()
        args = [(_dptr(a) if a is not None else None) for a in (u, v, pf)]
        check(g.sg_create(C.byref(self.p), args[0], args[1], args[2], C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            gpu().sg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self):
        return (self.p.nrows, self.p.ncols)

    def get(self, name):
        import numpy as np
        a = np.empty(self.shape(), dtype=np.float64)
        check(gpu().sg_download(self.h, FIELD_ID[name], _dptr(a)))
        return a

    def set(self, name, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape()
        check(gpu().sg_upload(self.h, FIELD_ID[name], _dptr(a)))

    def _res(self, fn):
        r = C.c_double()
        check(fn(self.h, C.byref(r)))
        return r.value

    def assemble_u(self): check(gpu().sg_assemble_u(self.h))
    def bc_u(self): check(gpu().sg_bc_u(self.h))
    def residual_u(self): return self._res(gpu().sg_residual_u)
    def sweep_u(self, n): check(gpu().sg_sweep_u(self.h, n))
    def assemble_v(self): check(gpu().sg_assemble_v(self.h))
    def bc_v(self): check(gpu().sg_bc_v(self.h))
    def residual_v(self): return self._res(gpu().sg_residual_v)
    def sweep_v(self, n): check(gpu().sg_sweep_v(self.h, n))
    def assemble_p(self): return self._res(gpu().sg_assemble_p)
    def solve_p(self): check(gpu().sg_solve_p(self.h))
    def correct(self): check(gpu().sg_correct(self.h))

    def outer_iteration(self, phase_s=None):
        import numpy as np
        r = np.zeros(3)
        ph = None if phase_s is None else _dptr(phase_s)
        check(gpu().sg_outer_iteration(self.h, _dptr(r), ph))
        return r

    def advance(self, n):
        """Up to n iterations by graph replay with the device-side stop rule.
        Returns (iterations done, log[done, 3], converged)."""
        import numpy as np
        log = np.zeros((n, 3))
        done = C.c_int()
        conv = C.c_int()
        check(gpu().sg_advance(self.h, n, C.byref(done), _dptr(log), C.byref(conv)))
        return done.value, log[:done.value].copy(), bool(conv.value)

    def set_phase_timing(self, on=True):
        check(gpu().sg_set_phase_timing(self.h, 1 if on else 0))

    def phase_times(self):
        """(accumulated device seconds per phase [NPHASES], timed iterations)"""
        import numpy as np
        a = np.zeros(NPHASES)
        n = C.c_int()
        check(gpu().sg_phase_times(self.h, _dptr(a), C.byref(n)))
        return a, n.value

    def iterations(self):
        k = C.c_int()
        check(gpu().sg_iterations(self.h, C.byref(k)))
        return k.value

    def probe(self, kernel, reps=20):
        t = C.c_double()
        b = C.c_double()
        n = C.c_int()
        check(gpu().sg_kernel_probe(self.h, kernel, reps, C.byref(t), C.byref(b), C.byref(n)))
        return t.value, b.value, n.value

    def coarse_sweeps(self):
        n = C.c_int()
        check(gpu().sg_mg_coarse_sweeps(self.h, C.byref(n)))
        return n.value

    def set_tuning(self, key, value):
        check(gpu().sg_set_tuning(self.h, key, value))

    def get_tuning(self, key):
        v = C.c_long()
        check(gpu().sg_get_tuning(self.h, key, C.byref(v)))
        return v.value

    def autotune(self, reps=20):
        buf = C.create_string_buffer(8192)
        check(gpu().sg_autotune(self.h, reps, buf, 8192))
        return buf.value.decode()

    def mg_levels(self):
        n = C.c_int()
        check(gpu().sg_mg_levels(self.h, C.byref(n)))
        return n.value


def _fptr(a):
    import numpy as np
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


class UNet:
    """Owner of an sg_unet handle (the learned warm start's network)."""

    def __init__(self, theta=None):
        h = C.c_void_p()
        check(gpu().sg_unet_create(C.byref(h)))
        self.h = h
        n = C.c_int()
        check(gpu().sg_unet_nparams(C.byref(n)))
        self.nparams = n.value
        if theta is not None:
            self.set_params(theta)

    def close(self):
        if self.h:
            gpu().sg_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta):
        import numpy as np
        t = np.ascontiguousarray(theta, dtype=np.float32)
        check(gpu().sg_unet_set_params(self.h, _fptr(t), t.size))

    def get_params(self):
        import numpy as np
        t = np.empty(self.nparams, np.float32)
        check(gpu().sg_unet_get_params(self.h, _fptr(t), t.size))
        return t

    def forward(self, x):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, h, w, _ = x.shape
        y = np.empty((n, h, w), np.float32)
        check(gpu().sg_unet_forward(self.h, n, h, w, _fptr(x), _fptr(y)))
        return y

    def forward_noalloc(self, nimg, h, w, x_dev, y_dev, stream=None):
        """sg_unet_forward_noalloc on device pointers (ints); enqueued on `stream`, not waited
        for."""
        check(gpu().sg_unet_forward_noalloc(self.h, nimg, h, w, x_dev, y_dev, stream))

    def loss_grad(self, x, t):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        n, h, w, _ = x.shape
        loss = C.c_float()
        g = np.empty(self.nparams, np.float32)
        check(gpu().sg_unet_loss_grad(self.h, n, h, w, _fptr(x), _fptr(t), C.byref(loss), _fptr(g)))
        return loss.value, g

    def train_step(self, x, t, lr):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        n, h, w, _ = x.shape
        loss = C.c_float()
        check(gpu().sg_unet_train_step(self.h, n, h, w, _fptr(x), _fptr(t), lr, C.byref(loss)))
        return loss.value

    def probe(self, n, h, w, reps=20):
        sec, fl = C.c_double(), C.c_double()
        check(gpu().sg_unet_mfma_probe(self.h, n, h, w, reps, C.byref(sec), C.byref(fl)))
        return sec.value, fl.value


def attach_warm_start(solver, net, level=-1):
    check(gpu().sg_attach_warm_start(solver.h, net.h if net is not None else None, level))


def warm_features(solver, level=2):
    import numpy as np
    p = solver.p
    b = 1 << level
    x = np.empty((p.NX // b, p.NY // b, 8), np.float32)
    sc = C.c_double()
    check(gpu().sg_warm_features(solver.h, level, _fptr(x), C.byref(sc)))
    return x, sc.value


def warm_level(solver):
    """The multigrid level the attached warm start runs on (-1: none attached)."""
    lv = C.c_int(0)
    check(gpu().sg_warm_level(solver.h, C.byref(lv)))
    return lv.value


def warm_stats(solver):
    a, r = C.c_int(), C.c_int()
    check(gpu().sg_warm_stats(solver.h, C.byref(a), C.byref(r)))
    return a.value, r.value


class _Borrowed(Solver):
    """A cavity of a Batch: the Solver interface without ownership."""

    def __init__(self, h, p):
        self.p = p.copy()
        self.h = h

    def close(self):
        self.h = None


class Batch:
    """Owner of an sg_batch handle: cavities advanced together, one launch per phase."""

    def __init__(self, plist):
        g = gpu()
        self.n = len(plist)
        arr = (SgParams * self.n)(*[q.copy() for q in plist])
        self.plist = [q.copy() for q in plist]
        h = C.c_void_p()
        check(g.sg_batch_create(arr, self.n, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            gpu().sg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cavity(self, k):
        h = C.c_void_p()
        check(gpu().sg_batch_solver(self.h, k, C.byref(h)))
        return _Borrowed(h, self.plist[k])

    def advance(self, n):
        """Up to n outer iterations; returns (iterations, status, last residuals) per cavity."""
        import numpy as np
        its = (C.c_int * self.n)()
        st = (C.c_int * self.n)()
        last = np.zeros((self.n, 3))
        check(gpu().sg_batch_advance(self.h, n, its, st, _dptr(last)))
        return list(its), list(st), last

    def launches(self):
        k = C.c_int()
        check(gpu().sg_batch_launches(self.h, C.byref(k)))
        return k.value

    def set_tuning(self, key, value):
        check(gpu().sg_batch_set_tuning(self.h, key, value))

    def get_tuning(self, key):
        v = C.c_long()
        check(gpu().sg_batch_get_tuning(self.h, key, C.byref(v)))
        return v.value

    def autotune(self, reps=5):
        buf = C.create_string_buffer(16384)
        check(gpu().sg_batch_autotune(self.h, reps, buf, 16384))
        return buf.value.decode()


WORKLOADS = os.path.join(ROOT, "simple-hpc_amd", "workloads")


def cavities(path=None):
    """The uneven-cavity workload as (N, Re) pairs: 64 square cavities, N uniform over
    {64, 96, 128, 192, 256}, then Re log-uniform in [100, 3200], drawn in that order from
    std::mt19937_64(0) by workloads/cavities.cpp (its output is committed)."""
    path = path or os.path.join(WORKLOADS, "cavities_seed0.txt")
    out = []
    with open(path) as f:
        for line in f:
            if line.strip():
                n, re = line.split()
                out.append((int(n), float(re)))
    return out


def batch_config(n, re, precision="mixed", tol=1e-5, maxit=10000000, check_every=200):
    """Config text of one cavity of the batched workload: QUICK, 2nd-order diffusion,
    multicolour momentum, multigrid p' (mixed or fp64), tolerance tol."""
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = {re!r}\nconvection_scheme = QUICK\n"
            f"diffusion_order = 2\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"error_tolerance = {tol!r}\nmaxit = {maxit}\nbackend = gpu\nwrite_fields = False\n"
            f"check_every = {check_every}\n")
    if precision == "mixed":
        text += "mg_precision = mixed\n"
    return text


def device_count():
    n = C.c_int(0)
    rc = gpu().sg_device_count(C.byref(n))
    return n.value if rc == 0 else 0
