"""Python binding of the gfx950 GL solver C ABI (include/pfdd_gl.h) via ctypes.

The solver itself is the HIP library ``lib/libpfdd_gl.so`` of this package;
this module only marshals arguments.  There is no CPU fallback: if the
library is missing, loading fails with an error that says how to build it.

Fields use the reference layout (``fft->xi[0]``, lanl PFDD-2.x
src/fft_fftw_slab.cpp:1760-1762): complex128 arrays of shape (S, nx, ny, nz).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libpfdd_gl.so")

PFDD_FP64, PFDD_FP32 = 0, 1
CORE_NONE, CORE_PERFECT, CORE_SINE, CORE_MPEA, CORE_SERIES = 0, 1, 3, 7, 8
CORE_EXTENDED, CORE_PYRII, CORE_BCC, CORE_FOURIER = 2, 4, 5, 9
MAX_SYSTEMS = 12
MAX_COEF = 8
NOALLOC_MAX_SYSTEMS = 1024  # PFDD_GL_NOALLOC_MAX_SYSTEMS
FLAG_STORED_B = 1
FLAG_AUTOTUNE = 2
STATUS = {0: "PFDD_OK", 1: "PFDD_ERR_ARG", 2: "PFDD_ERR_UNSUPPORTED", 3: "PFDD_ERR_HIP",
          4: "PFDD_ERR_NONFINITE", 5: "PFDD_ERR_STATE"}

_dp = ctypes.POINTER(ctypes.c_double)


class PfddError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class GLDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("ns", ctypes.c_int), ("precision", ctypes.c_int),
        ("four", ctypes.c_double * 9),
        ("C44", ctypes.c_double), ("C12", ctypes.c_double), ("C11", ctypes.c_double),
        ("S12", ctypes.c_double), ("S44", ctypes.c_double),
        ("dslip", ctypes.c_double),
        ("xn", _dp), ("xb", _dp),
        ("cddt", ctypes.c_double),
        ("core_model", ctypes.c_int), ("ncoef", ctypes.c_int),
        ("coef", _dp), ("amp", _dp), ("xo", _dp),
        ("device", ctypes.c_int), ("stream", ctypes.c_void_p),
        ("flags", ctypes.c_int),
    ]


class GLReduce(ctypes.Structure):
    _fields_ = [
        ("xinorm", ctypes.c_double), ("xiave", ctypes.c_double),
        ("e_elastic", ctypes.c_double), ("e_core", ctypes.c_double),
        ("e_stress", ctypes.c_double), ("e_total", ctypes.c_double),
        ("e_core_ref", ctypes.c_double),
        ("steps_done", ctypes.c_int), ("converged", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CondDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("four", ctypes.c_double * 9),
        ("sigma", _dp), ("dsigma", _dp),
        ("El", ctypes.c_double), ("epsilonzero", ctypes.c_double), ("rate", ctypes.c_double),
        ("device", ctypes.c_int), ("stream", ctypes.c_void_p),
    ]


class CondReduce(ctypes.Structure):
    _fields_ = [
        ("conductivity", ctypes.c_double), ("javerage_re", ctypes.c_double), ("javerage_im", ctypes.c_double),
        ("condxinorm", ctypes.c_double), ("condxiave", ctypes.c_double),
        ("conductivity_repeat", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = [
    "pfdd_gl_version", "pfdd_gl_last_error", "pfdd_gl_stream_rates", "pfdd_gl_create", "pfdd_gl_destroy",
    "pfdd_gl_set_xi", "pfdd_gl_get_xi", "pfdd_gl_xi_device", "pfdd_gl_set_xi_device",
    "pfdd_gl_get_xi_device",
    "pfdd_gl_set_tau", "pfdd_gl_set_core_coef", "pfdd_gl_step", "pfdd_gl_xi_sums",
    "pfdd_gl_xi_sums_noalloc",
    "pfdd_gl_diag", "pfdd_gl_diag_fields", "pfdd_gl_profile", "pfdd_gl_traffic",
    "pfdd_gl_tape_record", "pfdd_gl_tape_backward", "pfdd_gl_tape_backward_device",
    "pfdd_gl_create_batch", "pfdd_gl_graph_create", "pfdd_gl_graph_run", "pfdd_gl_graph_destroy",
    "pfdd_gl_autotune", "pfdd_gl_set_tuning", "pfdd_gl_theta",
    "pfdd_gl_cond_create", "pfdd_gl_cond_destroy", "pfdd_gl_cond_set", "pfdd_gl_cond_get", "pfdd_gl_cond_step",
]

_lib = None


def stream_rates(nbytes, reps=100, device=0):
    """GB/s of the library's nontemporal copy and add kernels over arrays of
    nbytes each (pfdd_gl_stream_rates): the practical streaming rates the
    field passes are compared with."""
    L = load_library()
    cp, ad = np.zeros(1), np.zeros(1)
    _check(L.pfdd_gl_stream_rates(int(device), int(nbytes), int(reps), _ptr(cp), _ptr(ad)))
    return float(cp[0]), float(ad[0])


def load_library(path: str = LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"PFDD GL HIP library not found at {path}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root")
    L = ctypes.CDLL(path)
    vp, ip, c_int, c_double = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double
    sig = {
        "pfdd_gl_version": (c_int, []),
        "pfdd_gl_last_error": (ctypes.c_char_p, []),
        "pfdd_gl_stream_rates": (c_int, [c_int, ctypes.c_longlong, c_int, _dp, _dp]),
        "pfdd_gl_create": (c_int, [ctypes.POINTER(GLDesc), ctypes.POINTER(vp)]),
        "pfdd_gl_destroy": (c_int, [vp]),
        "pfdd_gl_set_xi": (c_int, [vp, _dp]),
        "pfdd_gl_get_xi": (c_int, [vp, _dp]),
        "pfdd_gl_xi_device": (c_int, [vp, ctypes.POINTER(vp)]),
        "pfdd_gl_set_xi_device": (c_int, [vp, vp]),
        "pfdd_gl_get_xi_device": (c_int, [vp, vp]),
        "pfdd_gl_set_tau": (c_int, [vp, _dp]),
        "pfdd_gl_set_core_coef": (c_int, [vp, c_int, _dp]),
        "pfdd_gl_step": (c_int, [vp, c_int, c_double, ctypes.POINTER(GLReduce)]),
        "pfdd_gl_xi_sums": (c_int, [vp, _dp]),
        "pfdd_gl_xi_sums_noalloc": (c_int, [vp, _dp]),
        "pfdd_gl_diag": (c_int, [vp, _dp, _dp, _dp]),
        "pfdd_gl_diag_fields": (c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "pfdd_gl_profile": (c_int, [vp, c_int, _dp, ip]),
        "pfdd_gl_traffic": (c_int, [vp, _dp, _dp]),
        "pfdd_gl_tape_record": (c_int, [vp, c_int, c_int, ctypes.POINTER(GLReduce)]),
        "pfdd_gl_tape_backward": (c_int, [vp, _dp, _dp, _dp, _dp]),
        "pfdd_gl_tape_backward_device": (c_int, [vp, vp, vp, _dp, _dp]),
        "pfdd_gl_create_batch": (c_int, [ctypes.POINTER(GLDesc), c_int, ctypes.POINTER(vp)]),
        "pfdd_gl_graph_create": (c_int, [ctypes.POINTER(vp), c_int, c_int, ctypes.POINTER(vp)]),
        "pfdd_gl_graph_run": (c_int, [vp, c_int, c_double, ctypes.POINTER(GLReduce)]),
        "pfdd_gl_graph_destroy": (c_int, [vp]),
        "pfdd_gl_autotune": (c_int, [vp, c_int, ip]),
        "pfdd_gl_set_tuning": (c_int, [vp, ip]),
        "pfdd_gl_theta": (c_int, [vp, _dp]),
        "pfdd_gl_cond_create": (c_int, [ctypes.POINTER(CondDesc), ctypes.POINTER(vp)]),
        "pfdd_gl_cond_destroy": (c_int, [vp]),
        "pfdd_gl_cond_set": (c_int, [vp, _dp]),
        "pfdd_gl_cond_get": (c_int, [vp, _dp]),
        "pfdd_gl_cond_step": (c_int, [vp, c_int, ctypes.POINTER(CondReduce)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise PfddError(rc, load_library().pfdd_gl_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


class GLSolver:
    """One grid of S order parameters on one GPU (drop-in for the FFTW_Slab GL path)."""

    def __init__(self, shape, xn, xb, four, C44, C12, C11, dslip, cddt, *, S12=0.0, S44=0.0,
                 core_model=CORE_NONE, coef=None, ncoef=0, amp=None, xo=None,
                 precision=PFDD_FP64, device=0, stored_B=False, batch=1, autotune=False):
        L = load_library()
        self.shape = tuple(int(v) for v in shape)
        self._xn, self._xb = _f64(xn).reshape(-1, 3), _f64(xb).reshape(-1, 3)
        self.S = self._xn.shape[0]
        self.batch = int(batch)
        self.SM = self.S * self.batch
        self.precision = precision
        self._coef = _f64(coef).reshape(-1) if coef is not None else None
        self._amp = _f64(amp).reshape(-1) if amp is not None else None
        self._xo = _f64(xo).reshape(-1) if xo is not None else None
        # the library reads batch * S * N values from each per-cell array
        need = self.SM * int(np.prod(self.shape))
        for name, arr in (("amp", self._amp), ("xo", self._xo)):
            if arr is not None and arr.size != need:
                raise ValueError(f"{name} has {arr.size} values; a batch of {self.batch} x {self.S} systems "
                                 f"on {self.shape} needs {need} (one per member, system and cell)")
        d = GLDesc()
        d.nx, d.ny, d.nz = self.shape
        d.ns = self.S
        d.precision = precision
        for q, v in enumerate(_f64(four).reshape(-1)):
            d.four[q] = v
        d.C44, d.C12, d.C11, d.S12, d.S44, d.dslip = C44, C12, C11, S12, S44, dslip
        d.xn, d.xb = _ptr(self._xn), _ptr(self._xb)
        d.cddt = cddt
        d.core_model, d.ncoef = core_model, ncoef
        d.coef, d.amp, d.xo = _ptr(self._coef), _ptr(self._amp), _ptr(self._xo)
        d.device = device
        d.stream = None
        d.flags = (FLAG_STORED_B if stored_B else 0) | (FLAG_AUTOTUNE if autotune else 0)
        h = ctypes.c_void_p()
        _check(L.pfdd_gl_create_batch(ctypes.byref(d), self.batch, ctypes.byref(h)))
        self._h = h
        self._lib = L

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pfdd_gl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ field I/O
    def set_xi(self, xi):
        xi = np.ascontiguousarray(xi, dtype=np.complex128)
        assert xi.shape == (self.SM,) + self.shape, xi.shape
        _check(self._lib.pfdd_gl_set_xi(self._h, xi.ctypes.data_as(_dp)))

    def get_xi(self):
        out = np.empty((self.SM,) + self.shape, dtype=np.complex128)
        _check(self._lib.pfdd_gl_get_xi(self._h, out.ctypes.data_as(_dp)))
        return out

    def xi_device_ptr(self) -> int:
        p = ctypes.c_void_p()
        _check(self._lib.pfdd_gl_xi_device(self._h, ctypes.byref(p)))
        return p.value

    def set_xi_device(self, ptr: int):
        _check(self._lib.pfdd_gl_set_xi_device(self._h, ctypes.c_void_p(ptr)))

    def get_xi_device(self, ptr: int):
        _check(self._lib.pfdd_gl_get_xi_device(self._h, ctypes.c_void_p(ptr)))

    def set_tau(self, tau):
        t = _f64(tau).reshape(-1)
        assert t.size == self.SM
        _check(self._lib.pfdd_gl_set_tau(self._h, _ptr(t)))

    def set_core_coef(self, coef, ncoef):
        self._coef = _f64(coef).reshape(-1)
        _check(self._lib.pfdd_gl_set_core_coef(self._h, int(ncoef), _ptr(self._coef)))

    # ------------------------------------------------------------- stepping
    def step(self, nsteps=1, tol=0.0, sync=True):
        r = GLReduce()
        _check(self._lib.pfdd_gl_step(self._h, int(nsteps), float(tol),
                                      ctypes.byref(r) if sync else None))
        return r.as_dict() if sync else None

    def autotune(self, nsteps=4):
        """Pick the fastest LDS tile shape and radix plan per line pass
        (checked as a whole step) and the iteration variant; returns the
        choices (fwd z, fwd y, inv y: 0-2, or 3-5 for the same shapes with the
        radix-16 plan; inv z + update: 0-2; iteration: 0 five kernels, 1 five
        kernels with the pipelined x pass, 2 the three-kernel plane
        iteration)."""
        c = (ctypes.c_int * 5)()
        _check(self._lib.pfdd_gl_autotune(self._h, int(nsteps), c))
        return list(c)

    def set_tuning(self, choice):
        c = (ctypes.c_int * 5)(*[int(v) for v in choice])
        _check(self._lib.pfdd_gl_set_tuning(self._h, c))

    # --------------------------------------------------------- reverse mode
    def tape_record(self, nsteps, every):
        """nsteps GL iterations (as step(nsteps)) keeping checkpoints every `every`."""
        r = GLReduce()
        _check(self._lib.pfdd_gl_tape_record(self._h, int(nsteps), int(every), ctypes.byref(r)))
        return r.as_dict()

    def tape_backward(self, dxi_T, ncoef):
        """dL/dxi_T (complex, Re/Im independent) -> (dL/dxi_0, dL/dtau[S], dL/dcoef[S, ncoef])."""
        g = np.ascontiguousarray(dxi_T, dtype=np.complex128)
        assert g.shape == (self.S,) + self.shape, g.shape
        dxi0 = np.empty_like(g)
        dtau = np.zeros(self.S)
        dcoef = np.zeros(self.S * MAX_COEF)  # the library writes S * (its ncoef) <= S * MAX_COEF
        _check(self._lib.pfdd_gl_tape_backward(self._h, g.ctypes.data_as(_dp), dxi0.ctypes.data_as(_dp),
                                               _ptr(dtau), _ptr(dcoef)))
        return dxi0, dtau, dcoef[: self.S * int(ncoef)].reshape(self.S, int(ncoef))

    def tape_backward_device(self, dxi_T_ptr: int, dxi_0_ptr: int, ncoef):
        """As tape_backward with device buffers (complex128 layout)."""
        dtau = np.zeros(self.S)
        dcoef = np.zeros(self.S * MAX_COEF)
        _check(self._lib.pfdd_gl_tape_backward_device(self._h, ctypes.c_void_p(dxi_T_ptr),
                                                      ctypes.c_void_p(dxi_0_ptr), _ptr(dtau), _ptr(dcoef)))
        return dtau, dcoef[: self.S * int(ncoef)].reshape(self.S, int(ncoef))

    def xi_sums(self):
        out = np.zeros(self.SM)
        _check(self._lib.pfdd_gl_xi_sums(self._h, _ptr(out)))
        return out

    def xi_sums_noalloc(self):
        """pfdd_gl_xi_sums_noalloc: the same sums with no device allocation."""
        out = np.zeros(self.SM)
        _check(self._lib.pfdd_gl_xi_sums_noalloc(self._h, _ptr(out)))
        return out

    def diag(self, sigma_rot):
        ae, asg = np.zeros(9), np.zeros(9)
        _check(self._lib.pfdd_gl_diag(self._h, _ptr(_f64(sigma_rot).reshape(-1)), _ptr(ae), _ptr(asg)))
        return ae.reshape(3, 3), asg.reshape(3, 3)

    def diag_fields(self, sigma_rot):
        """Averages plus the output-step fields data_eps (complex) and
        data_sigma (real part; the reference's imaginary part is 0), each of
        shape (3 [j], 3 [i], nx, ny, nz) as the reference indexes them
        (r + i*N + j*3N)."""
        ae, asg = np.zeros(9), np.zeros(9)
        N = int(np.prod(self.shape))
        de = np.zeros(2 * 9 * N)
        ds = np.zeros(2 * 9 * N)
        _check(self._lib.pfdd_gl_diag_fields(self._h, _ptr(_f64(sigma_rot).reshape(-1)), _ptr(ae),
                                             _ptr(asg), _ptr(de), _ptr(ds)))
        return (ae.reshape(3, 3), asg.reshape(3, 3),
                de.view(np.complex128).reshape((3, 3) + self.shape),
                ds.view(np.complex128).reshape((3, 3) + self.shape))

    def profile(self, nsteps=10):
        ms = np.zeros(7)
        nk = ctypes.c_int()
        _check(self._lib.pfdd_gl_profile(self._h, int(nsteps), _ptr(ms), ctypes.byref(nk)))
        names = ["fwd_z", "fwd_y", "x_contract", "inv_y", "inv_z_update", "reduce", "step"]
        return dict(zip(names, ms.tolist()))

    def theta(self):
        """Character angle of gradient() for every system and cell (pfdd_gl_theta)."""
        out = np.zeros((self.SM,) + tuple(self.shape))
        _check(self._lib.pfdd_gl_theta(self._h, _ptr(out)))
        return out

    def traffic(self):
        k = np.zeros(5)
        tot = ctypes.c_double()
        _check(self._lib.pfdd_gl_traffic(self._h, _ptr(k), ctypes.byref(tot)))
        return k, tot.value


class GLGraph:
    """Several solvers stepped together by one hipGraph (pfdd_gl_graph_*)."""

    def __init__(self, solvers, iters=2):
        L = load_library()
        self._solvers = list(solvers)  # keep the handles alive
        arr = (ctypes.c_void_p * len(self._solvers))(*[s._h.value for s in self._solvers])
        g = ctypes.c_void_p()
        _check(L.pfdd_gl_graph_create(arr, len(self._solvers), int(iters), ctypes.byref(g)))
        self._g, self._lib = g, L

    def run(self, nsteps, tol=0.0):
        out = (GLReduce * len(self._solvers))()
        _check(self._lib.pfdd_gl_graph_run(self._g, int(nsteps), float(tol), out))
        return [r.as_dict() for r in out]

    def close(self):
        if getattr(self, "_g", None):
            self._lib.pfdd_gl_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CondSolver:
    """Conductivity field of solve_style GLCond on one GPU (pfdd_gl_cond_*):
    c has shape (nx, ny, nz), complex128; sigma, dsigma are the per-row layer
    conductivities (pfdd_host_cond_layers)."""

    def __init__(self, shape, four, sigma, dsigma, El, epsilonzero=8.8541878128e-12, rate=1e-22, device=0):
        L = load_library()
        self.shape = tuple(int(v) for v in shape)
        self._sig, self._dsig = _f64(sigma), _f64(dsigma)
        d = CondDesc()
        d.nx, d.ny, d.nz = self.shape
        d.four[:] = _f64(four).reshape(-1).tolist()
        d.sigma, d.dsigma = _ptr(self._sig), _ptr(self._dsig)
        d.El, d.epsilonzero, d.rate = float(El), float(epsilonzero), float(rate)
        d.device = device
        h = ctypes.c_void_p()
        _check(L.pfdd_gl_cond_create(ctypes.byref(d), ctypes.byref(h)))
        self._h, self._lib = h, L

    def set_c(self, c):
        a = np.ascontiguousarray(np.asarray(c, dtype=np.complex128).reshape(self.shape))
        _check(self._lib.pfdd_gl_cond_set(self._h, a.view(np.float64).ctypes.data_as(_dp)))

    def get_c(self):
        a = np.empty(self.shape, dtype=np.complex128)
        _check(self._lib.pfdd_gl_cond_get(self._h, a.view(np.float64).ctypes.data_as(_dp)))
        return a

    def step(self, nsteps=1):
        r = CondReduce()
        _check(self._lib.pfdd_gl_cond_step(self._h, int(nsteps), ctypes.byref(r)))
        return r.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pfdd_gl_cond_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
