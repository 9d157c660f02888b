"""ctypes front end of the CPU FP64 oracle (oracle/pfdd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.

Array conventions follow the reference (lanl PFDD-2.x, src/fft_fftw_slab.cpp):
``xi`` is complex128 of shape (S, nx, ny, nz) (bit-identical to the
reference's interleaved ``xi[0][2*(k + j*nz + i*ny*nz + a*N) + {0,1}]``),
``BB`` is float64 of shape (S_b, S_a, nx, ny, nz) i.e. flat index
``r + a*N + b*N*S`` (fft_fftw_slab.cpp:631).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -O2 -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_double = ctypes.c_int, ctypes.c_double
        sigs = {
            "orc_set_threads": (None, [c_int]),
            "orc_get_threads": (c_int, []),
            "orc_material": (None, [c_double, c_double, _dp]),
            "orc_voigt": (None, [c_double, c_double, c_double, _dp]),
            "orc_prim": (None, [_dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_set_scale": (None, [_dp, c_double, c_double, c_double, _dp, _dp, _dp]),
            "orc_rotate_slip": (None, [c_int, _dp, _dp, _dp, _dp]),
            "orc_rotate_stress": (None, [_dp, _dp, c_double, _dp, _dp]),
            "orc_resolss": (None, [c_int, _dp, _dp, _dp, c_double, _dp]),
            "orc_theta": (None, [c_int, c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp]),
            "orc_resolss_non_schmid": (None, [c_int, _dp, _dp, c_double, c_double, c_double, c_double,
                                              c_double, _dp]),
            "orc_frec": (None, [c_int, c_int, c_int, _dp, _dp, _dp, _dp]),
            "orc_stiffness": (None, [c_double, c_double, _dp]),
            "orc_bmatrix": (None, [c_int, c_int, c_int, c_int, c_double, c_double, c_double,
                                   c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_fmatrix": (None, [c_int, c_int, c_int, c_int, c_double, c_double, c_double,
                                   c_double, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_fft3d": (None, [c_int, c_int, c_int, _dp, c_int]),
            "orc_dft1d_naive": (None, [c_int, _dp, _dp, c_int]),
            "orc_core_energy": (c_double, [c_int, c_int, c_int, c_int, _dp, c_int, _dp, c_int,
                                           _dp, _dp, _dp]),
            "orc_internal_energy": (None, [c_int, c_int, c_int, c_int, _dp, _dp, _dp]),
            "orc_update": (None, [c_int, c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp,
                                  c_double, _dp]),
            "orc_gl_step": (None, [c_int, c_int, c_int, c_int, _dp, _dp, _dp, c_double, _dp,
                                   c_int, _dp, c_int, _dp, _dp, _dp]),
            "orc_diagnostics": (None, [c_int, c_int, c_int, c_int, _dp, _dp, _dp, _dp,
                                       c_double, c_double, c_double, c_double, c_double,
                                       c_double, _dp, _dp, _dp, _dp, _dp]),
            "orc_ranmar_init": (None, [c_int, _dp]),
            "orc_ranmar_uniform": (c_double, [_dp]),
            "orc_ranmar_fill": (None, [_dp, ctypes.c_long, _dp]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    if a.dtype == np.complex128:
        return a.ctypes.data_as(_dp)
    assert a.dtype == np.float64, a.dtype
    return a.ctypes.data_as(_dp)


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


# --------------------------------------------------------------- setup pieces
def material(mu: float, young: float) -> dict:
    out = np.zeros(8)
    lib().orc_material(mu, young, _p(out))
    return dict(zip(["C44", "C12", "C11", "nu", "S11", "S12", "S44", "ll"], out.tolist()))


def voigt(C44, C12, C11) -> dict:
    out = np.zeros(4)
    lib().orc_voigt(C44, C12, C11, _p(out))
    return dict(zip(["mu", "ll", "young", "xnu"], out.tolist()))


def prim(x, y, z):
    p, ip, nv = np.zeros(9), np.zeros(9), ctypes.c_double()
    lib().orc_prim(_p(_f64(x)), _p(_f64(y)), _p(_f64(z)), _p(p), _p(ip), ctypes.byref(nv))
    return p.reshape(3, 3), ip.reshape(3, 3), nv.value


def set_scale(prim_m, sa=1.0, sb=1.0, sc=1.0):
    four, nf, vol = np.zeros(9), np.zeros(3), ctypes.c_double()
    lib().orc_set_scale(_p(_f64(prim_m).reshape(-1)), sa, sb, sc, _p(four), _p(nf),
                        ctypes.byref(vol))
    return four.reshape(3, 3), nf, vol.value


def rotate_slip(xn, xb, four, normfour):
    xn = _f64(xn).copy()
    xb = _f64(xb).copy()
    lib().orc_rotate_slip(xn.shape[0], _p(xn), _p(xb), _p(_f64(four).reshape(-1)), _p(_f64(normfour)))
    return xn, xb


def rotate_stress(prim_m, invprim, norvol, sigma):
    out = np.zeros(9)
    lib().orc_rotate_stress(_p(_f64(prim_m).reshape(-1)), _p(_f64(invprim).reshape(-1)), norvol,
                            _p(_f64(sigma).reshape(-1)), _p(out))
    return out.reshape(3, 3)


def resolss(sigma_rot, xn, xb, dslip):
    xn = _f64(xn)
    tau = np.zeros(xn.shape[0])
    lib().orc_resolss(xn.shape[0], _p(_f64(sigma_rot).reshape(-1)), _p(xn), _p(_f64(xb)), dslip, _p(tau))
    return tau


def resolss_non_schmid(sigma_rot, xb, dslip, angle, w1, w2, w3):
    xb = _f64(xb).reshape(-1, 3)
    tau = np.zeros(xb.shape[0])
    lib().orc_resolss_non_schmid(xb.shape[0], _p(xb), _p(_f64(sigma_rot).reshape(-1)), float(dslip),
                                 float(angle), float(w1), float(w2), float(w3), _p(tau))
    return tau


def theta(xi, four, xn, xb):
    """Dislocation character angle of gradient() (quirk Q6 included)."""
    S, nx, ny, nz = xi.shape
    out = np.zeros((S, nx, ny, nz))
    lib().orc_theta(nx, ny, nz, S, _p(np.ascontiguousarray(xi)), _p(_f64(four).reshape(-1)),
                    _p(_f64(xn).reshape(-1)), _p(_f64(xb).reshape(-1)), _p(out))
    return out


def frec(nx, ny, nz, four):
    fx, fy, fz = (np.zeros((nx, ny, nz)) for _ in range(3))
    lib().orc_frec(nx, ny, nz, _p(_f64(four).reshape(-1)), _p(fx), _p(fy), _p(fz))
    return fx, fy, fz


def stiffness(C44, C12):
    C = np.zeros(81)
    lib().orc_stiffness(C44, C12, _p(C))
    return C.reshape(3, 3, 3, 3)


def bmatrix(shape, C44, C12, C11, dslip, xn, xb, four):
    """Returns (BB[S_b, S_a, nx, ny, nz], eps[S,3,3], moduli dict)."""
    nx, ny, nz = shape
    xn, xb = _f64(xn), _f64(xb)
    ns = xn.shape[0]
    fx, fy, fz = frec(nx, ny, nz, four)
    BB = np.zeros((ns, ns, nx, ny, nz))
    eps = np.zeros((ns, 3, 3))
    mod = np.zeros(4)
    lib().orc_bmatrix(nx, ny, nz, ns, C44, C12, C11, dslip, _p(xn), _p(xb), _p(fx), _p(fy),
                      _p(fz), _p(BB), _p(eps), _p(mod))
    return BB, eps, dict(zip(["mu", "ll", "young", "xnu"], mod.tolist()))


def fmatrix(shape, C44, C12, mu, xnu, eps, four, with_D=False):
    nx, ny, nz = shape
    eps = _f64(eps)
    ns = eps.shape[0]
    fx, fy, fz = frec(nx, ny, nz, four)
    FF = np.zeros((3, 3, ns, nx, ny, nz))  # [j][i][a][r]
    DD = np.zeros_like(FF) if with_D else None
    lib().orc_fmatrix(nx, ny, nz, ns, C44, C12, mu, xnu, _p(eps), _p(fx), _p(fy), _p(fz),
                      _p(FF), _p(DD))
    return (FF, DD) if with_D else FF


def fft3d(x: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised in-place-style 3-D DFT (FFTW 2 semantics) of a complex128 array."""
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n0, n1, n2 = y.shape
    lib().orc_fft3d(n0, n1, n2, _p(y), int(sign))
    return y


def dft1d_naive(x: np.ndarray, sign: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().orc_dft1d_naive(x.shape[0], _p(x), _p(out), int(sign))
    return out


def core_energy(xi, flag, par, nm=0, amp=None):
    S, nx, ny, nz = xi.shape
    dE = np.zeros((S, nx, ny, nz))
    ev = ctypes.c_double()
    par = _f64(par) if par is not None else np.zeros(1)
    amp_a = _f64(amp) if amp is not None else None
    eref = lib().orc_core_energy(nx, ny, nz, S, _p(np.ascontiguousarray(xi)), int(flag), _p(par),
                                 int(nm), _p(amp_a), _p(dE), ctypes.byref(ev))
    return dE, eref, ev.value


class GLState:
    """One reference GL run on a single rank (SolveGL::iterate, solve_GL.cpp:70-158)."""

    def __init__(self, xi, BB, cddt, core_flag, core_par, nm=0, amp=None, xo=None, four=None, xn=None, xb=None):
        self.xi = np.ascontiguousarray(xi, dtype=np.complex128).copy()
        self.S, self.nx, self.ny, self.nz = self.xi.shape
        self.BB = np.ascontiguousarray(BB, dtype=np.float64)
        assert self.BB.shape == (self.S, self.S, self.nx, self.ny, self.nz)
        self.cddt = float(cddt)
        self.core_flag = int(core_flag)
        self.core_par = _f64(core_par) if core_par is not None else np.zeros(1)
        self.nm = int(nm)
        self.amp = _f64(amp) if amp is not None else None
        self.xo = _f64(xo) if xo is not None else None
        self.work = np.zeros(2 * 2 * self.S * self.nx * self.ny * self.nz)
        self.four, self.xn, self.xb = four, xn, xb  # core flag 5 (theta)

    @property
    def data_fftw(self):
        """FFT of the pre-update xi of the last step (what the output step reads)."""
        n = self.S * self.nx * self.ny * self.nz
        return self.work[: 2 * n].view(np.complex128).reshape(self.S, self.nx, self.ny, self.nz)

    def step(self, tau) -> dict:
        out = np.zeros(8)
        tau = _f64(tau)
        if self.core_flag == 5:
            # gradient() runs on the field of this iteration before core_energy
            self.amp = _f64(theta(self.xi, self.four, self.xn, self.xb))
        lib().orc_gl_step(self.nx, self.ny, self.nz, self.S, _p(self.xi), _p(self.xo), _p(self.BB),
                          self.cddt, _p(tau), self.core_flag, _p(self.core_par), self.nm,
                          _p(self.amp), _p(self.work), _p(out))
        keys = ["xinorm", "xiave", "E_el", "E_core", "E_stress", "E_total", "E_core_ref"]
        return dict(zip(keys, out[:7].tolist()))


def diagnostics(data_fftw, xi, FF, eps, mat, mu_s, sigma_rot, fields=False):
    """Output-step ave_eps / ave_sigma (3x3 each); optional fields."""
    S, nx, ny, nz = xi.shape
    ae, asg = np.zeros(9), np.zeros(9)
    de = np.zeros(2 * 9 * nx * ny * nz) if fields else None
    ds = np.zeros(2 * 9 * nx * ny * nz) if fields else None
    lib().orc_diagnostics(nx, ny, nz, S, _p(np.ascontiguousarray(data_fftw)),
                          _p(np.ascontiguousarray(xi)), _p(_f64(FF)), _p(_f64(eps)),
                          mat["C44"], mat["C12"], mat["C11"], mat["S12"], mat["S44"], mu_s,
                          _p(_f64(sigma_rot).reshape(-1)), _p(ae), _p(asg), _p(de), _p(ds))
    res = {"ave_eps": ae.reshape(3, 3), "ave_sigma": asg.reshape(3, 3)}
    if fields:
        res["data_eps"] = de.view(np.complex128).reshape(3, 3, nx, ny, nz)
        res["data_sigma"] = ds.view(np.complex128).reshape(3, 3, nx, ny, nz)
    return res


class Ranmar:
    """RanMars (random_mars.cpp:41-101)."""

    def __init__(self, seed: int):
        self.st = np.zeros(103)
        lib().orc_ranmar_init(int(seed), _p(self.st))

    def uniform(self) -> float:
        return lib().orc_ranmar_uniform(_p(self.st))

    def fill(self, n: int) -> np.ndarray:
        out = np.empty(int(n))
        lib().orc_ranmar_fill(_p(self.st), int(n), _p(out))
        return out


def adjoint_rollout(xi0, BB, cddt, tau, coef, nsteps, dxi_T, amp=None, xo=None):
    """Reverse-mode derivative of `nsteps` GL iterations with the sine-series
    core energy (core flag 8: dE = A sum_m c_m m pi sin(2 m pi Re xi)).

    Not a reference function (the reference has no derivatives): the adjoint
    of orc_gl_step / orc_update (update_order_parameter,
    fft_fftw_slab.cpp:1746-1790) written out directly in numpy with all
    forward states held in memory, for small grids.  Re and Im parts are
    independent real variables.  Returns (dL/dxi_0, dL/dtau[S], dL/dcoef[S, nm]).
    """
    coef = np.atleast_2d(_f64(coef))
    S = xi0.shape[0]
    nm = coef.shape[1]
    st = GLState(xi0, BB, cddt, 8, coef.reshape(-1), nm=nm, amp=amp, xo=xo)
    states = [st.xi.copy()]
    for _ in range(nsteps):
        st.step(tau)
        states.append(st.xi.copy())
    U = np.ones(xi0.shape) if xo is None else (np.asarray(xo).reshape(xi0.shape) == 0).astype(float)
    A = np.ones(xi0.shape) if amp is None else np.asarray(amp, dtype=float).reshape(xi0.shape)
    BBt = np.ascontiguousarray(BB)  # BB[b, a]: ghat_a = sum_b BB[b, a] xihat_b
    lam = np.array(dxi_T, dtype=np.complex128)
    gtau = np.zeros(S)
    gc = np.zeros((S, nm))
    m = np.arange(1, nm + 1).reshape(nm, 1, 1, 1)
    for n in range(nsteps - 1, -1, -1):
        xr = states[n].real
        mu = U * lam
        muh = np.fft.fftn(mu, axes=(1, 2, 3))
        gh = np.einsum("ba...,b...->a...", BBt, muh)
        g = np.fft.ifftn(gh, axes=(1, 2, 3))  # = IFFT_unnormalised(gh) / N
        d2 = np.zeros(xi0.shape)
        for a in range(S):
            ang = 2.0 * np.pi * m * xr[a][None]
            d2[a] = A[a] * np.sum(coef[a].reshape(nm, 1, 1, 1) * 2.0 * (m * np.pi) ** 2 * np.cos(ang), axis=0)
            gc[a] -= cddt * np.sum(mu[a].real[None] * A[a][None] * m * np.pi * np.sin(ang), axis=(1, 2, 3))
            gtau[a] += cddt * np.sum(mu[a].real)
        lam = lam - cddt * (g + d2 * mu.real)
    return lam, gtau, gc, states[-1]


# ------------------------------------------------------------- conductivity
def _c_round(x: float) -> int:
    """C's round() for the non-negative layer counts: halves away from zero."""
    return int(np.floor(x + 0.5)) if x >= 0 else -int(np.floor(-x + 0.5))


def cond_layers(matA: dict, matB: dict, ny: int, temperature: float):
    """backwardcond's per-row layer algebra (fft_fftw_slab.cpp:2083-2159) and
    the bulk difference of its gradient term (:2171-2182): returns
    (sigma[ny], dsigma[ny]).  Materials as dicts with the keys condsigma,
    mfpel, alpha_resistivity, alpha_mfpel, Rgrain, beta_disloc, rho_disloc,
    dlayer, Nlayers, Ngrains, pinterface; A is the first material command."""
    Vf = matB["dlayer"] / (matA["dlayer"] + matB["dlayer"])
    Nl = int(matA["Nlayers"])
    prob = 1.0 - matA["pinterface"]
    dT = temperature - 300
    per = ny // Nl  # N2/(Nlayers): integer division
    dB, dA = _c_round(per * Vf), _c_round(per * (1.0 - Vf))
    bulk, sig = np.zeros(ny), np.zeros(ny)
    for j in range(ny):
        m = matB if j % (dA + dB) >= dA else matA
        bulk[j] = m["condsigma"] / (1.0 + m["alpha_resistivity"] * dT)
        ps = 2 * prob / m["dlayer"]
        lam = m["mfpel"] / (1.0 + m["alpha_mfpel"] * dT)
        ng = int(m["Ngrains"])
        ag = (lam * ng / m["dlayer"]) * m["Rgrain"] / (1 - m["Rgrain"])
        rhod = m["beta_disloc"] * m["rho_disloc"]
        if ps > 0.0:
            rhoi = (1 / bulk[j]) * ((3.0 / 16.0) * lam * ps)
            if ng > 1:
                rhog = (1 / bulk[j]) * (1.0 / (1.0 - 3.0 * ag / 2.0 + 3.0 * ag * ag -
                                               3.0 * ag ** 3 * np.log(1.0 + 1.0 / ag)) - 1.0)
                sig[j] = 1 / ((1 / bulk[j]) + (2 * rhoi + (ng - 1) * rhog) / (ng + 1) + rhod)
            else:
                sig[j] = 1 / ((1 / bulk[j]) + rhoi + rhod)
        else:
            sig[j] = bulk[j]
    dsig = np.zeros(ny)
    for j in range(ny):
        if j == 0:
            dsig[j] = bulk[j + 1] - bulk[j]
        elif j == ny - 1:
            dsig[j] = bulk[j] - bulk[j - 1]
        else:
            dsig[j] = (bulk[j + 1] - bulk[j - 1]) / 2.0
    return sig, dsig


def cond_step(c, four, sigma, dsigma, El, eps0, rate):
    """One iteration of the conductivity field c (nx, ny, nz) as
    SolveGLCond::iterate runs it: forwardcond (fft_fftw_slab.cpp:1999-2006),
    divcurrent (:1704-1741), backwardcond (:2057-2191),
    update_conduct_order_param (:1794-1835); then backwardcond once more on
    its own real-space output (solve_GL+cond.cpp:136), evaluated literally by
    a second inverse transform.  Returns (c_new, scalars)."""
    c = np.asarray(c, dtype=np.complex128)
    nx, ny, nz = c.shape
    N = nx * ny * nz
    ch = fft3d(c, -1)
    fx, fy, fz = frec(nx, ny, nz, four)
    f2 = fx * fx + fy * fy + fz * fz
    pos = f2 > 0.0
    a = np.zeros_like(ch)
    b = np.zeros_like(ch)
    a.real[pos] = -(ch.imag[pos] * fy[pos] / f2[pos])
    a.imag[pos] = ch.real[pos] * fy[pos] / f2[pos]
    b.real[pos] = ch.real[pos] * fy[pos] * fy[pos] / f2[pos]
    b.imag[pos] = ch.imag[pos] * fy[pos] * fy[pos] / f2[pos]
    ar, br = fft3d(a, +1), fft3d(b, +1)  # unnormalised inverse transforms
    s = np.asarray(sigma, dtype=np.float64)[None, :, None]
    ds = np.asarray(dsigma, dtype=np.float64)[None, :, None]
    jr = s * (El - ar.real / eps0)
    ji = -s * ar.imag / eps0
    rr = (s / eps0) * br.real + (ar.real / eps0 - El) * ds
    ri = (s / eps0) * br.imag + (ar.imag / eps0) * ds
    cn = np.empty_like(c)
    cn.real = c.real - rate * (rr / N)
    cn.imag = c.imag - rate * (ri / N)
    d = cn - c
    a2 = fft3d(ar, +1)
    jr2 = s * (El - a2.real / eps0)
    out = {"javerage_re": float(jr.sum() / N), "javerage_im": float(np.broadcast_to(ji, c.shape).sum() / N),
           "condxinorm": float(np.sqrt(np.sum(d.real ** 2 + d.imag ** 2))), "condxiave": float(cn.real.sum()),
           "conductivity_repeat": float(np.broadcast_to(jr2, c.shape).sum() / N / El)}
    out["conductivity"] = out["javerage_re"] / El
    return cn, out
