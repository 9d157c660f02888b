"""Host-side setup of GL problems through the C++ host helpers
(include/pfdd_host.h, lib/libpfdd_host.so): material constants, slip tables,
resolved shear stress, core-energy coefficients and seeded initial fields.
This is the Python face of the same host layer the deck driver uses."""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libpfdd_host.so")
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


class Material(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("a", "b", "mu", "young", "isf", "usf", "An", "Cn", "dslip", "C44", "C12", "C11",
                 "nu", "S11", "S12", "S44", "mu_s", "xnu_s")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"host helper library missing at {LIB_PATH}; run the package build")
        L = ctypes.CDLL(LIB_PATH)
        ci, cd, cll, cp = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_char_p
        mp = ctypes.POINTER(Material)
        ip = ctypes.POINTER(ctypes.c_int)
        for name, args in {
            "pfdd_host_material_preset": [cp, mp],
            "pfdd_host_material_derive": [mp],
            "pfdd_host_slip": [cp, _dp, _dp, ip, _dp, _dp, _dp, _dp, _dp],
            "pfdd_host_resolss": [cp, _dp, _dp, _dp, cd, _dp, _dp],
            "pfdd_host_core_coef": [mp, ci, _dp, ip],
            "pfdd_host_ranmar_fill": [ci, cll, _dp],
            "pfdd_host_init_random_loops": [ci, ci, ci, cp, ci, cd, cd, cd, ci, _dp, _dp],
            "pfdd_host_init_sxtal": [cp, ci, ci, ci, ci, _dp, _dp],
            "pfdd_host_cond_preset": [cp, _dp],
            "pfdd_host_cond_layers": [_dp, _dp, ci, cd, _dp, _dp],
            "pfdd_host_rotate_stress_raw": [_dp, _dp, cd, _dp, _dp],
            "pfdd_host_resolss_raw": [ci, _dp, _dp, _dp, cd, _dp],
            "pfdd_host_resolss_non_schmid_raw": [ci, _dp, _dp, cd, cd, cd, cd, cd, _dp],
            "pfdd_host_project_core_energy": [ci, ci, ci, ci, _dp, _dp, _dp, _dp],
        }.items():
            f = getattr(L, name)
            f.restype = ci
            f.argtypes = args
        _lib = L
    return _lib


def _ok(rc, what):
    if rc != 0:
        raise ValueError(f"{what} failed (rc={rc})")


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def material(preset=None, **over):
    m = Material()
    _ok(lib().pfdd_host_material_preset(preset.encode() if preset else None, ctypes.byref(m)), "preset")
    for k, v in over.items():
        setattr(m, k, float(v))
    _ok(lib().pfdd_host_material_derive(ctypes.byref(m)), "material")
    return m


def cu_like():
    """Benchmark material: mu 48 GPa, nu 0.34, a 0.3615 nm, b = a/sqrt2, isf/usf 41/158 mJ/m^2."""
    a = 0.3615e-9
    return material(None, mu=48e9, young=128.64e9, a=a, b=a / np.sqrt(2.0), isf=41e-3, usf=158e-3)


def slip(app, prim=None, scale=None):
    xn, xb, four = np.zeros(36), np.zeros(36), np.zeros(9)
    ns, cd, ts = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    pr = None if prim is None else np.ascontiguousarray(prim, dtype=np.float64).reshape(9)
    sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
    _ok(lib().pfdd_host_slip(app.encode(), _p(pr), _p(sc), ctypes.byref(ns), _p(xn), _p(xb), _p(four),
                             ctypes.byref(cd), ctypes.byref(ts)), "slip table " + app)
    S = ns.value
    return {"xn": xn[:3 * S].reshape(S, 3), "xb": xb[:3 * S].reshape(S, 3), "four": four.reshape(3, 3),
            "CD": cd.value, "timestep": ts.value, "S": S}


def resolved_shear(app, sigma_mu, dslip, prim=None, scale=None):
    sl = slip(app, prim, scale)
    tau, srot = np.zeros(sl["S"]), np.zeros(9)
    pr = None if prim is None else np.ascontiguousarray(prim, dtype=np.float64).reshape(9)
    sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
    _ok(lib().pfdd_host_resolss(app.encode(), _p(pr), _p(sc),
                                _p(np.ascontiguousarray(sigma_mu, dtype=np.float64).reshape(9)),
                                dslip, _p(tau), _p(srot)), "resolss")
    return tau, srot.reshape(3, 3)


def core_coef(m, flag):
    c = np.zeros(8)
    n = ctypes.c_int()
    _ok(lib().pfdd_host_core_coef(ctypes.byref(m), int(flag), _p(c), ctypes.byref(n)), "core coef")
    return c[:n.value].copy()


def ranmar(seed, n):
    out = np.empty(int(n))
    _ok(lib().pfdd_host_ranmar_fill(int(seed), int(n), _p(out)), "ranmar")
    return out


def random_loops(shape, app, nloops, rmin, rmax, noise, seed):
    sl = slip(app)
    nx, ny, nz = shape
    xi = np.empty((sl["S"], nx, ny, nz), dtype=np.complex128)
    loops = np.zeros(5 * max(nloops, 1))
    _ok(lib().pfdd_host_init_random_loops(nx, ny, nz, app.encode(), nloops, rmin, rmax, noise, seed,
                                          xi.ctypes.data_as(_dp), _p(loops)), "init loops")
    return xi, loops[:5 * nloops].reshape(nloops, 5)


def sxtal(style, shape, S, with_mask=False):
    nx, ny, nz = shape
    xi = np.empty((S, nx, ny, nz), dtype=np.complex128)
    xo = np.zeros((S, nx, ny, nz)) if with_mask else None
    _ok(lib().pfdd_host_init_sxtal(style.encode(), nx, ny, nz, S, xi.ctypes.data_as(_dp), _p(xo)),
        "init " + style)
    return (xi, xo) if with_mask else xi


COND_KEYS = ("condsigma", "mfpel", "alpha_resistivity", "alpha_mfpel", "Rgrain", "beta_disloc", "rho_disloc",
             "dlayer", "Nlayers", "Ngrains", "pinterface")


def cond_material(preset=None, **over):
    """Conduction constants of a material (defaults, preset, then overrides) as a dict."""
    par = np.zeros(11)
    _ok(lib().pfdd_host_cond_preset((preset or "").encode(), _p(par)), f"cond preset {preset}")
    d = dict(zip(COND_KEYS, par.tolist()))
    d.update(over)
    return d


def cond_layers(matA, matB, ny, temperature=300.0):
    """Per-row (sigma, dsigma) of the two-material layer stack (host C++)."""
    pa = np.array([float(matA[k]) for k in COND_KEYS])
    pb = np.array([float(matB[k]) for k in COND_KEYS])
    sig, dsig = np.zeros(ny), np.zeros(ny)
    _ok(lib().pfdd_host_cond_layers(_p(pa), _p(pb), int(ny), float(temperature), _p(sig), _p(dsig)), "cond layers")
    return sig, dsig


def fcc_loops_problem(shape, nloops=8, rmin=8, rmax=20, noise=1e-3, shear_pa=400e6, seed=576,
                      cddt=0.01):
    """The 128^3 x 12 production workload: fcc table, Cu-like material, seeded
    loops + noise, constant shear resolved on system 0 (sigma = tau0 (b0 n0 + n0 b0)),
    sine core (flag 3)."""
    m = cu_like()
    sl = slip("fcc")
    b0 = sl["xb"][0] / np.linalg.norm(sl["xb"][0])
    n0 = sl["xn"][0] / np.linalg.norm(sl["xn"][0])
    sigma = (shear_pa / m.mu) * (np.outer(b0, n0) + np.outer(n0, b0))
    tau, srot = resolved_shear("fcc", sigma, m.dslip)
    xi, loops = random_loops(shape, "fcc", nloops, rmin, rmax, noise, seed)
    return {"shape": tuple(shape), "xn": sl["xn"], "xb": sl["xb"], "four": sl["four"],
            "C44": m.C44, "C12": m.C12, "C11": m.C11, "S12": m.S12, "S44": m.S44, "dslip": m.dslip,
            "mu_s": m.mu_s, "cddt": cddt, "core_coef": core_coef(m, 3), "tau": tau, "sigma_rot": srot,
            "xi0": xi, "loops": loops, "material": m.as_dict()}


def lognormal_amplitude(shape, corr_len=4.0, cv=0.15, seed=576):
    """Per-cell GSF amplitude of an alloy realization: lognormal with mean 1 and
    coefficient of variation `cv` of a unit-variance periodic Gaussian random
    field whose covariance is exp(-r^2 / (2 corr_len^2)) (white noise filtered
    in Fourier space by the square root of that covariance's spectrum,
    exp(-pi^2 corr_len^2 |k|^2) with k in cycles per cell), standardised."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(tuple(shape))
    k = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    filt = np.exp(-np.pi ** 2 * corr_len ** 2 * (k[0] ** 2 + k[1] ** 2 + k[2] ** 2))
    g = np.fft.ifftn(np.fft.fftn(w) * filt).real
    g = (g - g.mean()) / g.std()
    s2 = np.log1p(cv * cv)
    return np.exp(-0.5 * s2 + np.sqrt(s2) * g)


def edge_dislocation_field(shape, S, a=0):
    """Straight edge dislocation pair on the slip plane k = nz/2: Re xi_a = 1 for
    i < nx/2 (the periodic box holds the two boundary lines x = 0 and x = nx/2)."""
    nx, ny, nz = shape
    xi = np.zeros((S,) + tuple(shape), dtype=np.complex128)
    xi[a, : nx // 2, :, nz // 2] = 1.0
    return xi


def alloy_ensemble(groups=((48, 16), (64, 24), (96, 8)), stress_pa=(200e6, 600e6), corr_len=4.0, cv=0.15,
                   seed=576, cddt=0.01):
    """Batched alloy ensemble: for each (n, count) group, `count` realizations on
    an n^3 grid with the three in-plane systems of basal3 (n = z; b at 0, 120,
    240 deg), Cu-like material, sine core scaled per cell by a lognormal
    amplitude field, a straight edge dislocation of system 0, and an applied
    shear sigma_i ~ U(stress_pa) acting on system 0.  Returns one dict per
    group with the concatenated member inputs of a batch handle."""
    m = cu_like()
    sl = slip("basal3")
    S = sl["xn"].shape[0]
    b0 = sl["xb"][0] / np.linalg.norm(sl["xb"][0])
    n0 = sl["xn"][0] / np.linalg.norm(sl["xn"][0])
    unit = np.outer(b0, n0) + np.outer(n0, b0)
    rng = np.random.default_rng(seed)
    coef = np.tile(core_coef(m, 3), S)
    out = []
    member = 0
    for n, count in groups:
        shape = (n, n, n)
        xi0 = np.concatenate([edge_dislocation_field(shape, S) for _ in range(count)])
        stresses = rng.uniform(stress_pa[0], stress_pa[1], size=count)
        tau = np.concatenate([resolved_shear("basal3", (sp / m.mu) * unit, m.dslip)[0] for sp in stresses])
        amp = np.concatenate([np.broadcast_to(lognormal_amplitude(shape, corr_len, cv, seed + 1000 + member + q),
                                              (S,) + shape) for q in range(count)])
        member += count
        out.append({"shape": shape, "batch": count, "S": S, "xn": sl["xn"], "xb": sl["xb"], "four": sl["four"],
                    "C44": m.C44, "C12": m.C12, "C11": m.C11, "S12": m.S12, "S44": m.S44, "dslip": m.dslip,
                    "mu": m.mu, "cddt": cddt, "coef": coef, "ncoef": 2, "tau": tau, "stress_pa": stresses,
                    "xi0": xi0, "amp": np.ascontiguousarray(amp).reshape(-1)})
    return out


def frank_read_skewed(shape=(512, 512, 256), seg=64, shear_pa=500e6, cddt=0.01):
    """Large single-grid run: Frank-Read source on a 60-degree skewed grid
    (primitive x = [1 0 0], y = [1/2 sqrt3/2 0], z = [0 0 1]), the three basal3
    systems (n = z, b at 0/120/240 deg in the lattice frame, rotated by the
    reciprocal matrix as set_slip does), Cu-like material, sine core.
    Initial field: Re xi_0 = 1 on the slip plane k = nz/2 for the `seg` cells
    i in [nx/2 - seg/2, nx/2 + seg/2) and the two rows j = ny/2 - 2, ny/2 - 1,
    so the row boundary at j = ny/2 is a straight segment of `seg` cells.
    Its ends are pinned by frozen cells (mask xo = 1, all systems): a 2 x 4
    block at each end, i = nx/2 -+ (seg/2 + 1..2), j = ny/2 - 2 .. ny/2 + 1.
    Applied shear sigma = tau0 (b0 n0 + n0 b0) on system 0 with tau0 = shear_pa."""
    nx, ny, nz = shape
    prim = ([1.0, 0.0, 0.0], [0.5, math.sqrt(3.0) / 2.0, 0.0], [0.0, 0.0, 1.0])
    m = cu_like()
    sl = slip("basal3", prim=prim)
    S = sl["xn"].shape[0]
    b0 = sl["xb"][0] / np.linalg.norm(sl["xb"][0])
    n0 = sl["xn"][0] / np.linalg.norm(sl["xn"][0])
    sigma = (shear_pa / m.mu) * (np.outer(b0, n0) + np.outer(n0, b0))
    tau, srot = resolved_shear("basal3", sigma, m.dslip, prim=prim)
    xi = np.zeros((S, nx, ny, nz), dtype=np.complex128)
    i0, i1 = nx // 2 - seg // 2, nx // 2 + seg // 2
    xi[0, i0:i1, ny // 2 - 2:ny // 2, nz // 2] = 1.0
    xo = np.zeros((S, nx, ny, nz))
    for a, b in ((i0 - 2, i0), (i1, i1 + 2)):
        xo[:, a:b, ny // 2 - 2:ny // 2 + 2, nz // 2] = 1.0
    return {"shape": tuple(shape), "xn": sl["xn"], "xb": sl["xb"], "four": sl["four"], "C44": m.C44,
            "C12": m.C12, "C11": m.C11, "S12": m.S12, "S44": m.S44, "dslip": m.dslip, "cddt": cddt,
            "core_coef": core_coef(m, 3), "tau": tau, "sigma_rot": srot, "xi0": xi, "xo": xo}
