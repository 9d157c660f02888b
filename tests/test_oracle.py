"""Pins of the CPU oracle (oracle/pfdd_oracle.c) against what the mathematics
fixes: the DFT definition (numpy, naive long-double DFT), the closed-form
isotropic elastic kernel of a single slip system, finite differences of the
energy, RANMAR's published test values, stress equilibrium and the analytic
averages of the output step.  The reference ships no golden vectors for this
path, so these are the checks that pin it (CPU only)."""
import math

import numpy as np
import pytest

import oracle
import problems as P


# ----------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 16, 12), (6, 10, 9), (4, 48, 3), (5, 7, 8)])
def test_fft3d_matches_numpy(shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = oracle.fft3d(x, -1)
    assert np.allclose(f, np.fft.fftn(x), rtol=0, atol=1e-12 * np.abs(x).sum())
    b = oracle.fft3d(x, +1)
    assert np.allclose(b, np.fft.ifftn(x) * x.size, rtol=0, atol=1e-12 * np.abs(x).sum())


@pytest.mark.parametrize("n", [8, 12, 48, 96, 128, 7])
def test_fft_line_matches_naive_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        ref = oracle.dft1d_naive(x, sign)
        got = oracle.fft3d(x.reshape(1, 1, n), sign).reshape(n)
        assert np.max(np.abs(got - ref)) < 5e-15 * n * np.max(np.abs(x))
        got = oracle.fft3d(x.reshape(n, 1, 1), sign).reshape(n)  # along the slow axis
        assert np.max(np.abs(got - ref)) < 5e-15 * n * np.max(np.abs(x))


def test_fft_impulse_and_plane_wave():
    n = (8, 12, 6)
    x = np.zeros(n, complex)
    x[0, 0, 0] = 1.0
    assert np.allclose(oracle.fft3d(x, -1), 1.0, atol=1e-15)
    i, j, k = np.meshgrid(*[np.arange(v) for v in n], indexing="ij")
    wave = np.exp(2j * np.pi * (3 * i / 8 + 5 * j / 12 + 1 * k / 6))
    f = oracle.fft3d(wave, -1)
    expect = np.zeros(n, complex)
    expect[3, 5, 1] = np.prod(n)
    assert np.allclose(f, expect, atol=1e-9)


# ------------------------------------------------------------- setup pieces
def test_reciprocal_matrix_and_frequencies():
    geo = P.geometry(((1, 0, 0), (0.5, math.sqrt(3) / 2, 0), (0, 0, 1)))
    sclprim = geo["prim"]  # scale 1
    # reciprocal lattice: four[i] . a_j = delta_ij (a_j = columns of sclprim)
    assert np.allclose(geo["four"] @ sclprim, np.eye(3), atol=1e-14)
    fx, fy, fz = oracle.frec(4, 6, 8, np.eye(3))
    assert fx[2, 0, 0] == -0.5 and fx[1, 0, 0] == 0.25 and fx[3, 0, 0] == -0.25
    assert fy[0, 3, 0] == -0.5 and fz[0, 0, 5] == -3 / 8


def test_material_constants():
    m = oracle.material(P.MU, P.YOUNG)
    assert m["C44"] == P.MU
    assert abs(m["nu"] - P.NU) < 1e-12
    # isotropy: C11 = C12 + 2 C44, Voigt mu equals C44 exactly
    assert abs(m["C11"] - (m["C12"] + 2 * m["C44"])) <= 1e-6
    v = oracle.voigt(m["C44"], m["C12"], m["C11"])
    assert v["mu"] == m["C44"] and abs(v["xnu"] - P.NU) < 1e-12
    # compliance is the inverse of the stiffness on symmetric tensors
    C = oracle.stiffness(m["C44"], m["C12"]).reshape(9, 9)
    S = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    S[i, j, k, l] = m["S44"] / 4 * ((i == k) * (j == l) + (i == l) * (j == k)) + \
                        m["S12"] * (i == j) * (k == l)
    eps = np.array([[1e-3, 2e-3, 0], [2e-3, -1e-3, 5e-4], [0, 5e-4, 3e-4]])
    sig = (C @ eps.reshape(9)).reshape(3, 3)
    back = (S.reshape(9, 9) @ sig.reshape(9)).reshape(3, 3)
    assert np.allclose(back, eps, rtol=1e-10, atol=1e-15)


# --------------------------------------------------------------------- B(k)
def _single_system_closed_form(nvec, bvec, f, xnu, d):
    """B(k) of one glide system (b perpendicular to n, |b| = |n| = 1), isotropic
    elasticity: (1/d^2)[1 - (n.k)^2 - (b.k)^2 + 2 (b.k)^2 (n.k)^2 / (1 - nu)]."""
    f2 = (f ** 2).sum(0)
    with np.errstate(invalid="ignore", divide="ignore"):
        kh = f / np.sqrt(f2)
    nk = np.tensordot(nvec, kh, 1)
    bk = np.tensordot(bvec, kh, 1)
    B = (1 - nk ** 2 - bk ** 2 + 2 * bk ** 2 * nk ** 2 / (1 - xnu)) / d ** 2
    B[f2 == 0] = 0.0
    return B


@pytest.mark.parametrize("nb", [((0, 0, 1), (1, 0, 0)), ((1, 1, 1), (-1, 1, 0))])
def test_bmatrix_single_system_closed_form(nb):
    n = np.array(nb[0], float); n /= np.linalg.norm(n)
    b = np.array(nb[1], float); b /= np.linalg.norm(b)
    m = P.material()
    shape = (8, 12, 16)
    BB, eps, mod = oracle.bmatrix(shape, m["C44"], m["C12"], m["C11"], P.DSLIP, n[None], b[None], np.eye(3))
    f = np.stack(oracle.frec(*shape, np.eye(3)))
    expect = _single_system_closed_form(n, b, f, mod["xnu"], P.DSLIP)
    assert np.max(np.abs(BB[0, 0] - expect)) < 1e-12 * np.max(np.abs(expect))
    assert BB[0, 0, 0, 0, 0] == 0.0


def _contracted_B(xn, xb, f, m, xnu, d):
    """mu B_ab = sig_a:eps_b - [2 u_a.u_b/f^2 - (u_a.f)(u_b.f)/((1-nu) f^4)]/(2 C44), u_a = sig_a f."""
    S = xn.shape[0]
    eps = np.einsum("ai,aj->aij", xb, xn) / d
    tr = np.trace(eps, axis1=1, axis2=2)
    sig = m["C44"] * (eps + eps.transpose(0, 2, 1)) + m["C12"] * tr[:, None, None] * np.eye(3)
    u = np.einsum("aij,j...->ai...", sig, f)
    uf = np.einsum("ai...,i...->a...", u, f)
    f2 = (f ** 2).sum(0)
    with np.errstate(invalid="ignore", divide="ignore"):
        B = np.einsum("aij,bij->ab", sig, eps)[:, :, None, None, None] - (
            2 * np.einsum("ai...,bi...->ab...", u, u) / f2
            - uf[:, None] * uf[None, :] / ((1 - xnu) * f2 ** 2)) / (2 * m["C44"])
    B = B / m["C44"]
    B[:, :, f2 == 0] = 0.0
    return B


def test_bmatrix_fcc_skewed_invariants():
    m = P.material()
    geo = P.geometry(((1, 0, 0), (0.5, math.sqrt(3) / 2, 0), (0, 0, 1)))
    xn, xb = oracle.rotate_slip(P.FCC_XN, P.FCC_XB, geo["four"], geo["normfour"])
    shape = (6, 8, 4)
    BB, eps, mod = oracle.bmatrix(shape, m["C44"], m["C12"], m["C11"], P.DSLIP, xn, xb, geo["four"])
    Bab = BB.transpose(1, 0, 2, 3, 4)  # [a][b] (BB is stored [b][a])
    # symmetric in (a, b)
    assert np.max(np.abs(Bab - Bab.transpose(1, 0, 2, 3, 4))) < 1e-13
    # even in k: B(-k) = B(k) (off the Nyquist planes, where frec maps +-n/2 to -1/2)
    neg = Bab[:, :, (-np.arange(6)) % 6][:, :, :, (-np.arange(8)) % 8][:, :, :, :, (-np.arange(4)) % 4]
    keep = np.ones(shape, bool)
    keep[3], keep[:, 4], keep[:, :, 2] = False, False, False
    assert np.max(np.abs(Bab[:, :, keep] - neg[:, :, keep])) < 1e-13
    # B(0) = 0 and positive semidefinite everywhere
    assert np.all(Bab[:, :, 0, 0, 0] == 0)
    w = np.linalg.eigvalsh(Bab.reshape(12, 12, -1).transpose(2, 0, 1))
    assert w.min() > -1e-12
    # independent contracted-vector form of the same kernel
    f = np.stack(oracle.frec(*shape, geo["four"]))
    ref = _contracted_B(xn, xb, f, m, mod["xnu"], P.DSLIP)
    assert np.max(np.abs(Bab - ref)) < 1e-12 * np.max(np.abs(ref))


# --------------------------------------------------------------- core energy
@pytest.mark.parametrize("flag", [1, 3, 8])
def test_core_energy_derivative_fd(flag):
    rng = np.random.default_rng(3)
    xi = (rng.random((2, 4, 4, 4)) * 2 - 0.5).astype(complex)
    if flag == 8:
        par, nm = rng.random(8) * 0.01, 4
    else:
        par, nm = np.array([0.004, 0.013]), 0
    dE, _, _ = oracle.core_energy(xi, flag, par, nm)
    h = 1e-6
    _, _, ep = oracle.core_energy(xi + h, flag, par, nm)
    _, _, em = oracle.core_energy(xi - h, flag, par, nm)
    assert abs((ep - em) / (2 * h) - dE.sum()) < 1e-7 * np.abs(dE).sum()


def _fd_amp(xi, flag, par, nm, amp, picks, h=1e-6):
    """Central differences of the variational core energy with respect to
    single elements (a, flat cell r) of Re xi."""
    out = []
    for (a, r) in picks:
        xp, xm = xi.copy(), xi.copy()
        xp.reshape(xi.shape[0], -1)[a, r] += h
        xm.reshape(xi.shape[0], -1)[a, r] -= h
        _, _, ep = oracle.core_energy(xp, flag, par, nm, amp)
        _, _, em = oracle.core_energy(xm, flag, par, nm, amp)
        out.append((ep - em) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("flag", [7, 8])
def test_core_energy_per_cell_amplitude(flag):
    """core_energy_mpea (flag 7, fft_fftw_slab.cpp:1196-1235: per-cell
    A_mpea[core_index], core_index = (lxs+i)*N2*N3 + j*N3 + k + isa*N1*N2*N3)
    and the amp factor of the sine series (flag 8):
      * the force is the derivative of the energy density: central
        differences of single elements with a non-uniform amplitude that
        differs between systems and cells;
      * index map: an amplitude that is 1 except one cell of one system (3)
        and a system with amplitude 0 reduces flag 7 to flag 1 (An' = 1)
        times that pattern, so a transposed or system-blind amp index fails;
      * flag 8 with one term of coefficient 1 equals flag 7 (forces and both
        energies); flag 8 with amp = 1 equals flag 8 without amp;
      * E_core as the reference accumulates it is the density sum / nx."""
    rng = np.random.default_rng(11)
    S, shape = 3, (4, 5, 6)
    xi = (rng.random((S,) + shape) * 2 - 0.5) + 1j * rng.random((S,) + shape)
    amp = np.exp(0.3 * rng.standard_normal((S,) + shape))
    if flag == 7:
        par, nm = np.zeros(1), 0
    else:
        par, nm = rng.random(S * 3) * 0.01, 3
    dE, eref, ev = oracle.core_energy(xi, flag, par, nm, amp)
    picks = [(0, 0), (1, 17), (2, 119), (2, 3), (0, 64), (1, 100)]
    fd = _fd_amp(xi, flag, par, nm, amp, picks)
    got = np.array([dE.reshape(S, -1)[a, r] for a, r in picks])
    assert np.allclose(got, fd, rtol=1e-7, atol=1e-9 * np.abs(dE).max())
    assert abs(eref - ev / shape[0]) <= 1e-14 * abs(ev)
    # index map through a pattern amplitude, against flag 1 with An' = 1
    pat = np.ones((S,) + shape)
    pat[2, 1, 3, 4] = 3.0
    pat[1] = 0.0
    d1, _, e1 = oracle.core_energy(xi, 1, np.array([1.0]))
    if flag == 7:
        dp, _, ep = oracle.core_energy(xi, 7, np.zeros(1), 0, pat)
        assert np.allclose(dp, pat * d1, rtol=1e-15, atol=0)
        dens1 = np.sin(np.pi * xi.real) ** 2
        assert abs(ep - np.sum(pat * dens1)) <= 1e-13 * abs(ep)
    else:
        one = np.ones(S)  # one term, c = 1 on every system
        d8, r8, e8 = oracle.core_energy(xi, 8, one, 1, amp)
        d7, r7, e7 = oracle.core_energy(xi, 7, np.zeros(1), 0, amp)
        assert np.allclose(d8, d7, rtol=1e-14, atol=1e-15) and abs(e8 - e7) <= 1e-14 * abs(e7)
        assert abs(r8 - r7) <= 1e-14 * abs(r7)
        dn, _, en = oracle.core_energy(xi, 8, par, nm, None)
        du, _, eu = oracle.core_energy(xi, 8, par, nm, np.ones((S,) + shape))
        assert np.array_equal(dn, du) and en == eu


def test_core_energy_sine_closed_form():
    isf, cn = 0.004, 0.013
    xi = np.array([0.0, 0.25, 0.5, 0.125]).reshape(1, 1, 1, 4).astype(complex)
    dE, eref, ev = oracle.core_energy(xi, 3, np.array([isf, cn]))
    s = np.sin
    expect = [0.0, isf * np.pi, 0.0, isf * np.pi * s(np.pi / 4) + cn * 2 * np.pi * s(np.pi / 2)]
    assert np.allclose(dE.reshape(-1), expect, rtol=1e-14, atol=1e-15)
    dens = isf * s(np.pi * xi.real) ** 2 + cn * s(2 * np.pi * xi.real) ** 2
    assert abs(ev - dens.sum()) < 1e-15 and abs(eref - dens.sum() / 1) < 1e-15  # nx = 1


def _elementwise_fd(xi, flag, par, picks, h=1e-6):
    out = []
    for (a, r) in picks:
        xp, xm = xi.copy(), xi.copy()
        xp.reshape(xi.shape[0], -1)[a, r] += h
        xm.reshape(xi.shape[0], -1)[a, r] -= h
        _, _, ep = oracle.core_energy(xp, flag, par)
        _, _, em = oracle.core_energy(xm, flag, par)
        out.append((ep - em) / (2 * h))
    return np.array(out)


def test_core_energy_extended_pins():
    """core_energy_extended (fft_fftw_slab.cpp:1243-1354), coupled gamma surface
    of the 3 systems of each plane: finite differences of the plane density
    per element, cyclic symmetry of the three systems, closed form at zero
    field, integer-shift invariance, E_core not accumulated."""
    rng = np.random.default_rng(5)
    par = rng.normal(size=7) * 0.01
    xi = (rng.random((6, 4, 4, 4)) * 2 - 0.5).astype(complex)
    dE, eref, ev = oracle.core_energy(xi, 2, par)
    assert eref == 0.0
    picks = [(a, int(r)) for a in range(6) for r in rng.integers(0, 64, 3)]
    fd = _elementwise_fd(xi, 2, par, picks)
    got = np.array([dE.reshape(6, -1)[a, r] for a, r in picks])
    assert np.allclose(fd, got, rtol=1e-7, atol=1e-9 * np.abs(got).max())
    # cyclic relabelling of the systems of each plane relabels dE
    perm = [1, 2, 0, 4, 5, 3]
    dEp, _, evp = oracle.core_energy(xi[perm], 2, par)
    assert np.allclose(dEp, dE[perm], rtol=1e-13, atol=1e-15) and abs(evp - ev) < 1e-12 * abs(ev)
    # zero field: density c0 + 3 c1 + 3 c2 + 3 c3 + 6 c4 per plane and cell, no force
    dE0, _, ev0 = oracle.core_energy(np.zeros_like(xi), 2, par)
    c0, c1, c2, c3, c4 = par[:5]
    assert abs(ev0 - 2 * 64 * (c0 + 3 * c1 + 3 * c2 + 3 * c3 + 6 * c4)) < 1e-12
    assert np.all(np.abs(dE0) < 1e-15)
    dE1, _, _ = oracle.core_energy(xi + 1.0, 2, par)
    assert np.allclose(dE1, dE, rtol=1e-9, atol=1e-12)


def test_core_energy_fourier_and_pyrII_as_executed():
    """core_energy_pyrII (:1429-1493): flag 9 = the Fourier series at every
    cell (finite differences, closed form at 0); flag 4 = the routine as
    executed, whose misplaced dE_core assignment (:1491) leaves the force only
    on the last cell of each system, with E_core not accumulated."""
    rng = np.random.default_rng(9)
    par = rng.normal(size=9) * 0.01
    xi = (rng.random((2, 4, 4, 4)) * 2 - 0.5).astype(complex)
    d9, e9ref, ev9 = oracle.core_energy(xi, 9, par)
    picks = [(a, int(r)) for a in range(2) for r in rng.integers(0, 64, 4)]
    fd = _elementwise_fd(xi, 9, par, picks)
    got = np.array([d9.reshape(2, -1)[a, r] for a, r in picks])
    assert np.allclose(fd, got, rtol=1e-7, atol=1e-9 * np.abs(got).max())
    d0, _, ev0 = oracle.core_energy(np.zeros_like(xi), 9, par)
    m = np.arange(1, 5)
    assert abs(ev0 - 128 * (par[0] + par[1:5].sum())) < 1e-12
    assert np.allclose(d0, 2 * np.pi * np.sum(m * par[5:9]), rtol=1e-14)
    d4, e4ref, ev4 = oracle.core_energy(xi, 4, par)
    assert e4ref == 0.0 and ev4 == ev9
    flat4, flat9 = d4.reshape(2, -1), d9.reshape(2, -1)
    assert np.all(flat4[:, :-1] == 0.0) and np.array_equal(flat4[:, -1], flat9[:, -1])


# ------------------------------------------------------------ the full step
def _energy(case, xi):
    st = P.oracle_state(dict(case, xi0=xi), case["_BB"])
    return st.step(case["tau"])["E_total"]


def test_step_is_gradient_flow_of_total_energy():
    case = P.make_case(shape=(8, 8, 8), noise=1e-2)
    case["_BB"], _, _ = P.oracle_B(case)
    st = P.oracle_state(case, case["_BB"])
    x0 = st.xi.copy()
    st.step(case["tau"])
    force = (x0 - st.xi) / case["cddt"]  # = dE/dRe xi + i dE/dIm xi
    rng = np.random.default_rng(7)
    h = 1e-6
    for _ in range(4):
        a, i, j, k = (rng.integers(0, v) for v in (3, 8, 8, 8))
        for comp in (1.0, 1j):
            xp, xm = x0.copy(), x0.copy()
            xp[a, i, j, k] += h * comp
            xm[a, i, j, k] -= h * comp
            fd = (_energy(case, xp) - _energy(case, xm)) / (2 * h)
            got = force[a, i, j, k].real if comp == 1.0 else force[a, i, j, k].imag
            assert abs(fd - got) < 2e-6 * max(1.0, abs(got)), (fd, got)


def test_energy_decreases_at_constant_stress():
    case = P.make_case(shape=(16, 16, 8), noise=1e-3, cddt=0.05)
    st = P.oracle_state(case)
    e = [st.step(case["tau"])["E_total"] for _ in range(30)]
    assert np.all(np.diff(e) < 0)


def test_step_invariant_under_unit_shift_and_translation():
    case = P.make_case(shape=(8, 8, 8), noise=1e-2)
    BB, _, _ = P.oracle_B(case)
    a = P.oracle_state(case, BB)
    b = P.oracle_state(dict(case, xi0=case["xi0"] + 1.0), BB)
    c = P.oracle_state(dict(case, xi0=np.roll(case["xi0"], (2, 3, 1), axis=(1, 2, 3))), BB)
    for _ in range(3):
        a.step(case["tau"]), b.step(case["tau"]), c.step(case["tau"])
    assert np.max(np.abs((b.xi - 1.0) - a.xi)) < 1e-12
    assert np.max(np.abs(np.roll(a.xi, (2, 3, 1), axis=(1, 2, 3)) - c.xi)) < 1e-12


def test_mask_quirk_q3_hand_example():
    # one system, 1x1x4 grid, cell 2 frozen: xinorm picks up (xi_2 - old xi_1)^2
    xi = np.array([0.1, 0.2, 0.7, 0.3], dtype=complex).reshape(1, 1, 1, 4)
    xo = np.array([0.0, 0.0, 1.0, 0.0]).reshape(1, 1, 1, 4)
    BB = np.zeros((1, 1, 1, 1, 4))
    st = oracle.GLState(xi, BB, 0.1, 0, None, 0, None, xo)
    tau = 0.5
    out = st.step([tau])
    new = xi.reshape(-1).real.copy()
    for q in (0, 1, 3):
        new[q] += 0.1 * tau
    expect = (0.05 ** 2) * 3 + (0.7 - 0.2) ** 2
    assert abs(out["xinorm"] - math.sqrt(expect)) < 1e-15
    assert np.allclose(st.xi.reshape(-1).real, new, atol=1e-16)
    assert abs(out["xiave"] - new.sum()) < 1e-15


# --------------------------------------------------------------- diagnostics
def test_diagnostics_averages_and_equilibrium():
    case = P.make_case(shape=(8, 8, 8), noise=1e-2, sigma_mu=0.004)
    m = case["mat"]
    BB, eps, mod = P.oracle_B(case)
    st = P.oracle_state(case, BB)
    st.step(case["tau"])
    FF = oracle.fmatrix(case["shape"], m["C44"], m["C12"], mod["mu"], mod["xnu"], eps, case["geo"]["four"])
    d = oracle.diagnostics(st.data_fftw, st.xi, FF, eps, m, mod["mu"], case["sigma_rot"], fields=True)
    # analytic averages: sym(S:sigma mu + sum_a eps_a <Re xi_a>) and sigma mu
    Sc = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    Sc[i, j, k, l] = m["S44"] / 4 * ((i == k) * (j == l) + (i == l) * (j == k)) + \
                        m["S12"] * (i == j) * (k == l)
    mean_xi = st.xi.real.mean(axis=(1, 2, 3))
    ep = np.einsum("ijkl,kl->ij", Sc, case["sigma_rot"] * mod["mu"]) + np.einsum("aij,a->ij", eps, mean_xi)
    assert np.allclose(d["ave_eps"], (ep + ep.T) / 2, rtol=1e-12, atol=1e-18)
    assert np.allclose(d["ave_sigma"], case["sigma_rot"] * mod["mu"], rtol=1e-10, atol=1e-6)
    # With strain and eigenstrain taken from the same field (the initial output
    # step, App::setup: data_fftw = FFT(xi)), the stress field is in
    # equilibrium: sum_j f_j sigmahat_ij(k) = 0 off k = 0 and the Nyquist planes.
    xh = np.stack([oracle.fft3d(x, -1) for x in st.xi])
    d = oracle.diagnostics(xh, st.xi, FF, eps, m, mod["mu"], case["sigma_rot"], fields=True)
    sig = d["data_sigma"].real
    sh = np.stack([np.stack([np.fft.fftn(sig[j, i]) for j in range(3)]) for i in range(3)])  # [i][j]
    f = np.stack(oracle.frec(*case["shape"], case["geo"]["four"]))
    div = np.einsum("ij...,j...->i...", sh, f)
    keep = np.ones(case["shape"], bool)
    keep[0, 0, 0] = False
    keep[4], keep[:, 4], keep[:, :, 4] = False, False, False
    assert np.max(np.abs(div[:, keep])) < 1e-9 * np.max(np.abs(sh[:, :, keep]))
    # and the quirk-Q2 output step (pre-update strain) is not: the check has teeth
    d2 = oracle.diagnostics(st.data_fftw, st.xi, FF, eps, m, mod["mu"], case["sigma_rot"], fields=True)
    sh2 = np.stack([np.stack([np.fft.fftn(d2["data_sigma"].real[j, i]) for j in range(3)]) for i in range(3)])
    div2 = np.einsum("ij...,j...->i...", sh2, f)
    assert np.max(np.abs(div2[:, keep])) > 1e-6 * np.max(np.abs(sh2[:, :, keep]))


# -------------------------------------------------------------------- RANMAR
def test_ranmar_published_test_values():
    # Marsaglia & Zaman test: seeds ij=1802, kl=9373; numbers 20001..20006 * 4096^2
    seed = 1802 * 30082 + 9373 + 1
    r = oracle.Ranmar(seed)  # RanMars::init already drew number 1
    for _ in range(19999):
        r.uniform()
    got = [r.uniform() * 4096.0 * 4096.0 for _ in range(6)]
    assert got == [6533892.0, 14220222.0, 7275067.0, 6172232.0, 8354498.0, 10633180.0]


def _rotation(axis, ang):
    a = np.asarray(axis, dtype=float)
    a /= np.linalg.norm(a)
    K = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return np.eye(3) + math.sin(ang) * K + (1.0 - math.cos(ang)) * K @ K


def test_rotated_crystal_resolved_shear_relation():
    """Pins the rotation conventions of rotate_stress, set_scale and the
    slip-vector rotation together, in the oracle and in the host layer.
    From the reference's own formulas, for a rotation R given as the
    primitive vectors (the columns of prim, fft.cpp:104-132):
      * four = prim^-1 = R^T (the cofactor formula of set_scale, app.cpp:411-421);
      * the slip vectors become four n, four b (app_fcc.cpp:288-313);
      * sigma_rot = prim sigma prim^-1 = R sigma R^T (fft_fftw_slab.cpp:252-261, |det| = 1).
    Hence tau(R, sigma) = (R^T b).(R sigma R^T)(R^T n) = tau(I, R^2 sigma R^2^T): the
    stress acts as if rotated twice. A transposed convention in any one of
    the four pieces gives tau(I, sigma) or tau(I, R^T^2 sigma R^2) instead,
    which the second assertion tells apart."""
    import pfdd_setup
    rng = np.random.default_rng(11)
    A = rng.standard_normal((3, 3))
    sigma = 0.01 * (A + A.T)
    xn0, xb0 = P.FCC_XN.copy(), P.FCC_XB.copy()

    def tau_oracle(R, sig):
        geo = P.geometry(tuple(map(tuple, R.T)))
        xn, xb = oracle.rotate_slip(xn0, xb0, geo["four"], geo["normfour"])
        return np.asarray(P.tau_of(geo, sig, xn, xb)[0])

    def tau_host(R, sig):
        return pfdd_setup.resolved_shear("fcc", sig, P.DSLIP, prim=R.T.reshape(9))[0]

    I = np.eye(3)
    for trial in range(4):
        R = _rotation(rng.standard_normal(3), rng.uniform(0.2, 3.0))
        R2 = R @ R
        for tau in (tau_oracle, tau_host):
            t = tau(R, sigma)
            scale = np.max(np.abs(t))
            assert np.max(np.abs(t - tau(I, R2 @ sigma @ R2.T))) <= 1e-13 * scale
            assert np.max(np.abs(t - tau(I, sigma))) > 1e-3 * scale


def test_resolss_non_schmid_closed_forms():
    """resolSS_non_Schmid (fft_fftw_slab.cpp:774-845) against the projection
    tensor written as outer products, S = b (x) n + w1 b (x) n' + w2 (n x b) (x) n
    + w3 (n' x b) (x) n', with n, n' the MRSS planes at angle + 120 k and +60 deg."""
    rng = np.random.default_rng(17)
    xb = rng.normal(size=(3, 3))
    srot = rng.normal(size=(3, 3))
    for angle, w in ((0.0, (0.0, 0.0, 0.0)), (27.5, (0.3, -0.2, 0.45))):
        tau = oracle.resolss_non_schmid(srot, xb, P.DSLIP, angle, *w)
        for k in range(3):
            th = np.radians(angle + 120.0 * k)
            n = np.array([-np.sin(th), 0.0, np.cos(th)])
            npr = np.array([-np.sin(th + np.pi / 3), 0.0, np.cos(th + np.pi / 3)])
            b = xb[k]
            S = (np.outer(b, n) + w[0] * np.outer(b, npr) + w[1] * np.outer(np.cross(n, b), n)
                 + w[2] * np.outer(np.cross(npr, b), npr))
            assert abs(tau[k] - np.sum(srot * S) / P.DSLIP) < 1e-13 * (1 + abs(tau[k]))
    # w = 0, angle 0: system 0 is the Schmid factor on the plane z
    tau0 = oracle.resolss_non_schmid(srot, xb, P.DSLIP, 0.0, 0, 0, 0)[0]
    assert abs(tau0 - xb[0] @ srot[:, 2] / P.DSLIP) < 1e-14


def test_theta_character_angle():
    """gradient() (fft_fftw_slab.cpp:869-1065): edge segment -> pi/2, screw ->
    0 or pi, no line -> -1, and the in-place transform to global axes (quirk
    Q6) on a sheared reciprocal matrix, by hand."""
    xi = np.zeros((1, 8, 8, 4), dtype=complex)
    xi[0, :4] = 1.0  # slipped for i < 4 on every plane: lines along y at i = 3.5 and 7.5
    n = np.array([[0.0, 0.0, 1.0]])
    th = oracle.theta(xi, np.eye(3), n, np.array([[1.0, 0.0, 0.0]]))
    assert np.allclose(th[0, [3, 4, 7, 0]], np.pi / 2)
    assert np.all(th[0, [1, 2, 5, 6]] == -1.0)
    ths = oracle.theta(xi, np.eye(3), n, np.array([[0.0, 1.0, 0.0]]))
    assert np.all(np.isclose(ths[0, [3, 4]], 0.0) | np.isclose(ths[0, [3, 4]], np.pi))
    # four = [[1, s, 0], [0, 1, 0], [0, 0, 1]]: at i = 3 the raw gradient is (-1/2, 0, 0);
    # in place: gx' = gx, gy' = gx' s, gz' = 0, so tang = (gy', -gx', 0)
    s_ = 0.5
    four = np.array([[1.0, s_, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    b = np.array([1.0, 0.0, 0.0])
    thq = oracle.theta(xi, four, n, b[None])
    gx = -0.5
    t = np.array([gx * s_, -gx, 0.0])
    t /= np.linalg.norm(t)
    want = np.arccos(t @ b)
    if np.cross(b, t) @ n[0] < 0:
        want = np.pi - want
    assert np.allclose(thq[0, 3], want, rtol=1e-14)


def test_oracle_reproduces_committed_fixtures():
    """tests/golden/gl_fixtures.npz (written by tests/golden/make_fixtures.py):
    the seeded recipe still gives the same initial fields (checksums) and the
    oracle still gives the committed fields and scalars."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_fixtures as F
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gl_fixtures.npz"))
    for name, (kw, steps) in F.cases().items():
        c = P.make_case(**kw)
        assert np.allclose(F.checksum(c["xi0"]), fx[name + "/xi0_checksum"], rtol=1e-13, atol=0), name
        _, xi, r = F.run_oracle(kw, steps)
        ref = fx[name + "/xi"]
        assert np.max(np.abs(xi - ref)) / np.max(np.abs(ref)) < 1e-12, name
        sc = fx[name + "/scalars"]
        assert int(sc[0]) == steps
        for v, w in zip((r["xinorm"], r["xiave"], r["E_total"]), sc[1:]):
            assert abs(v - w) <= 1e-12 * max(abs(w), 1.0), (name, v, w)


def test_c0_fixture_recipe_unchanged():
    """The committed 2000-state oracle run of the 64^3 x 3 ramp
    (tests/golden/c0_ramp_64cubed.npz) was made from tests/problems.py's
    seeded recipe; the recipe must still give the same initial field, and the
    stored last stress state must be the ramp's (1999 increments on xz)."""
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_fullsize_fixtures as F
    fx = np.load(os.path.join(here, "golden", "c0_ramp_64cubed.npz"))
    case = F.c0_case()
    got = [np.sum(case["xi0"].real), np.sum(np.abs(case["xi0"]) ** 2)]
    assert np.allclose(fx["xi0_checksum"], got, rtol=1e-14, atol=0)
    _, srot = F.tau_of_state(case, F.STATES - 1)
    assert np.allclose(fx["sigma_rot_last"], srot, rtol=1e-15, atol=0)
    assert abs(srot[0, 2] - (F.STATES - 1) * F.DELTA_MU) < 1e-18
    assert fx["scalars"].shape == (F.STATES // F.EVERY, 8) and fx["scalars"][-1, 0] == F.STATES
