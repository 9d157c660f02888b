"""Reverse mode of the GL iteration (pfdd_gl_tape_record / _backward).

The reference has no derivatives, so the checks are mathematical:
CPU: the oracle's adjoint (numpy, all states in memory) against central
finite differences of the oracle's forward model (orc_gl_step, the
restatement of solve_GL / update_order_parameter) -- with and without an
obstacle mask.
GPU: the checkpointed HIP adjoint against the oracle adjoint (FP64 1e-9),
against central finite differences of the GPU forward run (1e-6 relative,
above the difference quotient's rounding floor), FP32
against FP64, and the torch.autograd.Function wrapper.
Loss: L = sum |xi_T - target|^2, so dL/dxi_T = 2 (xi_T - target) with Re and
Im independent.
"""
import numpy as np
import pytest

import oracle
import problems as P


def small_case(shape=(8, 8, 8), masked=False, noise=1e-2, sigma_mu=0.005):
    nx, ny, nz = shape
    xo = None
    if masked:
        xo = np.zeros((3,) + tuple(shape))
        xo[0, nx // 4, :, nz // 2] = 1.0
        xo[1, :, ny // 3, nz // 2] = 1.0
    return P.make_case(shape=shape, noise=noise, sigma_mu=sigma_mu, core="series", ncoef=4,
                       loops=[(0, 2, nz // 2, nx / 2, ny / 2, max(2, nx // 4))], xo=xo)


def target_of(case, seed=7):
    rng = np.random.default_rng(seed)
    return case["xi0"] + 0.05 * (rng.standard_normal(case["xi0"].shape)
                                 + 1j * rng.standard_normal(case["xi0"].shape))


def oracle_forward(case, BB, xi0, tau, coef, nsteps):
    st = oracle.GLState(xi0, BB, case["cddt"], 8, np.asarray(coef).reshape(-1), nm=case["nm"],
                        xo=case["xo"])
    for _ in range(nsteps):
        st.step(tau)
    return st.xi


def loss(xi, target):
    return float(np.sum(np.abs(xi - target) ** 2))


EPS = np.finfo(np.float64).eps


def fd_err(g_fd, g_adj, L, h, what, rtol=1e-6):
    """Scale-aware comparison of a central difference with the adjoint:
    |g_fd - g_adj| <= rtol |g_fd| + floor, where floor = 64 eps |L| / h bounds
    the rounding error of the difference quotient (the loss L is an FP64 sum
    whose inputs carry the FP64 rounding of the forward run).  Returns the
    error in units of that tolerance (<= 1 passes) and prints the numbers."""
    floor = 64.0 * EPS * abs(L) / h
    tol = rtol * abs(g_fd) + floor
    err = abs(g_fd - g_adj)
    print(f"  {what}: fd {g_fd:+.10e} adjoint {g_adj:+.10e} rel {err / max(abs(g_fd), 1e-300):.2e} "
          f"floor {floor:.1e}", flush=True)
    return err / tol


def richardson(f, h):
    """Richardson-extrapolated central difference (4 D(h/2) - D(h)) / 3 of
    f at 0, with D(h) = (f(h) - f(-h)) / (2h): truncation O(h^4).  Its
    rounding error is about 3x that of D(h), which fd_err covers when given
    the effective step h / 3."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


@pytest.mark.parametrize("masked", [False, True])
def test_oracle_adjoint_matches_finite_differences(masked):
    case = small_case(masked=masked)
    BB, _, _ = P.oracle_B(case)
    nsteps = 6
    coef = case["core_par"].reshape(case["S"], case["nm"])
    tau = np.asarray(case["tau"]) + np.array([1e-3, -2e-3, 5e-4])
    target = target_of(case)
    xT = oracle_forward(case, BB, case["xi0"], tau, coef, nsteps)
    lam0, gtau, gc, xT2 = oracle.adjoint_rollout(case["xi0"], BB, case["cddt"], tau, coef, nsteps,
                                                 2.0 * (xT - target), xo=case["xo"])
    assert np.array_equal(xT, xT2)
    h = 1e-6
    L0 = loss(xT, target)

    def fd(f):
        return (f(h) - f(-h)) / (2 * h)

    for a in range(3):
        def f(d, a=a):
            t = tau.copy()
            t[a] += d
            return loss(oracle_forward(case, BB, case["xi0"], t, coef, nsteps), target)
        g = fd(f)
        assert fd_err(g, gtau[a], L0, h, f"dL/dtau[{a}]") <= 1.0
    for a in range(3):
        for m in range(case["nm"]):
            def f(d, a=a, m=m):
                c = coef.copy()
                c[a, m] += d
                return loss(oracle_forward(case, BB, case["xi0"], tau, c, nsteps), target)
            g = fd(f)
            assert fd_err(g, gc[a, m], L0, h, f"dL/dc[{a},{m}]") <= 1.0
    rng = np.random.default_rng(3)
    for _ in range(2):
        v = rng.standard_normal(case["xi0"].shape) + 1j * rng.standard_normal(case["xi0"].shape)

        def f(d):
            return loss(oracle_forward(case, BB, case["xi0"] + d * v, tau, coef, nsteps), target)
        g = fd(f)
        pred = float(np.sum(lam0.real * v.real + lam0.imag * v.imag))
        assert fd_err(g, pred, L0, h, "dL/dxi0 . v") <= 1.0
    if masked:
        # frozen cells: their Re/Im carry no tau or core sensitivity, but the
        # elastic coupling reaches them
        assert np.all(np.isfinite(lam0))


def gpu_case_solver(case, precision=0, stored_B=False):
    s = P.gpu_solver(case, precision=precision, stored_B=stored_B)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("masked,every,stored_B", [(False, 5, False), (True, 4, False), (False, 12, True)])
def test_gpu_adjoint_matches_oracle_adjoint(masked, every, stored_B):
    print(f"case: adjoint 16^3 x 3 series nm=4, 12 steps, every={every}, masked={masked}, "
          f"stored_B={stored_B}", flush=True)
    case = small_case(shape=(16, 16, 16), masked=masked)
    BB, _, _ = P.oracle_B(case)
    nsteps = 12
    coef = case["core_par"].reshape(case["S"], case["nm"])
    target = target_of(case)
    s = gpu_case_solver(case, stored_B=stored_B)
    s.tape_record(nsteps, every)
    xT = s.get_xi()
    ref_lam, ref_tau, ref_c, ref_xT = oracle.adjoint_rollout(case["xi0"], BB, case["cddt"], case["tau"], coef,
                                                             nsteps, 2.0 * (xT - target), xo=case["xo"])
    assert np.max(np.abs(xT - ref_xT)) <= 1e-12 * np.max(np.abs(ref_xT))
    lam, dtau, dc = s.tape_backward(2.0 * (xT - target), case["nm"])
    s.close()
    assert np.max(np.abs(lam - ref_lam)) <= 1e-9 * np.max(np.abs(ref_lam))
    assert np.allclose(dtau, ref_tau, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref_tau)))
    assert np.allclose(dc, ref_c, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref_c)))


@pytest.mark.gpu
def test_gpu_adjoint_matches_gpu_finite_differences():
    print("case: adjoint 32^3 x 3 vs central FD of the GPU forward, 40 steps, every 10", flush=True)
    case = small_case(shape=(32, 32, 32), noise=1e-2, sigma_mu=0.006)
    nsteps, every = 40, 10
    coef = case["core_par"].reshape(case["S"], case["nm"])
    target = target_of(case)
    s = gpu_case_solver(case)

    def run(xi0=None, tau=None, c=None):
        s.set_xi(case["xi0"] if xi0 is None else xi0)
        s.set_tau(case["tau"] if tau is None else tau)
        s.set_core_coef((coef if c is None else c).reshape(-1), case["nm"])
        s.step(nsteps)
        return loss(s.get_xi(), target)

    s.set_core_coef(coef.reshape(-1), case["nm"])
    s.set_xi(case["xi0"])
    s.set_tau(case["tau"])
    s.tape_record(nsteps, every)
    xT = s.get_xi()
    lam, dtau, dc = s.tape_backward(2.0 * (xT - target), case["nm"])
    h = 1e-6
    L0 = loss(xT, target)
    worst = 0.0
    for a in range(3):
        def f(d, a=a):
            t = np.array(case["tau"], dtype=float)
            t[a] += d
            return run(tau=t)
        worst = max(worst, fd_err(richardson(f, h), dtau[a], L0, h / 3, f"dL/dtau[{a}]"))
    for a, m in [(0, 0), (1, 1), (2, 3), (0, 2)]:
        def f(d, a=a, m=m):
            c = coef.copy()
            c[a, m] += d
            return run(c=c)
        worst = max(worst, fd_err(richardson(f, h), dc[a, m], L0, h / 3, f"dL/dc[{a},{m}]"))
    rng = np.random.default_rng(11)
    v = rng.standard_normal(case["xi0"].shape) + 1j * rng.standard_normal(case["xi0"].shape)
    pred = float(np.sum(lam.real * v.real + lam.imag * v.imag))
    worst = max(worst, fd_err(richardson(lambda d: run(xi0=case["xi0"] + d * v), h), pred, L0, h / 3,
                              "dL/dxi0 . v"))
    s.close()
    print(f"adjoint vs central FD: worst error {worst:.2f} of the tolerance", flush=True)
    assert worst <= 1.0


@pytest.mark.gpu
def test_gpu_adjoint_fp32_close_to_fp64():
    print("case: adjoint FP32 vs FP64, 32^3 x 3, 30 steps", flush=True)
    case = small_case(shape=(32, 32, 32))
    target = target_of(case)
    out = {}
    for prec in (0, 1):
        s = gpu_case_solver(case, precision=prec)
        s.tape_record(30, 7)
        xT = s.get_xi()
        seed = 2.0 * (xT - target) if prec == 0 else out[0][3]
        lam, dtau, dc = s.tape_backward(seed, case["nm"])
        out[prec] = (lam, dtau, dc, seed)
        s.close()
    (l64, t64, c64, _), (l32, t32, c32, _) = out[0], out[1]
    e_lam = np.max(np.abs(l32 - l64)) / np.max(np.abs(l64))
    e_tau = np.max(np.abs(t32 - t64)) / np.max(np.abs(t64))
    e_c = np.max(np.abs(c32 - c64)) / np.max(np.abs(c64))
    print(f"FP32 adjoint relative error: dxi0 {e_lam:.2e} dtau {e_tau:.2e} dcoef {e_c:.2e}", flush=True)
    assert e_lam < 1e-4 and e_tau < 1e-4 and e_c < 1e-4


@pytest.mark.gpu
def test_gpu_tape_errors():
    print("case: tape error paths", flush=True)
    import pfdd_gl
    case = small_case(shape=(16, 16, 16))
    s = gpu_case_solver(case)
    with pytest.raises(pfdd_gl.PfddError, match="STATE"):
        s.tape_backward(np.zeros_like(case["xi0"]), case["nm"])
    with pytest.raises(pfdd_gl.PfddError, match="ARG"):
        s.tape_record(0, 1)
    s.close()


@pytest.mark.gpu
def test_torch_autograd_function():
    print("case: torch.autograd GLRollout 16^3 x 3, 15 steps, every 4; stress scalar through torch", flush=True)
    import torch
    import pfdd_torch
    case = small_case(shape=(16, 16, 16))
    nsteps, every = 15, 4
    target = target_of(case)
    s = gpu_case_solver(case)
    # tau = sigma * r: resolved shear of a scalar shear stress (linear, as resolSS_Schmid)
    r = torch.tensor(np.asarray(case["tau"]) / 0.005)
    coef0 = case["core_par"].reshape(case["S"], case["nm"])

    def loss_of(sig_val):
        xi0 = torch.tensor(case["xi0"], requires_grad=True)
        sig = torch.tensor(sig_val, dtype=torch.float64, requires_grad=True)
        coef = torch.tensor(coef0, requires_grad=True)
        xT = pfdd_torch.gl_rollout(s, xi0, sig * r, coef, nsteps, every)
        L = (xT - torch.tensor(target)).abs().pow(2).sum()
        return L, xi0, sig, coef

    L, xi0, sig, coef = loss_of(0.005)
    L.backward()
    h = 1e-7
    Lp = loss_of(0.005 + h)[0].item()
    Lm = loss_of(0.005 - h)[0].item()
    g_fd = (Lp - Lm) / (2 * h)
    g = sig.grad.item()
    print(f"dL/dsigma autograd {g:.10e} central FD {g_fd:.10e}", flush=True)
    assert abs(g - g_fd) <= 1e-6 * abs(g_fd)
    # same numbers as the plain C ABI reverse pass
    s.set_core_coef(coef0.reshape(-1), case["nm"])
    s.set_xi(case["xi0"])
    s.set_tau(0.005 * r.numpy())
    s.tape_record(nsteps, every)
    xT = s.get_xi()
    lam, dtau, dc = s.tape_backward(2.0 * (xT - target), case["nm"])
    s.close()
    assert np.allclose(xi0.grad.numpy(), lam, rtol=1e-12, atol=1e-12 * np.max(np.abs(lam)))
    assert np.allclose(coef.grad.numpy(), dc, rtol=1e-12, atol=1e-14)
    assert abs(g - float(np.dot(dtau, r.numpy()))) <= 1e-10 * abs(g)


@pytest.mark.gpu
@pytest.mark.parametrize("core", ["fourier", "pyrII", "extended"])
def test_gpu_adjoint_fourier_models_match_gpu_fd(core):
    """Reverse mode of the Fourier core models (flag 9 on every cell, flag 4
    as executed: force on each system's last cell only) and of the extended
    model (flag 2: the three systems of a plane coupled through one density)
    against central differences of the GPU forward run: the shared
    coefficients (aa_m, bb_m; c0..c4, a1, a3: gradient = sum over systems),
    the stresses and a field direction."""
    print(f"case: adjoint {core} core 16^3 x 3 vs central FD, 16 steps, every 4", flush=True)
    import pfdd_gl
    shape, nsteps, every = (16, 16, 16), 16, 4
    case = P.make_case(shape=shape, noise=1e-2, sigma_mu=0.006, core=core,
                       loops=[(0, 2, 8, 8.0, 8.0, 4)])
    target = target_of(case)
    m = case["mat"]
    par = case["core_par"]

    def solver(c):
        gc = dict(case["gpu_core"], coef=c)
        s = pfdd_gl.GLSolver(shape, case["xn"], case["xb"], case["geo"]["four"], m["C44"], m["C12"], m["C11"],
                             m["dslip"], case["cddt"], S12=m["S12"], S44=m["S44"], **gc)
        s.set_xi(case["xi0"])
        s.set_tau(case["tau"])
        return s

    def run(c=par, xi0=None, tau=None):
        s = solver(c)
        if xi0 is not None:
            s.set_xi(xi0)
        if tau is not None:
            s.set_tau(tau)
        s.step(nsteps)
        out = loss(s.get_xi(), target)
        s.close()
        return out

    s = solver(par)
    s.tape_record(nsteps, every)
    xT = s.get_xi()
    nc, off = (7, 0) if core == "extended" else (8, 1)  # c0..a3 = par[0..6]; aa1..bb4 = par[1..8]
    lam, dtau, dc = s.tape_backward(2.0 * (xT - target), nc)
    s.close()
    assert dc.shape == (3, nc)
    h, worst = 1e-6, 0.0
    L0 = loss(xT, target)
    for q in range(nc):
        def f(d, q=q):
            c = par.copy()
            c[off + q] += d
            return run(c=c)
        worst = max(worst, fd_err(richardson(f, h), dc[:, q].sum(), L0, h / 3, f"dL/dcoef[{q}]"))
    for a in range(3):
        def f(d, a=a):
            t = np.array(case["tau"], dtype=float)
            t[a] += d
            return run(tau=t)
        worst = max(worst, fd_err(richardson(f, h), dtau[a], L0, h / 3, f"dL/dtau[{a}]"))
    rng = np.random.default_rng(5)
    v = rng.standard_normal(case["xi0"].shape) + 1j * rng.standard_normal(case["xi0"].shape)
    pred = float(np.sum(lam.real * v.real + lam.imag * v.imag))
    worst = max(worst, fd_err(richardson(lambda d: run(xi0=case["xi0"] + d * v), h), pred, L0, h / 3,
                              "dL/dxi0 . v"))
    print(f"adjoint vs central FD ({core}): worst error {worst:.2f} of the tolerance", flush=True)
    assert worst <= 1.0
