"""GPU parity of the HIP GL step against the FP64 CPU oracle, through the C ABI.

Tolerances (floating point, written per test):
  FP64 plain path: xi, energies and xinorm within 1e-9 relative of the oracle
                   after the stated number of iterations;
  FP32 mode:       within 1e-5 relative.
Relative error of a field = max|gpu - ref| / max|ref|.
"""
import sys

import numpy as np
import pytest

import oracle
import problems as P

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def scal_rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def run_pair(case, nsteps, precision=0, tau_seq=None, check_every=None, tol_field=1e-9,
             tol_scal=1e-9, stored_B=False):
    st = P.oracle_state(case)
    g = P.gpu_solver(case, precision, stored_B)
    try:
        for s in range(nsteps):
            tau = case["tau"] if tau_seq is None else tau_seq[s]
            if tau_seq is not None:
                g.set_tau(tau)
            r_ref = st.step(tau)
            r_gpu = g.step(1)
            if check_every and (s % check_every == 0 or s == nsteps - 1):
                for key, gk in (("xinorm", "xinorm"), ("xiave", "xiave"), ("E_el", "e_elastic"),
                                ("E_core", "e_core"), ("E_stress", "e_stress"),
                                ("E_total", "e_total"), ("E_core_ref", "e_core_ref")):
                    if abs(r_ref[key]) > 1e-12:
                        assert scal_rel(r_gpu[gk], r_ref[key]) < tol_scal, (s, key, r_gpu[gk], r_ref[key])
        xg = g.get_xi()
        assert relerr(xg, st.xi) < tol_field, relerr(xg, st.xi)
        return st, g, r_ref, r_gpu
    finally:
        g.close()


def test_one_step_basal3_16():
    _say("case: 16^3 x 3 basal, sine core, 1 step, FP64")
    case = P.make_case(shape=(16, 16, 16), noise=1e-2)
    run_pair(case, 1, check_every=1, tol_field=1e-12, tol_scal=1e-11)


def test_loop_basal3_32_200steps():
    _say("case: 32^3 x 3 basal loop r=6 + noise, 200 steps, FP64")
    case = P.make_case(shape=(32, 32, 32), loops=[(0, 2, 16, 16, 16, 6)], noise=1e-3)
    run_pair(case, 200, check_every=50)


@pytest.mark.parametrize("stored_B", [False, True])
def test_fcc12_16_20steps(stored_B):
    _say(f"case: 16^3 x 12 fcc, 20 steps, FP64, stored_B={stored_B}")
    case = P.make_case(shape=(16, 16, 16), systems="fcc12", noise=1e-3, sigma_mu=0.008,
                       loops=[(0, 2, 8, 8, 8, 4), (5, 0, 4, 8, 8, 3)])
    run_pair(case, 20, check_every=5, stored_B=stored_B)


def test_stored_B_fp32_and_skewed():
    _say("case: 16x16x32 x 3 skewed grid, stored packed B, FP32, 30 steps")
    case = P.make_case(shape=(16, 16, 32), noise=1e-2,
                       prim_vectors=((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1)))
    run_pair(case, 30, precision=1, check_every=10, tol_field=1e-5, tol_scal=1e-5, stored_B=True)


@pytest.mark.parametrize("shape", [(48, 16, 32), (16, 96, 8), (8, 32, 48), (40, 16, 80), (16, 160, 8),
                                   (192, 8, 16), (80, 40, 8), (256, 8, 8), (512, 8, 8)])
def test_mixed_radix_grids(shape):
    _say(f"case: {shape} x 3 basal, 10 steps, FP64")
    nx, ny, nz = shape
    case = P.make_case(shape=shape, noise=1e-2, loops=[(1, 2, nz // 2, nx / 2, ny / 2, 3)])
    run_pair(case, 10, check_every=3)


@pytest.mark.parametrize("shape,precision", [((64, 64, 16), 0), ((64, 64, 16), 1), ((16, 64, 64), 0),
                                             ((16, 64, 64), 1), ((8, 512, 16), 0), ((8, 512, 16), 1),
                                             ((8, 16, 512), 0), ((8, 16, 512), 1), ((16, 256, 8), 0),
                                             ((8, 8, 256), 1)])
def test_line_lengths_of_the_large_configs(shape, precision):
    """The line lengths of the 64^3 configurations (64 on every axis) and of
    the 512 x 512 x 256 grid (512 along y and z, 256 along y and z), which
    the other parity shapes put on x only: every pass kernel at those
    lengths in both precisions."""
    _say(f"case: {shape} x 3 basal, 10 steps, precision {precision}")
    nx, ny, nz = shape
    case = P.make_case(shape=shape, noise=1e-2, loops=[(0, 2, nz // 2, nx / 2, ny / 2, min(nx, ny) // 3)])
    tol = 1e-9 if precision == 0 else 1e-5
    run_pair(case, 10, precision=precision, check_every=3, tol_field=tol, tol_scal=tol)


@pytest.mark.parametrize("precision", [0, 1])
def test_skewed_masked_256_z_lines(precision):
    """Skewed (60 deg) grid with a pinned segment, 256 cells along z: the
    update kernel's 256-point inverse transform with the obstacle mask."""
    _say(f"case: 16x16x256 x 3 skewed, mask, 12 steps, precision {precision}")
    shape = (16, 16, 256)
    xo = np.zeros((3,) + shape)
    xo[:, 3, 7:9, 128] = 1.0
    xo[:, 12, 7:9, 128] = 1.0
    case = P.make_case(shape=shape, noise=1e-2, sigma_mu=0.01, xo=xo, loops=[(0, 2, 128, 8, 8, 5)],
                       prim_vectors=((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1)))
    tol = 1e-9 if precision == 0 else 1e-5
    run_pair(case, 12, precision=precision, check_every=4, tol_field=tol, tol_scal=tol)


@pytest.mark.parametrize("nx", [64, 96, 128])
@pytest.mark.parametrize("precision", [0, 1])
def test_fcc12_x_pass_tiles(nx, precision):
    """12 systems with 64-, 96- and 128-point x lines: the x pass at its
    12-system tiles (48 lines at 64 points; 24 lines at 96 and 128, which in
    FP64 use the swizzled tile layout, through the generic addressing of the
    radix-8/4/3 stages at 96 and the folded one of the radix-16/8 stages at
    128)."""
    _say(f"case: {nx}x16x16 x 12 fcc, 8 steps, precision {precision}")
    case = P.make_case(shape=(nx, 16, 16), systems="fcc12", noise=1e-2, sigma_mu=0.004,
                       loops=[(0, 2, 8, nx // 2, 8, 5), (7, 0, 20, 8, 8, 4)])
    tol = 1e-9 if precision == 0 else 1e-5
    run_pair(case, 8, precision=precision, check_every=4, tol_field=tol, tol_scal=tol)


@pytest.mark.parametrize("shape", [(48, 16, 16), (80, 16, 8), (128, 16, 8), (256, 8, 8)])
def test_radix16_x_lengths_fp32(shape):
    """The x pass takes a radix-16 first stage at these lengths; FP32 mode
    against the FP64 oracle at the FP32 tolerance."""
    _say(f"case: {shape} x 3 basal, 10 steps, FP32 (radix-16 x pass)")
    nx, ny, nz = shape
    case = P.make_case(shape=shape, noise=1e-2, loops=[(1, 2, nz // 2, nx / 2, ny / 2, 3)])
    run_pair(case, 10, precision=1, check_every=3, tol_field=1e-5, tol_scal=1e-5)


def test_skewed_grid_masked_frank_read_like():
    _say("case: 32x32x16 skewed (60 deg) grid, mask, 30 steps, FP64")
    shape = (32, 32, 16)
    S = 3
    xo = np.zeros((S,) + shape)
    xo[:, :, 1:16, :] = 1.0          # frozen band (quirk Q3 path)
    xo[:, 4, 20, 8] = 1.0            # isolated pinning points
    xo[:, 28, 20, 8] = 1.0
    case = P.make_case(shape=shape, noise=1e-3, sigma_mu=0.01, xo=xo,
                       prim_vectors=((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1)),
                       loops=[(0, 2, 8, 16, 20, 6)])
    st, g, r_ref, r_gpu = run_pair(case, 30, check_every=10)


def test_series_core_and_ramp():
    _say("case: 16x32x16 x 3, 4-term sine series, stress ramp, 40 steps, FP64")
    case = P.make_case(shape=(16, 32, 16), noise=1e-2, core="series", ncoef=4)
    taus = [case["tau"] * (s / 40.0) for s in range(40)]
    run_pair(case, 40, tau_seq=taus, check_every=10)


def test_fp32_mode_matches_fp64_oracle():
    _say("case: 32^3 x 3 basal loop, 200 steps, FP32 mode vs FP64 oracle")
    case = P.make_case(shape=(32, 32, 32), loops=[(0, 2, 16, 16, 16, 6)], noise=1e-3)
    run_pair(case, 200, precision=1, check_every=50, tol_field=1e-5, tol_scal=1e-5)


def test_device_convergence_predicate():
    _say("case: 16^3 x 3, tol-driven early stop")
    case = P.make_case(shape=(16, 16, 16), noise=1e-3, sigma_mu=0.0)
    st = P.oracle_state(case)
    norms = []
    for _ in range(60):
        norms.append(st.step(case["tau"])["xinorm"])
    tol = norms[25] * 1.0000001  # first iteration below tol is #26 (index 25) if decreasing
    first = next(i for i, v in enumerate(norms) if v < tol)
    g = P.gpu_solver(case)
    try:
        r = g.step(60, tol=tol)
        assert r["converged"] == 1 and r["steps_done"] == first + 1
        st2 = P.oracle_state(case)
        for _ in range(first + 1):
            st2.step(case["tau"])
        assert relerr(g.get_xi(), st2.xi) < 1e-9
    finally:
        g.close()


def test_diag_averages_match_oracle():
    _say("case: 16^3 x 3 diagnostics averages after 5 steps")
    case = P.make_case(shape=(16, 16, 16), noise=1e-2, sigma_mu=0.004)
    m = case["mat"]
    BB, eps, mod = P.oracle_B(case)
    st = P.oracle_state(case, BB)
    g = P.gpu_solver(case)
    try:
        for _ in range(5):
            st.step(case["tau"])
            g.step(1)
        FF = oracle.fmatrix(case["shape"], m["C44"], m["C12"], mod["mu"], mod["xnu"], eps,
                            case["geo"]["four"])
        d = oracle.diagnostics(st.data_fftw, st.xi, FF, eps, m, mod["mu"], case["sigma_rot"])
        ae, asg = g.diag(case["sigma_rot"])
        scale_e = np.max(np.abs(d["ave_eps"]))
        scale_s = np.max(np.abs(d["ave_sigma"]))
        # absolute floor: components that are round-off in the reference
        assert np.max(np.abs(ae - d["ave_eps"])) < 1e-9 * scale_e
        assert np.max(np.abs(asg - d["ave_sigma"])) < 1e-9 * scale_s
        sums = g.xi_sums()
        assert np.allclose(sums, st.xi.real.sum(axis=(1, 2, 3)), rtol=1e-11, atol=1e-9)
    finally:
        g.close()


@pytest.mark.parametrize("precision,steps,skew", [(0, 0, False), (0, 4, False), (0, 3, True), (1, 4, False)])
def test_diag_fields_match_oracle(precision, steps, skew):
    _say(f"case: 16x16x32 x 3 output-step strain/stress fields, precision {precision}, {steps} steps, skew {skew}")
    prim = ((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1)) if skew else ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    case = P.make_case(shape=(16, 16, 32), noise=1e-2, sigma_mu=0.004, prim_vectors=prim,
                       loops=[(0, 2, 16, 8, 8, 5)])
    m = case["mat"]
    BB, eps, mod = P.oracle_B(case)
    st = P.oracle_state(case, BB)
    g = P.gpu_solver(case, precision)
    try:
        for _ in range(steps):
            st.step(case["tau"])
            g.step(1)
        FF = oracle.fmatrix(case["shape"], m["C44"], m["C12"], mod["mu"], mod["xnu"], eps,
                            case["geo"]["four"])
        pre = st.data_fftw if steps else np.stack([oracle.fft3d(x, -1) for x in st.xi])
        d = oracle.diagnostics(pre, st.xi, FF, eps, m, mod["mu"], case["sigma_rot"], fields=True)
        ae, asg, de, ds = g.diag_fields(case["sigma_rot"])
        tol = 1e-9 if precision == 0 else 1e-5
        assert relerr(de, d["data_eps"]) < tol, relerr(de, d["data_eps"])
        assert relerr(ds, d["data_sigma"]) < tol, relerr(ds, d["data_sigma"])
        assert np.max(np.abs(ae - d["ave_eps"])) < tol * np.max(np.abs(d["ave_eps"]))
    finally:
        g.close()


def test_nonfinite_is_reported():
    _say("case: NaN in the initial field -> PFDD_ERR_NONFINITE")
    import pfdd_gl
    case = P.make_case(shape=(8, 8, 8), noise=0.0)
    case["xi0"][0, 1, 2, 3] = np.nan
    g = P.gpu_solver(case)
    try:
        with pytest.raises(pfdd_gl.PfddError) as ei:
            g.step(3)
        assert ei.value.code == 4
    finally:
        g.close()


def test_full_size_c1_properties_fp64():
    """128^3 x 12 (production size): size-independent properties of one GPU
    run: translation covariance and monotone energy decrease at constant stress."""
    _say("case: 128^3 x 12 fcc FP64, 3 steps, roll covariance + energy decrease")
    shape = (128, 128, 128)
    case = P.make_case(shape=shape, systems="fcc12", noise=1e-3, sigma_mu=400e6 / P.MU,
                       loops=[(0, 2, 64, 64, 64, 12), (4, 0, 40, 60, 70, 9)])
    g = P.gpu_solver(case)
    try:
        e = [g.step(1)["e_total"] for _ in range(3)]
        x1 = g.get_xi()
        assert e[0] > e[1] > e[2]
        shift = (5, -7, 3)
        g.set_xi(np.roll(case["xi0"], shift, axis=(1, 2, 3)))
        for _ in range(3):
            g.step(1)
        x2 = g.get_xi()
        assert relerr(x2, np.roll(x1, shift, axis=(1, 2, 3))) < 1e-12
    finally:
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("core,systems,precision,skew", [("extended", "basal3", 0, False),
                                                         ("extended", "fcc12", 0, False),
                                                         ("pyrII", "basal3", 0, False), ("fourier", "basal3", 0, False),
                                                         ("extended", "basal3", 1, False),
                                                         ("fourier", "basal3", 1, False),
                                                         ("bcc", "basal3", 0, False), ("bcc", "basal3", 0, True),
                                                         ("bcc", "basal3", 1, True)])
def test_other_core_models_match_oracle(core, systems, precision, skew):
    print(f"case: core model {core}, 16^3 {systems}, precision {precision}, skew {skew}, 20 steps", flush=True)
    prim = ((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1)) if skew else ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    case = P.make_case(shape=(16, 16, 16), systems=systems, core=core, noise=2e-2, sigma_mu=0.004,
                       loops=[(0, 2, 8, 8, 8, 4)], prim_vectors=prim)
    st = P.oracle_state(case)
    s = P.gpu_solver(case, precision=precision)
    for _ in range(20):
        ref = st.step(case["tau"])
    r = s.step(20)
    x = s.get_xi()
    s.close()
    tol = 1e-9 if precision == 0 else 1e-5
    err = np.max(np.abs(x - st.xi)) / np.max(np.abs(st.xi))
    assert err < tol, err
    assert abs(r["e_core"] - ref["E_core"]) <= tol * max(1.0, abs(ref["E_core"])) * 10
    assert abs(r["e_core_ref"] - ref["E_core_ref"]) <= tol * max(1.0, abs(ref["E_core_ref"]))
    assert abs(r["e_total"] - ref["E_total"]) <= tol * max(1.0, abs(ref["E_total"])) * 10


def test_gpu_matches_committed_fixtures():
    """The GPU path against the committed FP64 fixtures (tests/golden), which
    the oracle reproduces in the CPU suite: fields and scalars within 1e-9."""
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_fixtures as F
    fx = np.load(os.path.join(here, "golden", "gl_fixtures.npz"))
    for name, (kw, steps) in F.cases().items():
        _say(f"case: committed fixture {name}, {steps} steps, FP64")
        c = P.make_case(**kw)
        g = P.gpu_solver(c)
        try:
            r = g.step(steps)
            xi = g.get_xi()
        finally:
            g.close()
        assert relerr(xi, fx[name + "/xi"]) < 1e-9, (name, relerr(xi, fx[name + "/xi"]))
        sc = fx[name + "/scalars"]
        for v, w in zip((r["xinorm"], r["xiave"], r["e_total"]), sc[1:]):
            assert scal_rel(v, w) < 1e-9, (name, v, w)


@pytest.mark.parametrize("precision", [0, 1])
def test_theta_matches_oracle(precision):
    """pfdd_gl_theta (gradient(), fft_fftw_slab.cpp:869-1065) against the
    oracle: of the current field before any step, then of the field before
    the last update.  acos amplifies rounding near |cos| = 1 and the
    |tang| > 0.1 switch can flip on a rounding difference, so a few cells may
    differ; all others agree within 1e-9 (FP64) / 1e-3 (FP32)."""
    _say(f"case: theta of a skewed 32x32x16 x 3 field, precision {precision}")
    prim = ((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1))
    case = P.make_case(shape=(32, 32, 16), noise=2e-2, prim_vectors=prim,
                       loops=[(0, 2, 8, 16, 16, 6), (1, 2, 8, 10, 20, 4)])
    g = P.gpu_solver(case, precision)
    tol = 1e-9 if precision == 0 else 1e-3
    try:
        xi_prev = case["xi0"]
        for nsteps in (0, 3):
            if nsteps:
                g.step(nsteps - 1)
                xi_prev = g.get_xi()
                g.step(1)
            want = oracle.theta(xi_prev, case["geo"]["four"], case["xn"], case["xb"]).reshape(g.theta().shape)
            got = g.theta()
            bad = np.abs(got - want) > tol
            assert bad.mean() < 2e-3, (nsteps, bad.mean())
            assert np.any(want != -1.0) and np.any(want == -1.0)
    finally:
        g.close()


def test_radix5_fcc12_grid():
    _say("case: 40x40x16 x 12 fcc (radix 5 with the two-column x pass), 8 steps, FP64 and FP32")
    case = P.make_case(shape=(40, 40, 16), systems="fcc12", noise=1e-2, sigma_mu=0.004)
    run_pair(case, 8, check_every=4)
    run_pair(case, 8, precision=1, tol_field=1e-5, tol_scal=1e-5)
