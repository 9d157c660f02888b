"""The three-kernel plane iteration (2-D FFT of each y-z plane in LDS, x pass
in the plane layout, 2-D inverse FFT fused with the update) against the
five-kernel iteration and the FP64 oracle, through the C ABI.

Every line is transformed with the same radix plan and twiddles in both
iterations and the update is the same code, so the field must be identical
bit for bit; the reductions are summed in another order (1e-12 relative).
Against the oracle: FP64 within 1e-9 relative, FP32 within 1e-5.
"""
import sys

import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu

FIVE = [1, 1, 1, 1, 0]
PLANE = [1, 1, 1, 1, 2]
KEYS = ("xinorm", "xiave", "e_elastic", "e_core", "e_stress", "e_total", "e_core_ref")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _pair(case, nsteps, precision, batch_cases=None):
    out = []
    for choice in (FIVE, PLANE):
        g = P.gpu_solver(case, precision=precision)
        try:
            g.set_tuning(choice)
            r = g.step(nsteps)
            out.append((g.get_xi(), r))
        finally:
            g.close()
    return out


@pytest.mark.parametrize("precision,shape,systems,core", [
    (1, (16, 16, 16), "basal3", "sine"),
    (1, (32, 32, 32), "fcc12", "sine"),
    (1, (48, 48, 48), "basal3", "mpea"),
    (1, (64, 64, 64), "basal3", "sine"),
    (1, (96, 96, 96), "basal3", "series"),
    (1, (32, 80, 80), "basal3", "sine"),
    (1, (128, 128, 128), "fcc12", "sine"),
    (0, (16, 16, 16), "fcc12", "sine"),
    (0, (32, 32, 32), "basal3", "perfect"),
    (0, (48, 48, 48), "basal3", "mpea"),
    (0, (64, 64, 64), "basal3", "sine"),
    (0, (80, 80, 80), "basal3", "sine"),
    (0, (32, 32, 32), "basal3", "extended"),
    (1, (32, 32, 32), "basal3", "fourier"),
    (0, (32, 32, 32), "basal3", "pyrII"),
])
def test_plane_iteration_equals_five_kernel_iteration(precision, shape, systems, core):
    nsteps = 4 if shape[0] == 128 else 12
    _say(f"case: plane vs five-kernel {shape} {systems} {core} precision {precision}, {nsteps} steps")
    case = P.make_case(shape=shape, systems=systems, core=core, noise=1e-2, sigma_mu=0.004)
    (x5, r5), (xp, rp) = _pair(case, nsteps, precision)
    assert np.array_equal(xp, x5), relerr(xp, x5)
    for k in KEYS:
        if abs(r5[k]) > 1e-14:
            assert abs(rp[k] - r5[k]) <= 1e-12 * abs(r5[k]), (k, rp[k], r5[k])


@pytest.mark.parametrize("precision", [0, 1])
def test_plane_iteration_masked_skewed(precision):
    _say(f"case: plane iteration, skewed 32^3 grid with a pinned segment, precision {precision}")
    shape = (32, 32, 32)
    xn, xb = P.basal3()
    xo = np.zeros((3,) + shape)
    xo[:, 8, 10:22, 16] = 1.0  # obstacle cells
    case = P.make_case(shape=shape, noise=1e-2, sigma_mu=0.004, xo=xo,
                       prim_vectors=((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1)))
    (x5, r5), (xp, rp) = _pair(case, 10, precision)
    assert np.array_equal(xp, x5)
    assert abs(rp["xinorm"] - r5["xinorm"]) <= 1e-12 * abs(r5["xinorm"])


def test_plane_iteration_matches_oracle_fp64():
    _say("case: plane iteration 32^3 x 3 FP64, 60 steps vs the oracle (1e-9)")
    case = P.make_case(shape=(32, 32, 32), noise=1e-2)
    st = P.oracle_state(case)
    g = P.gpu_solver(case, precision=0)
    try:
        g.set_tuning(PLANE)
        for _ in range(60):
            rr = st.step(case["tau"])
        rg = g.step(60)
        assert relerr(g.get_xi(), st.xi) < 1e-9
        for kg, kr in (("xinorm", "xinorm"), ("e_elastic", "E_el"), ("e_core", "E_core"), ("e_total", "E_total")):
            assert abs(rg[kg] - rr[kr]) <= 1e-9 * abs(rr[kr]), (kg, rg[kg], rr[kr])
    finally:
        g.close()


def test_plane_iteration_matches_oracle_fp32():
    _say("case: plane iteration 32^3 x 12 FP32, 30 steps vs the FP64 oracle (1e-5)")
    case = P.make_case(shape=(32, 32, 32), systems="fcc12", noise=1e-2, sigma_mu=0.004)
    st = P.oracle_state(case)
    g = P.gpu_solver(case, precision=1)
    try:
        g.set_tuning(PLANE)
        for _ in range(30):
            rr = st.step(case["tau"])
        rg = g.step(30)
        assert relerr(g.get_xi(), st.xi) < 1e-5
        assert abs(rg["e_total"] - rr["E_total"]) <= 1e-5 * abs(rr["E_total"])
    finally:
        g.close()


def test_plane_iteration_batch_members():
    _say("case: plane iteration, batch of 3 members (48^3 x 3, own amplitude each) vs single runs")
    shape = (48, 48, 48)
    cases = [P.make_case(shape=shape, core="mpea", noise=1e-2, sigma_mu=0.002 * (m + 1), seed=100 + m)
             for m in range(3)]
    import pfdd_gl
    c0 = cases[0]
    m = c0["mat"]
    amp = np.concatenate([c["amp"] for c in cases])
    b = pfdd_gl.GLSolver(shape, c0["xn"], c0["xb"], c0["geo"]["four"], m["C44"], m["C12"], m["C11"], m["dslip"],
                         c0["cddt"], S12=m["S12"], S44=m["S44"], precision=1, batch=3, core_model=7, amp=amp)
    try:
        b.set_tuning(PLANE)
        b.set_xi(np.concatenate([c["xi0"] for c in cases]))
        b.set_tau(np.concatenate([c["tau"] for c in cases]))
        b.step(10)
        xb = b.get_xi()
    finally:
        b.close()
    for q, c in enumerate(cases):
        g = P.gpu_solver(c, precision=1)
        try:
            g.set_tuning(FIVE)
            g.step(10)
            assert np.array_equal(xb[3 * q:3 * q + 3], g.get_xi()), q
        finally:
            g.close()


def test_plane_iteration_rejected_where_it_does_not_fit():
    import pfdd_gl
    case = P.make_case(shape=(32, 32, 16), noise=1e-2)
    g = P.gpu_solver(case, precision=1)
    try:
        with pytest.raises(pfdd_gl.PfddError) as ei:
            g.set_tuning(PLANE)
        assert ei.value.code == 2
    finally:
        g.close()
