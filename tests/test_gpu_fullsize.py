"""GPU parity at the benchmark configurations' own sizes and step counts.

* 64^3 x 3 loop under the 0 -> 300 MPa shear ramp, 2000 stress states of one
  GL iteration (SolveGL::iterate, solve_GL.cpp:70-158, :151-155): through
  the C ABI and through the deck driver (decks/in.loop64-ramp), against the
  committed FP64 oracle run (tests/golden/c0_ramp_64cubed.npz, written by
  tests/golden/make_fullsize_fixtures.py): final field summaries, xinorm /
  xiave / energies every 50 states and the output-step averages, 1e-9.
* 128^3 x 12 production workload (fcc table, 8 seeded loops + noise, 400 MPa):
  FP64 against the oracle run live for 3 iterations, 1e-9; FP32 mode against
  the GPU FP64 path after 1000 iterations, 1e-5 in xi and total energy.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import problems as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "phase-field-dislocation-dynamics-pfdd-2.x_amd")
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_fullsize_fixtures as F  # noqa: E402


def _say(msg):
    print(msg, flush=True)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _fixture():
    return np.load(os.path.join(HERE, "golden", "c0_ramp_64cubed.npz"))


def _check_summaries(xi, fx, tol=1e-9):
    got = F.summarize(xi)
    scale = np.max(np.abs(fx["xi/plane"]))
    e_plane = np.max(np.abs(got["plane"] - fx["xi/plane"])) / scale
    e_sample = np.max(np.abs(got["sample"] - fx["xi/sample"])) / scale
    sums = fx["xi/sums"]  # rows: sum Re xi_a, sum Im xi_a, sum |xi_a|^2
    # scale of the linear sums: sqrt(N sum |xi_a|^2) >= sum |xi_a|
    lin = np.sqrt(np.prod(F.SHAPE) * sums[2])
    e_sums = max(np.max(np.abs(got["sums"][:2] - sums[:2]) / lin),
                 np.max(np.abs(got["sums"][2] - sums[2]) / sums[2]))
    e_chk = np.max(np.abs(got["checksum"] - fx["xi/checksum"]) / np.abs(fx["xi/checksum"]).max())
    _say(f"final field vs oracle: slip plane {e_plane:.3e}, sampled cells {e_sample:.3e}, "
         f"per-system sums {e_sums:.3e}, weighted checksum {e_chk:.3e}")
    assert e_plane < tol and e_sample < tol and e_sums < tol and e_chk < tol


def test_c0_ramp_2000_states_c_abi():
    _say("case: 64^3 x 3 loop, 0 -> 300 MPa ramp over 2000 states (C ABI), FP64 vs committed oracle run")
    fx = _fixture()
    case = F.c0_case()
    assert np.allclose(fx["xi0_checksum"], [np.sum(case["xi0"].real), np.sum(np.abs(case["xi0"]) ** 2)],
                       rtol=1e-14, atol=0)
    g = P.gpu_solver(case)
    rows = {int(r[0]): r for r in fx["scalars"]}
    worst = 0.0
    try:
        srot = None
        for s in range(F.STATES):
            tau, srot = F.tau_of_state(case, s)
            g.set_tau(tau)
            r = g.step(1)
            want = rows.get(s + 1)
            if want is not None:
                for v, w in zip((r["xinorm"], r["xiave"], r["e_elastic"], r["e_core"], r["e_stress"],
                                 r["e_total"], r["e_core_ref"]), want[1:]):
                    if abs(w) > 1e-12:
                        worst = max(worst, abs(v - w) / abs(w))
        _say(f"scalars every {F.EVERY} states: worst relative difference {worst:.3e}")
        assert worst < 1e-9
        _check_summaries(g.get_xi(), fx)
        assert np.allclose(srot, fx["sigma_rot_last"], rtol=1e-15, atol=0)
        ae, asg = g.diag(srot)
        for got, want in ((ae, fx["ave_eps"]), (asg, fx["ave_sigma"])):
            err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            _say(f"output-step average vs oracle: {err:.3e}")
            assert err < 1e-9
    finally:
        g.close()


def test_c0_ramp_deck_driver(tmp_path):
    """The shipped C0 deck through the deck driver (the reference's SolveGL
    loop with the stress ramp and output steps), final field vs the oracle."""
    _say("case: decks/in.loop64-ramp through pfdd_hip (2000 states), final field vs committed oracle run")
    deck = open(os.path.join(PKG, "decks", "in.loop64-ramp")).read()
    deck = deck.replace("run              2000", "write_xi xi_final.bin\nrun              2000")
    assert "write_xi" in deck
    (tmp_path / "in.c0").write_text(deck)
    exe = os.path.join(PKG, "lib", "pfdd_hip")
    r = subprocess.run([exe, "-in", "in.c0", "-log", "log.c0"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    xi = np.fromfile(tmp_path / "xi_final.bin", dtype=np.float64).view(np.complex128).reshape((3,) + F.SHAPE)
    _check_summaries(xi, _fixture())
    assert "Loop time of" in r.stdout


def _threads():
    oracle.set_threads(max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))))


def _c1_solver(pb, precision):
    import pfdd_gl
    s = pfdd_gl.GLSolver(pb["shape"], pb["xn"], pb["xb"], pb["four"], pb["C44"], pb["C12"], pb["C11"], pb["dslip"],
                         pb["cddt"], S12=pb["S12"], S44=pb["S44"], core_model=3, coef=pb["core_coef"],
                         precision=precision)
    s.set_xi(pb["xi0"])
    s.set_tau(pb["tau"])
    return s


def test_c1_three_steps_match_oracle():
    """FP64 within 1e-9 and FP32 within 1e-5 of the FP64 oracle run live for
    3 iterations at 128^3 x 12 (the largest size the oracle runs in a test):
    every reduced scalar per iteration and the final field."""
    import pfdd_setup
    _say("case: 128^3 x 12 fcc production workload, FP64 and FP32, 3 iterations vs the oracle run live")
    _threads()
    pb = pfdd_setup.fcc_loops_problem((128, 128, 128))
    BB, _, _ = oracle.bmatrix(pb["shape"], pb["C44"], pb["C12"], pb["C11"], pb["dslip"], pb["xn"], pb["xb"],
                              pb["four"])
    st = oracle.GLState(pb["xi0"], BB, pb["cddt"], 3, pb["core_coef"])
    del BB
    s = {prec: _c1_solver(pb, prec) for prec in (0, 1)}
    tol = {0: 1e-9, 1: 1e-5}
    worst = {0: 0.0, 1: 0.0}
    try:
        for it in range(3):
            ref = st.step(pb["tau"])
            for prec in (0, 1):
                got = s[prec].step(1)
                for key, gk in (("xinorm", "xinorm"), ("xiave", "xiave"), ("E_el", "e_elastic"),
                                ("E_core", "e_core"), ("E_stress", "e_stress"), ("E_total", "e_total")):
                    e = abs(got[gk] - ref[key]) / abs(ref[key])
                    worst[prec] = max(worst[prec], e)
                    assert e < tol[prec], (prec, it, key, got[gk], ref[key])
        for prec in (0, 1):
            err = relerr(s[prec].get_xi(), st.xi)
            _say(f"{'FP64' if prec == 0 else 'FP32'}: xi rel err after 3 iterations {err:.3e}, "
                 f"worst scalar {worst[prec]:.3e}")
            assert err < tol[prec]
    finally:
        for v in s.values():
            v.close()


def test_c1_fp32_1000_steps_vs_fp64():
    """FP32 mode after the stated 1000 iterations at 128^3 x 12, against the
    GPU FP64 path (itself pinned to the oracle above and at every smaller
    size): relative error in xi and in the total energy below 1e-5."""
    import pfdd_setup
    _say("case: 128^3 x 12 fcc production workload, 1000 iterations, FP32 vs FP64")
    pb = pfdd_setup.fcc_loops_problem((128, 128, 128))
    out = {}
    for prec in (0, 1):
        s = _c1_solver(pb, prec)
        try:
            r = s.step(1000)
            out[prec] = (s.get_xi(), r)
        finally:
            s.close()
    e_xi = relerr(out[1][0], out[0][0])
    e_en = abs(out[1][1]["e_total"] - out[0][1]["e_total"]) / abs(out[0][1]["e_total"])
    _say(f"FP32 vs FP64 after 1000 iterations: xi {e_xi:.3e}, total energy {e_en:.3e}")
    assert e_xi < 1e-5 and e_en < 1e-5
