"""The per-cell GSF amplitude path on the GPU against the FP64 oracle:
core_energy_mpea (core flag 7, fft_fftw_slab.cpp:1196-1235, per-cell file
material.cpp:837-861) and the amp-weighted sine series (flag 8), the core
model of the batched alloy ensemble.

Tolerances (floating point, written per test): FP64 within 1e-9 relative of
the oracle in xi and the energies; FP32 mode within 1e-5.  Relative error of
a field = max|gpu - ref| / max|ref|.
"""
import os
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


def _threads():
    oracle.set_threads(max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))))


def _compare(res_gpu, res_ref, tol):
    for key, gk in (("xinorm", "xinorm"), ("xiave", "xiave"), ("E_el", "e_elastic"), ("E_core", "e_core"),
                    ("E_stress", "e_stress"), ("E_total", "e_total"), ("E_core_ref", "e_core_ref")):
        ref = res_ref[key]
        if abs(ref) > 1e-12:
            assert abs(res_gpu[gk] - ref) <= tol * abs(ref), (key, res_gpu[gk], ref)


@pytest.mark.parametrize("core,precision,masked", [("mpea", 0, False), ("mpea", 1, False), ("mpea", 0, True),
                                                   ("series", 0, False), ("series", 1, True)])
def test_amplitude_core_matches_oracle(core, precision, masked):
    shape = (32, 16, 32)
    _say(f"case: per-cell amplitude, core {core}, {shape} x 3, precision {precision}, masked {masked}, 40 steps")
    _threads()
    xo = None
    if masked:
        xo = np.zeros((3,) + shape)
        xo[:, 3:6, :, 16] = 1.0      # a frozen band on the slip plane (quirk Q3 path)
        xo[1, 20, 9, 16] = 1.0
    amp = P.amplitude_field(shape, 3, seed=31) if core == "series" else None
    case = P.make_case(shape=shape, core=core, ncoef=3, noise=1e-2, sigma_mu=0.004, xo=xo, amp=amp,
                       loops=[(0, 2, 16, 16, 8, 6), (2, 2, 16, 8, 4, 3)])
    assert case["amp"] is not None and case["amp"].std() > 0.05 * case["amp"].mean()
    st = P.oracle_state(case)
    g = P.gpu_solver(case, precision)
    tol = 1e-9 if precision == 0 else 1e-5
    try:
        for s in range(40):
            ref = st.step(case["tau"])
            got = g.step(1)
            if s % 10 == 9:
                _compare(got, ref, tol)
        err = relerr(g.get_xi(), st.xi)
        _say(f"xi rel err {err:.3e}")
        assert err < tol, err
    finally:
        g.close()


def _member_cases(shape, n, core, masked_member=None):
    nx, ny, nz = shape
    out = []
    for m in range(n):
        xo = None
        if m == masked_member:
            xo = np.zeros((3,) + shape)
            xo[:, nx // 3, :, nz // 2] = 1.0
        amp = P.amplitude_field(shape, 3, seed=100 + m) if core == "series" else None
        kw = dict(amp=P.amplitude_field(shape, 3, seed=200 + m)) if core == "mpea" else dict(amp=amp)
        out.append(P.make_case(shape=shape, core=core, ncoef=2, noise=1e-2, sigma_mu=0.002 + 0.0015 * m,
                               seed=576 + 13 * m, xo=xo, loops=[(m % 3, 2, nz // 2, nx / 2 + m, ny / 2 - m, 4)],
                               **kw))
    return out


def _batch_of(cases, precision):
    import pfdd_gl
    c0, m = cases[0], cases[0]["mat"]
    xo = None
    if any(c["xo"] is not None for c in cases):
        xo = np.concatenate([c["xo"] if c["xo"] is not None else np.zeros((3,) + c["shape"]) for c in cases])
    gc = dict(c0["gpu_core"])
    gc["amp"] = np.concatenate([c["amp"] for c in cases])
    if gc["core_model"] == 8:
        # the series coefficients are shared by the members of a batch
        for c in cases[1:]:
            c["core_par"] = c0["core_par"]
    s = pfdd_gl.GLSolver(c0["shape"], c0["xn"], c0["xb"], c0["geo"]["four"], m["C44"], m["C12"], m["C11"],
                         m["dslip"], c0["cddt"], S12=m["S12"], S44=m["S44"], xo=xo, precision=precision,
                         batch=len(cases), **gc)
    s.set_xi(np.concatenate([c["xi0"] for c in cases]))
    s.set_tau(np.concatenate([c["tau"] for c in cases]))
    return s


@pytest.mark.parametrize("core,precision", [("mpea", 0), ("series", 0), ("mpea", 1)])
def test_batch_members_with_own_amplitude_match_oracle(core, precision):
    """A batch of 3 members, each with its own amplitude field, stress,
    initial field (and one with a mask): every member against its own
    oracle run."""
    shape = (16, 32, 16)
    _say(f"case: batch 3 x ({shape} x 3) with per-member amplitude, core {core}, precision {precision}, 25 steps")
    _threads()
    cases = _member_cases(shape, 3, core, masked_member=1)
    b = _batch_of(cases, precision)
    tol = 1e-9 if precision == 0 else 1e-5
    try:
        b.step(25)
        xb = b.get_xi()
    finally:
        b.close()
    for m, c in enumerate(cases):
        st = P.oracle_state(c)
        for _ in range(25):
            st.step(c["tau"])
        err = relerr(xb[3 * m:3 * m + 3], st.xi)
        _say(f"member {m}: xi rel err {err:.3e}")
        assert err < tol, (m, err)
    # the members really differ in their amplitude
    assert not np.allclose(cases[0]["amp"], cases[2]["amp"])


_ENSEMBLE_REF = {}


def _ensemble_reference(groups):
    """Oracle runs of every member of the reduced alloy ensemble (cached)."""
    import pfdd_setup
    key = tuple(groups)
    if key in _ENSEMBLE_REF:
        return _ENSEMBLE_REF[key]
    _threads()
    ens = pfdd_setup.alloy_ensemble(groups=groups, seed=576)
    refs = []
    for g in ens:
        BB, _, _ = oracle.bmatrix(g["shape"], g["C44"], g["C12"], g["C11"], g["dslip"], g["xn"], g["xb"], g["four"])
        S, M = g["S"], g["batch"]
        amp = g["amp"].reshape((M, S) + g["shape"])
        out = []
        for m in range(M):
            st = oracle.GLState(g["xi0"][S * m:S * (m + 1)], BB, g["cddt"], 8, g["coef"], nm=g["ncoef"], amp=amp[m])
            r = None
            for _ in range(20):
                r = st.step(g["tau"][S * m:S * (m + 1)])
            out.append((st.xi.copy(), r))
        refs.append(out)
    _ENSEMBLE_REF[key] = (ens, refs)
    return ens, refs


@pytest.mark.parametrize("precision", [1, 0])
def test_reduced_alloy_ensemble_in_one_graph(precision):
    """The batched alloy ensemble, reduced to 2 realizations each of 48^3,
    64^3 and 96^3 (3 systems, lognormal per-cell amplitude of the sine core,
    straight edge dislocation, stress uniform in [200, 600] MPa): three batch
    handles stepped by one hipGraph for 20 iterations, every realization
    against its own oracle run (FP32 is the ensemble's bulk precision)."""
    import pfdd_gl
    groups = ((48, 2), (64, 2), (96, 2))
    _say(f"case: reduced alloy ensemble {groups} in one hipGraph, 20 steps, precision {precision}")
    ens, refs = _ensemble_reference(groups)
    solvers = []
    try:
        for g in ens:
            s = pfdd_gl.GLSolver(g["shape"], g["xn"], g["xb"], g["four"], g["C44"], g["C12"], g["C11"], g["dslip"],
                                 g["cddt"], S12=g["S12"], S44=g["S44"], core_model=8, coef=g["coef"],
                                 ncoef=g["ncoef"], amp=g["amp"], precision=precision, batch=g["batch"])
            s.set_xi(g["xi0"])
            s.set_tau(g["tau"])
            solvers.append(s)
        graph = pfdd_gl.GLGraph(solvers, iters=4)
        res = graph.run(20)
        graph.close()
        tol = 1e-9 if precision == 0 else 1e-5
        for g, s, r, ref in zip(ens, solvers, res, refs):
            assert r["steps_done"] == 20
            x = s.get_xi()
            S = g["S"]
            for m, (xr, rr) in enumerate(ref):
                err = relerr(x[S * m:S * (m + 1)], xr)
                _say(f"{g['shape'][0]}^3 member {m}: xi rel err {err:.3e}")
                assert err < tol, (g["shape"], m, err)
    finally:
        for s in solvers:
            s.close()


def test_deck_mpea_file_run_matches_oracle(tmp_path):
    """core_energy 7 through the deck driver: `material ... mpea FILE` with
    the reference's file format (header n1 n2 n3 ns, then lines i j k s
    value in mJ/m^2, material.cpp:837-861), 20 stress states of one
    iteration, final field against the oracle."""
    import test_driver as D
    _say("case: driver deck 16^3 x 3, core_energy 7 from an MPEA file, 20 states")
    shape, S = (16, 16, 16), 3
    A = 158.0 * P.amplitude_field(shape, S, seed=77)  # mJ/m^2
    lines = ["16 16 16 3"]
    for s in range(S):
        for i in range(16):
            for j in range(16):
                for k in range(16):
                    lines.append(f"{i} {j} {k} {s} {A[s, i, j, k]:.17g}")
    (tmp_path / "usfe.txt").write_text("\n".join(lines) + "\n")
    s0, ds = "0 0 0 0 0.002 0", "0 0 0 0 0.0001 0"
    text = D.deck_text(max_iter=1, tol=0, run=20, s0=s0, ds=ds, noise=1e-2, extra="write_xi xi_final.bin\n")
    text = text.replace("core_energy 3", "core_energy 7").replace("usf 158e-3", "usf 158e-3 mpea usfe.txt")
    r = D.run_exe(text, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    xi = np.fromfile(tmp_path / "xi_final.bin", dtype=np.float64).view(np.complex128).reshape((S,) + shape)
    case = P.make_case(shape=shape, noise=1e-2, loops=[(0, 2, 8, 8, 8, 4)], sigma_mu=0.0, core="mpea",
                       amp=A * 1e-3 / P.USF)
    st = P.oracle_state(case)
    sigma, dsig, geo = D.sig_from6(D.six(s0)), D.sig_from6(D.six(ds)), case["geo"]
    for _ in range(20):
        srot = oracle.rotate_stress(geo["prim"], geo["invprim"], geo["norvol"], sigma)
        st.step(oracle.resolss(srot, case["xn"], case["xb"], P.DSLIP))
        sigma = sigma + dsig
    err = relerr(xi, st.xi)
    _say(f"xi rel err {err:.3e}")
    assert err < 1e-9, err
