"""The deck driver pfdd_hip (host C++ mirror of SolveGL/FFT/App/Output/Input).

CPU: `-check` prints the setup the host layer derives from a deck (material,
reciprocal lattice, rotated slip systems, resolved shear, core coefficients);
it is compared with the oracle's restatement of the same reference code, and
malformed decks must fail like Error::all (exit 1, "ERROR: ...").
GPU: a deck run end to end against the oracle emulating SolveGL::iterate.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "phase-field-dislocation-dynamics-pfdd-2.x_amd")
EXE = os.path.join(PKG, "lib", "pfdd_hip")

DECK = """seed 576
app_style {app} core_energy 3
fft_style hip pbc 1 1 1 mode 1 primitive x {px} y {py} z {pz}
dimension 3
lattice sc/6n 1
region box block 0 {nx} 0 {ny} 0 {nz}
create_box box
create_sites box
material mu 48e9 young 128.64e9 a 0.3615e-9 isf 41e-3 usf 158e-3
solve_style GL max_iter {max_iter} tol {tol}
scale 1.0 1.0 1.0
sigma initial {s0} delta {ds}
mobility 0.01 1.0
initial noise {noise}
initial disc 0 2 {k0} {c0} {c1} {r}
stats 1000
{extra}run {run}
"""


def deck_text(**kw):
    d = dict(app="basal3", px="1 0 0", py="0 1 0", pz="0 0 1", nx=16, ny=16, nz=16, max_iter=1, tol=0,
             s0="0 0 0 0 0 0", ds="0 0 0 0 0 0", noise=1e-3, k0=8, c0=8, c1=8, r=4, run=10, extra="")
    d.update(kw)
    return DECK.format(**d)


def run_exe(text, tmp_path, check=False):
    f = tmp_path / "in.test"
    f.write_text(text)
    cmd = [EXE, "-in", str(f), "-log", str(tmp_path / "log.pfdd")] + (["-check"] if check else [])
    return subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path, timeout=600)


def parse_check(out):
    res = {}
    for line in out.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "slip":
            res.setdefault("xn", []).append([float(v) for v in w[3:6]])
            res.setdefault("xb", []).append([float(v) for v in w[7:10]])
        elif w[0] in ("four", "tau", "coef"):
            res[w[0]] = np.array([float(v) for v in w[1:]])
        elif w[0] == "material":
            res["material"] = {w[i]: float(w[i + 1]) for i in range(1, len(w), 2)}
    return res


pytestmark_cpu = pytest.mark.skipif(not os.path.exists(EXE), reason="pfdd_hip not built")


def six(s):
    return [float(v) for v in s.split()]


def sig_from6(v):
    s = np.zeros((3, 3))
    s[0, 0], s[1, 1], s[2, 2] = v[0], v[1], v[2]
    s[0, 1] = s[1, 0] = v[3]
    s[0, 2] = s[2, 0] = v[4]
    s[1, 2] = s[2, 1] = v[5]
    return s


@pytestmark_cpu
@pytest.mark.parametrize("app,prim,s0", [
    ("basal3", ("1 0 0", "0 1 0", "0 0 1"), "0 0 0 0 0.004 0"),
    ("fcc", ("1 0 0", "0.5 0.8660254037844386 0", "0 0 1"), "0.001 -0.002 0.0005 0.003 -0.004 0.002"),
    ("fcc", ("1 -1 0", "1 1 -2", "1 1 1"), "0 0 0 0.001 0.002 0.003"),
])
def test_check_mode_setup_matches_oracle(tmp_path, app, prim, s0):
    r = run_exe(deck_text(app=app, px=prim[0], py=prim[1], pz=prim[2], s0=s0), tmp_path, check=True)
    assert r.returncode == 0, r.stderr
    got = parse_check(r.stdout)
    vecs = [six(p) for p in prim]
    geo = {}
    geo["prim"], geo["invprim"], geo["norvol"] = oracle.prim(*vecs)
    four, nf, _ = oracle.set_scale(geo["prim"])
    assert np.allclose(got["four"].reshape(3, 3), four, rtol=1e-15, atol=1e-15)
    xn0, xb0 = (P.basal3() if app == "basal3" else (P.FCC_XN, P.FCC_XB))
    xn, xb = oracle.rotate_slip(xn0, xb0, four, nf)
    assert np.allclose(np.array(got["xn"]), xn, rtol=1e-15, atol=1e-15)
    assert np.allclose(np.array(got["xb"]), xb, rtol=1e-15, atol=1e-15)
    mat = P.material()
    for k in ("C44", "C12", "C11", "S12", "S44"):
        assert abs(got["material"][k] - mat[k]) <= 1e-15 * abs(mat[k])
    assert np.allclose(got["coef"], P.core_coef_sine(mat), rtol=1e-15)
    srot = oracle.rotate_stress(geo["prim"], geo["invprim"], geo["norvol"], sig_from6(six(s0)))
    tau = oracle.resolss(srot, xn, xb, P.DSLIP)
    assert np.allclose(got["tau"], tau, rtol=1e-14, atol=1e-18)


@pytestmark_cpu
@pytest.mark.parametrize("text,msg", [
    ("seed 1\nbogus_command 3\n", "Unknown command"),
    ("app_style nope\n", "Unrecognized app style"),
    ("app_style fcc\nrun 10\n", "Cannot run application until simulation box is defined"),
    ("region box block 0 16 0 16 0 16\ncreate_sites box\n", "Cannot create sites before simulation box"),
    ("sigma initial 1 2\n", "Illegal sigma command"),
])
def test_malformed_decks_fail_like_error_all(tmp_path, text, msg):
    r = run_exe(text, tmp_path, check=True)
    assert r.returncode == 1
    assert "ERROR:" in r.stderr and msg in r.stderr


def oracle_solvegl(case, stopsteps, max_iter, tol, s0, ds):
    """SolveGL::iterate (solve_GL.cpp:70-158) driving the oracle."""
    st = P.oracle_state(case)
    sigma = sig_from6(six(s0))
    dsig = sig_from6(six(ds))
    geo = case["geo"]
    for _ in range(stopsteps):
        srot = oracle.rotate_stress(geo["prim"], geo["invprim"], geo["norvol"], sigma)
        tau = oracle.resolss(srot, case["xn"], case["xb"], P.DSLIP)
        for it in range(max_iter):
            r = st.step(tau)
            if r["xinorm"] < tol:
                break
        sigma = sigma + dsig
    return st.xi


@pytest.mark.gpu
@pytest.mark.parametrize("max_iter,tol,run", [(1, 0, 40), (25, 3e-4, 3)])
def test_deck_run_matches_oracle(tmp_path, max_iter, tol, run):
    print(f"case: driver deck 16^3 x 3, max_iter {max_iter} tol {tol} run {run}", flush=True)
    s0, ds = "0 0 0 0 0.002 0", "0 0 0 0 0.0001 0"
    text = deck_text(max_iter=max_iter, tol=tol, run=run, s0=s0, ds=ds, noise=1e-2,
                     extra="write_xi xi_final.bin\n")
    r = run_exe(text, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    xi = np.fromfile(tmp_path / "xi_final.bin", dtype=np.float64).view(np.complex128).reshape(3, 16, 16, 16)
    case = P.make_case(shape=(16, 16, 16), noise=1e-2, loops=[(0, 2, 8, 8, 8, 4)], sigma_mu=0.0)
    ref = oracle_solvegl(case, run, max_iter, tol, s0, ds)
    err = np.max(np.abs(xi - ref)) / np.max(np.abs(ref))
    assert err < 1e-9, err
    assert "Loop time of" in r.stdout


def read_text_dumps(path):
    snaps, lines = [], open(path).read().splitlines()
    q = 0
    while q < len(lines):
        assert lines[q] == "ITEM: TIMESTEP"
        idump, t = lines[q + 1].split()
        n = int(lines[q + 3])
        cols = lines[q + 8].split()[2:]
        rows = np.array([[float(v) for v in lines[q + 9 + r].split()] for r in range(n)])
        assert lines[q + 9].endswith(" ")  # "%g " per value, as DumpText::write_text
        snaps.append((int(idump), float(t), cols, rows))
        q += 9 + n
    return snaps


def read_bin_dumps(path):
    raw, off, snaps = open(path, "rb").read(), 0, []
    while off < len(raw):
        idump, = np.frombuffer(raw, np.int32, 1, off); off += 4
        t, = np.frombuffer(raw, np.float64, 1, off); off += 8
        ndump, = np.frombuffer(raw, np.int32, 1, off); off += 4
        box = np.frombuffer(raw, np.float64, 6, off); off += 48
        size_one, nprocs = np.frombuffer(raw, np.int32, 2, off); off += 8
        n, = np.frombuffer(raw, np.int32, 1, off); off += 4
        assert n == ndump * size_one and nprocs == 1
        data = np.frombuffer(raw, np.float64, n, off).reshape(ndump, size_one); off += 8 * n
        snaps.append((int(idump), float(t), box, data))
    return snaps


@pytest.mark.gpu
def test_deck_text_and_binary_dumps(tmp_path):
    print("case: driver deck 16^3 x 3 with text and .bin dumps every 5 of 10 steps", flush=True)
    fields = "x y z xir1 xii1 xir2 pxz exz pxx"
    extra = (f"dump 1 text 5 d.txt {fields}\n"
             f"dump 2 text 5 d.bin {fields}\n"
             "write_xi xi_final.bin\n")
    text = deck_text(max_iter=1, tol=0, run=10, s0="0 0 0 0 0.002 0", ds="0 0 0 0 0 0", noise=1e-2, extra=extra)
    r = run_exe(text, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ts = read_text_dumps(tmp_path / "d.txt")
    bs = read_bin_dumps(tmp_path / "d.bin")
    # with max_iter 1 every stress state ends in an output event that writes
    # all dumps (SolveGL::iterate, solve_GL.cpp:108): setup + 10 snapshots
    assert len(ts) == len(bs) == 11
    for q, ((i1, t1, cols, rows), (i2, t2, box, data)) in enumerate(zip(ts, bs)):
        assert i1 == i2 == q and t1 == pytest.approx(t2)
        assert cols == fields.split()
        assert np.allclose(box, [0, 16, 0, 16, 0, 16])
        assert np.allclose(rows, data, rtol=1e-5, atol=1e-300 + 1e-5 * np.abs(data).max(axis=0))
    xi = np.fromfile(tmp_path / "xi_final.bin", dtype=np.float64).view(np.complex128).reshape(3, 16 ** 3)
    last = bs[-1][3]
    assert np.array_equal(last[:, 3], xi[0].real) and np.array_equal(last[:, 4], xi[0].imag)
    assert np.array_equal(last[:, 5], xi[1].real)
    assert np.array_equal(last[:, 0], np.repeat(np.arange(16.0), 256))


@pytestmark_cpu
def test_mpea_file_must_match_the_box(tmp_path):
    """Material::read_mpea_file (material.cpp:848-853) rejects a per-cell
    USFE file whose header does not match the box and slip systems."""
    lines = ["8 16 16 3"] + [f"{i} {j} {k} {s} 158.0" for s in range(3) for i in range(8) for j in range(16)
                             for k in range(16)]
    (tmp_path / "usfe.txt").write_text("\n".join(lines) + "\n")
    text = deck_text().replace("core_energy 3", "core_energy 7").replace("usf 158e-3", "usf 158e-3 mpea usfe.txt")
    r = run_exe(text, tmp_path, check=True)
    assert r.returncode == 1
    assert "MPEA core energy file size does not match box 16 8 16 16 16 16 3 3" in r.stderr


@pytestmark_cpu
@pytest.mark.parametrize("preset,flag,names", [
    ("Cu", 2, [344.948e-3, -85.399e-3, -40.421e-3, 12.396e-3, -0.674e-3, -131.393e-3, -19.816e-3]),
    ("Ti", 4, [427.0e-3, -165.9e-3, -221.9e-3, -26.29e-3, -13.42e-3, -53.09e-3, 119.1e-3, -14.25e-3, -17.77e-3]),
    ("MgZn", 9, [216.6e-3, -82.84e-3, -116.8e-3, -5.873e-3, -11.27e-3, -51.94e-3, 22.04e-3, 9.44e-3, -0.2501e-3]),
])
def test_check_mode_core_models_from_presets(tmp_path, preset, flag, names):
    """Material presets (material.cpp:176-835) feed core_energy_extended (flag 2)
    and core_energy_pyrII (flag 4; 9 = its Fourier model on every cell) with
    coefficients over (mu dslip b) as fft_fftw_slab.cpp:1281-1287, 1468-1476."""
    text = deck_text(app="fcc").replace("core_energy 3", f"core_energy {flag}")
    text = text.replace("material mu 48e9 young 128.64e9 a 0.3615e-9 isf 41e-3 usf 158e-3",
                        f"material preset {preset}")
    r = run_exe(text, tmp_path, check=True)
    assert r.returncode == 0, r.stderr
    got = parse_check(r.stdout)
    mu_young = {"Cu": (54.5e9, 144.4e9, 3.64e-10), "Ti": (44.6e9, 118.67e9, 2.95e-10), "MgZn": (18.0e9, 46.8e9, 3.19e-10)}
    mu, young, a = mu_young[preset]
    mat = oracle.material(mu, young)
    mod = oracle.voigt(mat["C44"], mat["C12"], mat["C11"])
    mdb = mod["mu"] * P.DSLIP * a / np.sqrt(2.0)
    assert np.allclose(got["coef"], np.array(names) / mdb, rtol=1e-14)


@pytestmark_cpu
def test_check_mode_non_schmid_and_usfe_angle_tau(tmp_path):
    """non_Schmid (app.cpp:446-460) -> resolSS_non_Schmid (fft_fftw_slab.cpp:
    774-845) against the oracle, and core_energy 6 (USFE(angle, tau),
    :1500-1545) from the Ta preset's fit."""
    s0 = "0.001 -0.002 0.0005 0.003 -0.004 0.002"
    text = deck_text(app="fcc", s0=s0).replace("core_energy 3", "core_energy 6")
    text = text.replace("material mu 48e9 young 128.64e9 a 0.3615e-9 isf 41e-3 usf 158e-3", "material preset Ta")
    text = text.replace("mobility 0.01 1.0", "mobility 0.01 1.0\nnon_Schmid 25 0.3 -0.2 0.45")
    r = run_exe(text, tmp_path, check=True)
    assert r.returncode == 0, r.stderr
    got = parse_check(r.stdout)
    xb = np.array(got["xb"])
    geo = P.geometry()
    srot = oracle.rotate_stress(geo["prim"], geo["invprim"], geo["norvol"], sig_from6(six(s0)))
    tau = oracle.resolss_non_schmid(srot, xb, P.DSLIP, 25.0, 0.3, -0.2, 0.45)
    assert np.allclose(got["tau"], tau, rtol=1e-14, atol=1e-18)
    mat = oracle.material(70.67e9, 189.24e9)
    mu_s = oracle.voigt(mat["C44"], mat["C12"], mat["C11"])["mu"]
    s21 = srot[2, 1] * mu_s * 1e-9
    g = ((51.59816407 * s21 - 2.27173784) * np.sin(0.9979452 * 25 * np.pi / 180) + (-4.16536 * s21 + 1133.961364)) * 1e-3
    b = np.sqrt(3.0) * 3.3058e-10 / 2.0
    assert np.allclose(got["coef"], [g / (mu_s * P.DSLIP * b)], rtol=1e-13)


@pytest.mark.gpu
def test_deck_non_schmid_usfe_run_matches_oracle(tmp_path):
    print("case: driver deck 16^3 x 3, non_Schmid + core_energy 6, ramp, 12 states", flush=True)
    s0, ds = "0 0 0 0.001 0.002 0.0015", "0 0 0 0.0001 0.0001 0.0002"
    text = deck_text(max_iter=2, tol=0, run=12, s0=s0, ds=ds, noise=1e-2, extra="write_xi xi_final.bin\n")
    text = text.replace("core_energy 3", "core_energy 6")
    text = text.replace("material mu 48e9 young 128.64e9 a 0.3615e-9 isf 41e-3 usf 158e-3", "material preset Fe")
    text = text.replace("mobility 0.01 1.0", "mobility 0.01 1.0\nnon_Schmid 15 0.2 0.1 -0.3")
    r = run_exe(text, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    xi = np.fromfile(tmp_path / "xi_final.bin", dtype=np.float64).view(np.complex128).reshape(3, 16, 16, 16)
    # oracle: SolveGL::iterate with resolSS_non_Schmid and the per-state An'
    mat = oracle.material(87.2796e9, 225.1814e9)
    mod = oracle.voigt(mat["C44"], mat["C12"], mat["C11"])
    mat.update(dslip=P.DSLIP, b=np.sqrt(3.0) * 2.8665e-10 / 2.0, mu_s=mod["mu"], xnu=mod["xnu"])
    case = P.make_case(shape=(16, 16, 16), noise=1e-2, loops=[(0, 2, 8, 8, 8, 4)], sigma_mu=0.0, core="perfect")
    case["mat"] = mat
    BB, _, _ = P.oracle_B(case)
    st = oracle.GLState(case["xi0"], BB, case["cddt"], 1, np.zeros(1))
    sigma, dsig, geo = sig_from6(six(s0)), sig_from6(six(ds)), case["geo"]
    for _ in range(12):
        srot = oracle.rotate_stress(geo["prim"], geo["invprim"], geo["norvol"], sigma)
        tau = oracle.resolss_non_schmid(srot, case["xb"], P.DSLIP, 15.0, 0.2, 0.1, -0.3)
        s21 = srot[2, 1] * mod["mu"] * 1e-9
        g = ((15.524609942496577 * s21 + 0.004388435835439675) * np.sin(0.9934680885709689 * 15 * np.pi / 180)
             + (-0.9065590784654822 * s21 + 944.0633238797747)) * 1e-3
        st.core_par = np.array([g / (mod["mu"] * P.DSLIP * mat["b"])])
        for _ in range(2):
            st.step(tau)
        sigma = sigma + dsig
    err = np.max(np.abs(xi - st.xi)) / np.max(np.abs(st.xi))
    assert err < 1e-9, err


@pytest.mark.gpu
def test_deltao_ddelta_dumps_as_executed(tmp_path):
    """project_core_energy (fft_fftw_slab.cpp:1550-1624) runs after the output
    of each iteration of a non_Schmid run and init_loop zeroes delta / ddelta
    at the start of the next, so dumps inside a run read zeros; the setup
    snapshot of a second run command reads the last projection: for the one
    system of 1L1S100B, deltao = Re xi_1 and ddelta the forward x difference
    on the line j = ny/2, k = nz/2."""
    print("case: driver deck 1L1S100B 16^3, non_Schmid, deltao/ddelta dumps over two runs", flush=True)
    text = deck_text(app="1L1S100B", max_iter=1, tol=0, run=3, s0="0 0 0 0 0.002 0", noise=1e-2,
                     extra="non_Schmid 0.0 0.1 0.2 0.3\ndump 1 text 1 d.txt x y z xir1 deltao ddelta\n")
    text = text.replace("app_style 1L1S100B core_energy 3", "app_style 1L1S100B core_energy 1")
    text = text.replace("run 3\n", "run 3\nrun 2\n")
    r = run_exe(text, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    snaps = read_text_dumps(tmp_path / "d.txt")
    # run 1: setup + 3 states; run 2 (stoptime 3 + 2 = 5, and App::init_app
    # recomputes stopsteps = floor(stoptime / timestep) = 5): setup + 5 states
    assert len(snaps) == 10
    for q, (_, _, cols, rows) in enumerate(snaps):
        assert cols == ["x", "y", "z", "xir1", "deltao", "ddelta"]
        if q != 4:
            assert np.all(rows[:, 4] == 0) and np.all(rows[:, 5] == 0), q
    rows = snaps[4][3]
    assert np.any(rows[:, 3] != 0) and np.allclose(rows[:, 4], rows[:, 3], rtol=1e-5, atol=1e-6)
    x = rows[:, 3].reshape(16, 16, 16)
    want = np.zeros((16, 16, 16))
    want[:-1, 8, 8] = x[1:, 8, 8] - x[:-1, 8, 8]
    assert np.allclose(rows[:, 5].reshape(16, 16, 16), want, rtol=1e-4, atol=1e-5)


APP_STYLES = {  # AppStyle key -> (slip systems, set_slip rotates by four/normfour)
    "fcc": (12, True), "3d-hcp-basal": (3, True), "FrankRead": (2, False), "1L1S100B": (1, False),
    "cond100B": (1, False), "2d-bcc-1lOrtho": (1, True), "1L1S111Bbcc": (1, False), "1L1S110Bfcc": (1, False),
    "1L1S110Bfccnonortho": (1, False), "1L1S110Binclined": (1, False), "1L1S111B": (1, False),
    "1SBCC": (1, False), "3SBCC": (3, False), "BCCEdge": (3, False), "BCCEdgePlanar": (1, False),
    "2L2S110B": (2, False), "2d": (1, True), "3d-2slip": (1, True), "3d2s2lOrtho": (2, True),
    "3d2sStraightNonOrtho": (2, True), "3d2sStraightOrtho": (2, True), "3d3sHCPnotch": (3, True),
}


def _slips(out):
    xn, xb = [], []
    for line in out.splitlines():
        w = line.split()
        if w and w[0] == "slip":
            xn.append([float(v) for v in w[3:6]])
            xb.append([float(v) for v in w[7:10]])
    return np.array(xn), np.array(xb)


@pytestmark_cpu
@pytest.mark.parametrize("app", sorted(APP_STYLES))
def test_every_reference_app_style_parses(tmp_path, app):
    """Every AppStyle key of the reference is accepted: its slip systems are
    unit plane normals perpendicular to their Burgers vectors on the cubic
    grid, and on a skewed grid they are rotated by four/normfour exactly when
    the app's set_slip rotates (checked with the oracle's rotate_slip)."""
    S, rot = APP_STYLES[app]
    base = deck_text(app=app).replace(f"app_style {app} core_energy 3", f"app_style {app} core_energy 1")
    r = run_exe(base, tmp_path, check=True)
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
    xn, xb = _slips(r.stdout)
    assert xn.shape == (S, 3)
    assert np.allclose(np.linalg.norm(xn, axis=1), 1.0, atol=2e-4)
    assert np.allclose(np.sum(xn * xb, axis=1), 0.0, atol=2e-4)
    skew = base.replace("primitive x 1 0 0 y 0 1 0 z 0 0 1", "primitive x 1 0 0 y 0.5 0.8660254 0 z 0 0 1")
    r2 = run_exe(skew, tmp_path, check=True)
    assert r2.returncode == 0
    xn2, xb2 = _slips(r2.stdout)
    if rot:
        geo = P.geometry(((1, 0, 0), (0.5, 0.8660254, 0), (0, 0, 1)))
        wn, wb = oracle.rotate_slip(xn, xb, geo["four"], geo["normfour"])
        assert np.allclose(xn2, wn, rtol=1e-12, atol=1e-15) and np.allclose(xb2, wb, rtol=1e-12, atol=1e-15)
    else:
        assert np.array_equal(xn2, xn) and np.array_equal(xb2, xb)


@pytest.mark.parametrize("style,S,systems", [
    ("mode1", 2, {0, 1}), ("mode2", 2, {0, 1}), ("1LOrthoFCC", 1, {0}), ("1LNonOrthoFCC", 1, {0}),
    ("1LInclinedFCC", 1, {0}), ("2LOrtho", 2, {0, 1}), ("3D3ShcpBasal", 3, {0}), ("3D3ShcpNotch", 3, {0}),
    ("StraightOrtho", 2, {0}), ("StraightNonOrtho", 1, {0}), ("1LOrthoBCC", 1, {0})])
def test_reference_initial_configurations(style, S, systems):
    """The hard-coded initial fields of FFTW_Slab (fft_fftw_slab.cpp:2196-3247):
    zero imaginary part, Re xi in {0, 1}, set only on the systems the
    reference seeds; single-plane seeds lie on one plane (k = nz/2 for the
    {100}/{0001} loops, i + j + k = nx for the inclined {111} loop); the
    notch also freezes the faces i, j, k < 5."""
    import pfdd_setup
    n = 96 if style == "mode2" else 32  # mode2 centres its loops for a 96^3 grid
    with_mask = style == "3D3ShcpNotch"
    out = pfdd_setup.sxtal(style, (n, n, n), S, with_mask=with_mask)
    xi, xo = out if with_mask else (out, None)
    assert np.all(xi.imag == 0) and set(np.unique(xi.real)) <= {0.0, 1.0}
    assert {a for a in range(S) if xi[a].real.any()} == systems
    i, j, k = np.nonzero(xi[0].real)
    if style in ("1LOrthoFCC", "1LNonOrthoFCC", "3D3ShcpBasal", "StraightOrtho", "StraightNonOrtho", "1LOrthoBCC",
                 "mode1", "3D3ShcpNotch"):
        assert np.all(k == n // 2)
    if style in ("1LInclinedFCC", "mode2"):
        assert np.all(i + j + k == n)
    if with_mask:
        assert xo[:, :5].all() and xo[:, :, :5].all() and xo[:, :, :, :5].all()
        assert np.all(xo[0][i, j, k] == 1)
