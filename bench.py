"""GL steps/s at 128^3 x 12 slip systems on one MI355X: FP64 plain path and
FP32 mode, with the HBM roofline of the dominant kernel and a CPU baseline.

Workload (the production configuration of the GL solver): 128x128x128 grid,
FCC table with 4 {111} planes x 3 <110> = 12 slip systems, Cu-like isotropic
material (mu 48 GPa, nu 0.34, b = 0.3615/sqrt2 nm), sine GSF core energy
(isf/usf 41/158 mJ/m^2), 8 seeded circular loops (radii 8-20 cells) plus
uniform noise 1e-3, constant 400 MPa shear resolved on system 0, explicit
Euler with CD*dt = 0.01.  One "step" is one GL iteration (SolveGL::iterate
inner loop body) over all 12 order parameters; inputs are resident in HBM
before the timed region.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "phase-field-dislocation-dynamics-pfdd-2.x_amd")
sys.path.insert(0, os.path.join(PKG, "python"))

import pfdd_gl  # noqa: E402
import pfdd_setup as setup  # noqa: E402

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s


def make_workload(n=128, seed=576):
    prob = setup.fcc_loops_problem((n, n, n), nloops=8, rmin=8, rmax=20, noise=1e-3,
                                   shear_pa=400e6, seed=seed)
    return prob


STORED_B = False  # --stored-b: stream the packed B(k) instead of evaluating it in the x pass


def solver_for(prob, precision):
    s = pfdd_gl.GLSolver(prob["shape"], prob["xn"], prob["xb"], prob["four"], prob["C44"],
                         prob["C12"], prob["C11"], prob["dslip"], prob["cddt"], S12=prob["S12"],
                         S44=prob["S44"], core_model=pfdd_gl.CORE_SINE, coef=prob["core_coef"],
                         precision=precision, stored_B=STORED_B)
    s.set_xi(prob["xi0"])
    s.set_tau(prob["tau"])
    # tuning last (it restores the field): the GPU goes from the tuner's
    # iterations straight into the warm-up without an idle gap
    s.autotune()
    return s


def time_steps(s, steps, warmup):
    import ctypes
    if warmup:
        s.step(warmup)
    # Synchronise (step() with sync waits for the GPU), then time K iterations
    # enqueued back to back and one final wait.
    s.step(0)
    t0 = time.perf_counter()
    r = s.step(steps)
    t1 = time.perf_counter()
    return (t1 - t0), r


def roofline(s, prof_steps=50, warm_steps=50):
    # steady state first (the clocks settle within a few tens of iterations of
    # continuous load), then per-kernel device times from HIP events that the
    # kernel dispatches record on the solver's stream
    s.step(warm_steps)
    prof = s.profile(prof_steps)
    kbytes, step_bytes = s.traffic()
    slots = ["fwd_z", "fwd_y", "x_contract", "inv_y", "inv_z_update"]
    plane = kbytes[1] == 0  # the three-kernel plane iteration has no separate y passes
    names = ["fwd_yz", None, "x_contract", None, "inv_yz_update"] if plane else slots
    live = [q for q in range(5) if names[q]]
    times = {q: prof[slots[q]] for q in live}
    dom = max(live, key=lambda q: times[q])
    achieved = kbytes[dom] / (times[dom] * 1e-3)
    per_kernel = {names[q]: {"ms": times[q], "bytes": kbytes[q],
                             "GBps": kbytes[q] / (times[q] * 1e-3) / 1e9} for q in live}
    return {
        "kernel": names[dom],
        "bound": "hbm",
        "achieved": achieved / 1e9,
        "peak": PEAK_HBM / 1e9,
        "unit": "GB/s",
        "frac": achieved / PEAK_HBM,
        "traffic": None,
        "per_kernel": per_kernel,
        "step_bytes": step_bytes,
        "step_ms_events": prof["step"],
        "field_bytes": kbytes[0] / 2,
        "iteration": "three-kernel plane" if plane else "five-kernel",
    }


def load_pmc_traffic(tag, kernel):
    """HBM bytes per launch of the kernel instance `kernel` (its template name,
    e.g. "k_update<float, 128, 2, false>") from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated with a copy kernel), or None
    when that instance was not measured."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{tag}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel)
        return k["hbm_bytes_per_launch"] if k else None
    except Exception:
        return None


def kernel_instance(s, role, tiles):
    """Template name of the kernel that runs `role` for solver s with the
    tuning choice `tiles` (fwd z, fwd y, inv y, inv z + update, iteration),
    as rocprofv3 reports it (sine-series core: the update's GEN flag is false)."""
    T = "double" if s.precision == pfdd_gl.PFDD_FP64 else "float"
    nx, ny, nz = s.shape
    S = s.S
    if role == "fwd_yz":
        return f"k_plane_fwd<{T}, {ny}, {nz}>"
    if role == "inv_yz_update":
        return f"k_plane_upd<{T}, {ny}, {nz}, false>"
    if role == "x_contract":
        if STORED_B:
            return f"k_mid<{T}, {T}, {nx}, {S}>"
        return {0: "k_mid_otf", 1: "k_mid_otf_pipe", 2: "k_mid_otf_pl"}[tiles[4]] + f"<{T}, {nx}, {S}>"
    if role == "inv_z_update":
        return f"k_update<{T}, {nz}, {tiles[3]}, false>"
    n, inv, tq = {"fwd_z": (nz, "false", tiles[0]), "fwd_y": (ny, "false", tiles[1]),
                  "inv_y": (ny, "true", tiles[2])}[role]
    return f"k_fft_pass<{T}, {n}, {inv}, {tq}>"


# kernel families of the GL iteration and the roles they play
KERNEL_ROLES = {"k_mid_otf": "x_contract", "k_mid<": "x_contract", "k_update<": "inv_z_update",
                "k_plane_fwd": "fwd_yz", "k_plane_upd": "inv_yz_update", "k_fft_pass": "line_pass"}


PROBE_STEPS = 3


def pmc_probe(size):
    """Workload for the rocprofv3 --pmc runs: a plain element-wise copy of known
    size (torch's element-wise multiply by 1: ordinary loads and stores, the
    calibration kernel), then 3 GL iterations in each precision."""
    import torch
    n = 1 << 28  # float32 elements: 1 GiB
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(f"pmc-probe: copied {4 * n} bytes x3", flush=True)
    del x, y
    torch.cuda.empty_cache()
    prob = make_workload(size)
    for prec in (pfdd_gl.PFDD_FP64, pfdd_gl.PFDD_FP32):
        s = solver_for(prob, prec)
        s.step(PROBE_STEPS)
        s.close()
        print(f"pmc-probe: precision {prec} done", flush=True)
    return 0


def pmc_summarize(fetch_dir, write_dir, size, out_dir):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) from two rocprofv3 --pmc runs
    of --pmc-probe, calibrated by the copy kernel, for the unique-named GL
    kernels; writes profiles/pmc_traffic_{fp64,fp32}.json."""
    import csv
    import glob

    def read(d, counter):
        rows = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    if r.get("Counter_Name") != counter:
                        continue
                    key = (r["Dispatch_Id"], r["Kernel_Name"])
                    rows[key] = rows.get(key, 0.0) + float(r["Counter_Value"])
        return rows

    fetch, write = read(fetch_dir, "FETCH_SIZE"), read(write_dir, "WRITE_SIZE")
    nbytes = float(1 << 30)

    def per_kernel(rows, match, last=None):
        # the probe's own GL iterations are the last dispatches of each kernel
        # role (the autotuner's candidate launches come before them)
        v = sorted(((int(d), name, val) for (d, name), val in rows.items() if match(name)))
        if last:
            v = v[-last:]
        return (sum(x[2] for x in v) / len(v), v[-1][1]) if v else (None, None)

    is_copy = lambda n: "elementwise" in n and "Mul" in n
    copy_f, _ = per_kernel(fetch, is_copy)
    copy_w, _ = per_kernel(write, is_copy)
    copy_name = next((name for (d, name) in fetch if is_copy(name)), None)
    sf = nbytes / (copy_f * 1024.0) if copy_f else 1.0
    sw = nbytes / (copy_w * 1024.0) if copy_w else 1.0
    prob = make_workload(size)
    short = lambda n: (re.search(r"(k_\w+<[^>]*>)", n) or re.search(r"(.*)", n)).group(1)
    out = {}
    for prec, tag, tname in ((pfdd_gl.PFDD_FP64, "fp64", "double"), (pfdd_gl.PFDD_FP32, "fp32", "float")):
        s = solver_for(prob, prec)
        kbytes, _ = s.traffic()
        s.close()
        # algorithmic bytes per launch: 2 S N e for a line pass or the x pass,
        # 3 S N e for an update (the plane kernels: 2 and 3 S N e as well)
        two, three = kbytes[0], kbytes[4]
        res = {"copy_calibration": {"kernel": copy_name, "bytes_read": nbytes, "bytes_written": nbytes,
                                    "fetch_kib": copy_f, "write_kib": copy_w,
                                    "fetch_scale": sf, "write_scale": sw}, "kernels": {}}
        # every instance the probe ran (the tuner's candidates included),
        # averaged over its last PROBE_STEPS dispatches
        names = sorted({short(n) for (_, n) in fetch if f"<{tname}," in n and "k_" in n})
        for kn in names:
            role = next((r for sub, r in KERNEL_ROLES.items() if kn.startswith(sub)), None)
            if role is None:
                continue
            m = lambda n, kn=kn: short(n) == kn
            (f, _), (w, _) = per_kernel(fetch, m, PROBE_STEPS), per_kernel(write, m, PROBE_STEPS)
            if f is None or w is None:
                continue
            algo = three if role in ("inv_z_update", "inv_yz_update") else two
            res["kernels"][kn] = {"role": role, "launches": PROBE_STEPS,
                                  "fetch_bytes": f * 1024.0 * sf, "write_bytes": w * 1024.0 * sw,
                                  "hbm_bytes_per_launch": f * 1024.0 * sf + w * 1024.0 * sw,
                                  "algorithmic_bytes": algo}
        out[tag] = res
        os.makedirs(out_dir, exist_ok=True)
        path = os.path.join(out_dir, f"pmc_traffic_{tag}.json")
        with open(path, "w") as fh:
            json.dump(res, fh, indent=1)
        print(f"wrote {path}")
    return 0


def cpu_baseline(prob, nsteps=3):
    """The FP64 CPU oracle (a C restatement of the reference's GL step) timed on
    the host's cores on a bounded sample: B(k) setup untimed, `nsteps` GL
    iterations of the same 128^3 x 12 workload timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    oracle.set_threads(cores)
    BB, _, _ = oracle.bmatrix(prob["shape"], prob["C44"], prob["C12"], prob["C11"], prob["dslip"],
                              prob["xn"], prob["xb"], prob["four"])
    st = oracle.GLState(prob["xi0"], BB, prob["cddt"], 3, prob["core_coef"], 0, None, None)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        st.step(prob["tau"])
    dt = time.perf_counter() - t0
    del BB, st
    return {"value": nsteps / dt, "unit": "GL steps/s", "cores": cores, "kind": "port",
            "sample": f"{nsteps} GL iterations of the 128^3 x 12 workload (B(k) setup untimed), "
                      f"FP64 C oracle with {cores} OpenMP threads"}


def bench_ensemble(args):
    """Batched alloy ensemble: 16 x 48^3 + 24 x 64^3 + 8 x 96^3 realizations, 3 systems
    each, FP32, per-cell lognormal GSF amplitude, per-realization shear in
    [200, 600] MPa, straight edge dislocation; every GL iteration of all 48
    realizations is one hipGraph (3 batch handles) replayed back to back."""
    groups = setup.alloy_ensemble()
    prec = pfdd_gl.PFDD_FP64 if args.precision == "fp64" else pfdd_gl.PFDD_FP32
    solvers = []
    for g in groups:
        s = pfdd_gl.GLSolver(g["shape"], g["xn"], g["xb"], g["four"], g["C44"], g["C12"], g["C11"], g["dslip"],
                             g["cddt"], S12=g["S12"], S44=g["S44"], core_model=pfdd_gl.CORE_SERIES,
                             coef=g["coef"], ncoef=g["ncoef"], amp=g["amp"], precision=prec, batch=g["batch"],
                             autotune=True)
        s.set_xi(g["xi0"])
        s.set_tau(g["tau"])
        solvers.append(s)
    graph = pfdd_gl.GLGraph(solvers, iters=10)
    graph.run(args.warmup - args.warmup % 10 or 10)   # capture + warm-up (even count: same buffers)
    for s, g in zip(solvers, groups):
        s.set_xi(g["xi0"])
    nsteps = args.steps
    half = nsteps // 2
    # two timed halves; the per-system sums of Re xi between them give each
    # realization's glide over the second half (after the initial core relaxation)
    t0 = time.perf_counter()
    graph.run(nsteps - half)
    dt = time.perf_counter() - t0
    mid = [s.xi_sums() for s in solvers]
    t0 = time.perf_counter()
    res = graph.run(half)
    dt += time.perf_counter() - t0
    end = [s.xi_sums() for s in solvers]
    nreal = sum(g["batch"] for g in groups)
    step_bytes = sum(s.traffic()[1] for s in solvers)
    members = []
    for g, a, b in zip(groups, mid, end):
        n = g["shape"][0]
        for q in range(g["batch"]):
            # swept area of system 0 / plane area; two dislocation lines per box
            glide = (b[q * g["S"]] - a[q * g["S"]]) / (n * n) * n / 2.0
            members.append({"grid": n, "stress_MPa": float(g["stress_pa"][q] / 1e6),
                            "velocity_cells_per_step": glide / half})
    for s in solvers:
        s.close()
    graph.close()
    # disorder-free twin of every realization (amplitude 1, same grid, stress and
    # protocol; untimed): glide velocity depends on the box size through the
    # restoring force of the periodic dipole, so each member is compared with
    # its own twin only
    clean = setup.alloy_ensemble(cv=0.0)
    twins = []
    for g in clean:
        s = pfdd_gl.GLSolver(g["shape"], g["xn"], g["xb"], g["four"], g["C44"], g["C12"], g["C11"], g["dslip"],
                             g["cddt"], S12=g["S12"], S44=g["S44"], core_model=pfdd_gl.CORE_SERIES,
                             coef=g["coef"], ncoef=g["ncoef"], amp=g["amp"], precision=prec, batch=g["batch"])
        s.set_xi(g["xi0"])
        s.set_tau(g["tau"])
        s.step(nsteps - half)
        a = s.xi_sums()
        s.step(half)
        b = s.xi_sums()
        n = g["shape"][0]
        twins += [(b[q * g["S"]] - a[q * g["S"]]) / (2.0 * n) / half for q in range(g["batch"])]
        s.close()
    for mbr, v0 in zip(members, twins):
        mbr["velocity_disorder_free_cells_per_step"] = v0
        mbr["mobility_ratio"] = mbr["velocity_cells_per_step"] / v0 if v0 > 0 else None
    line = {
        "metric": "realization GL steps/s, batched alloy ensemble in one hipGraph",
        "value": nreal * nsteps / dt, "unit": "realization-steps/s", "n_gpus": 1, "steps": nsteps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / nsteps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic (lognormal GSF amplitude fields, uniform stresses)",
        "config": {"workload": "16 x 48^3 + 24 x 64^3 + 8 x 96^3 realizations x 3 basal systems, sine core "
                               "x per-cell lognormal amplitude (mean 1, CV 0.15, correlation length 4), "
                               "shear U[200,600] MPa, straight edge dislocation, 10 GL iterations per graph launch",
                   "realizations": nreal},
        "hbm": {"bytes_per_step": step_bytes, "achieved_GBps": step_bytes * nsteps / dt / 1e9,
                "frac": step_bytes * nsteps / dt / PEAK_HBM},
        "steps_done": [r["steps_done"] for r in res],
        "critical_stress": "per realization from `bench.py --workload depinning` (stress bracketing)",
        "realizations": members,
    }
    print(json.dumps(line))
    return 0


def bench_depinning(args, levels=8, rounds=3, smax_pa=700e6):
    """Critical (depinning) stress of every realization of the alloy ensemble
    by stress bracketing: each realization runs at `levels` shear stresses at
    once (batch members sharing its amplitude field) for --steps iterations.
    A level depins if the dislocation pair has swept a quarter of the box
    (each line glided n/8 cells or more: the slipped area of system 0 on the
    slip plane grew by n^2/4) or is still gliding in the last quarter of the
    run (more than 0.05 cells per line); otherwise the pair is pinned.  The
    critical stress lies between the highest level below which every level
    stays pinned and the lowest level that depins.  The first round spaces the
    levels over [0, smax], each later round over the bracket of the previous
    one."""
    base = setup.alloy_ensemble()
    prec = pfdd_gl.PFDD_FP64 if args.precision == "fp64" else pfdd_gl.PFDD_FP32
    m = setup.cu_like()
    sl = setup.slip("basal3")
    b0 = sl["xb"][0] / np.linalg.norm(sl["xb"][0])
    n0 = sl["xn"][0] / np.linalg.norm(sl["xn"][0])
    unit = np.outer(b0, n0) + np.outer(n0, b0)
    brackets = [[(0.0, smax_pa)] * g["batch"] for g in base]
    tsteps, tsec, nmem = 0, 0.0, 0
    for rnd in range(rounds):
        for gi, g in enumerate(base):
            n, cnt, S = g["shape"][0], g["batch"], g["S"]
            N = n ** 3
            lv = []
            for q in range(cnt):
                lo, hi = brackets[gi][q]
                lv.append(np.linspace(lo, hi, levels) if rnd == 0 else np.linspace(lo, hi, levels + 2)[1:-1])
            amp_q = g["amp"].reshape(cnt, S * N)
            amp = np.ascontiguousarray(np.repeat(amp_q, levels, axis=0)).reshape(-1)
            xi0 = np.concatenate([setup.edge_dislocation_field(g["shape"], S)] * (cnt * levels))
            tau = np.concatenate([setup.resolved_shear("basal3", (sp / m.mu) * unit, m.dslip)[0]
                                  for q in range(cnt) for sp in lv[q]])
            s = pfdd_gl.GLSolver(g["shape"], g["xn"], g["xb"], g["four"], g["C44"], g["C12"], g["C11"], g["dslip"],
                                 g["cddt"], S12=g["S12"], S44=g["S44"], core_model=pfdd_gl.CORE_SERIES,
                                 coef=g["coef"], ncoef=g["ncoef"], amp=amp, precision=prec,
                                 batch=cnt * levels)
            s.set_xi(xi0)
            s.set_tau(tau)
            a = s.xi_sums()
            last = args.steps // 4
            t0 = time.perf_counter()
            s.step(args.steps - last)
            tsec += time.perf_counter() - t0
            c = s.xi_sums()
            t0 = time.perf_counter()
            s.step(last)
            tsec += time.perf_counter() - t0
            b = s.xi_sums()
            s.close()
            tsteps += args.steps * cnt * levels
            nmem += cnt * levels
            for q in range(cnt):
                glide = [(b[(q * levels + k) * S] - a[(q * levels + k) * S]) / (2.0 * n) for k in range(levels)]
                tail = [(b[(q * levels + k) * S] - c[(q * levels + k) * S]) / (2.0 * n) for k in range(levels)]
                moved = [gl >= n / 8.0 or tl > 0.05 for gl, tl in zip(glide, tail)]
                first = next((k for k in range(levels) if moved[k]), None)
                lo_old, hi_old = brackets[gi][q]
                if first is None:
                    brackets[gi][q] = (float(lv[q][-1]), hi_old)
                elif first == 0:
                    brackets[gi][q] = (lo_old, float(lv[q][0]))
                else:
                    brackets[gi][q] = (float(lv[q][first - 1]), float(lv[q][first]))
    out = []
    for gi, g in enumerate(base):
        for q in range(g["batch"]):
            lo, hi = brackets[gi][q]
            out.append({"grid": g["shape"][0], "critical_stress_MPa": 0.5 * (lo + hi) / 1e6,
                        "bracket_MPa": [lo / 1e6, hi / 1e6]})
    crit = np.array([r["critical_stress_MPa"] for r in out])
    line = {
        "metric": "critical (depinning) stress per alloy realization, stress bracketing in batched runs",
        "value": tsteps / tsec, "unit": "realization-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic (lognormal GSF amplitude fields)",
        "config": {"workload": "the 48 alloy realizations of the ensemble workload (16 x 48^3 + 24 x 64^3 + "
                               "8 x 96^3, 3 basal systems, per-cell lognormal amplitude, straight edge pair), "
                               f"{levels} shear levels per realization and round, {rounds} rounds, "
                               f"{args.steps} iterations per level",
                   "depinned_if": "each line glides >= n/8 cells within the run, or > 0.05 cells in its last quarter"},
        "members_per_round": nmem // rounds,
        "critical_stress_MPa": {"mean": float(crit.mean()), "std": float(crit.std()),
                                "min": float(crit.min()), "max": float(crit.max())},
        "realizations": out,
    }
    print(json.dumps(line))
    return 0


def bench_adjoint(args):
    """Reverse mode on the 64^3 x 3 loop problem: 500 GL iterations forward with
    checkpoints every 50, then the adjoint sweep; cost ratio and a central
    finite-difference check of dL/dcoef and dL/dsigma (h = 1e-6; its O(h^2)
    truncation also removed by Richardson extrapolation with h/2)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    n, nsteps, every = 64, args.steps, 50
    m = setup.cu_like()
    sl = setup.slip("basal3")
    S = 3
    rng = np.random.default_rng(576)
    base = setup.core_coef(m, 3)
    c_true = np.zeros((S, 4))
    c_true[:, 0], c_true[:, 1] = base[0], base[1]
    c_true[:, 2] = 0.1 * base[0]
    c_true[:, 3] = -0.05 * base[0]
    c_true *= 1.0 + 0.2 * (2.0 * rng.random((S, 4)) - 1.0)
    c_guess = c_true.copy()
    c_guess[:, 0], c_guess[:, 1] = base[0], base[1]
    c_guess[:, 2], c_guess[:, 3] = 0.1 * base[0], -0.05 * base[0]
    b0 = sl["xb"][0] / np.linalg.norm(sl["xb"][0])
    n0 = sl["xn"][0] / np.linalg.norm(sl["xn"][0])
    unit = np.outer(b0, n0) + np.outer(n0, b0)
    sig_mu = 250e6 / m.mu

    def tau_of(sig):
        return setup.resolved_shear("basal3", sig * unit, m.dslip)[0]

    x0 = np.zeros((S, n, n, n), dtype=np.complex128)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x0[0, :, :, n // 2][(ii - n / 2) ** 2 + (jj - n / 2) ** 2 <= 144] = 1.0
    x0 += 2e-3 * (setup.ranmar(576, x0.size).reshape(x0.shape) - 0.5)
    out = {}
    for prec, tag in ((pfdd_gl.PFDD_FP64, "fp64"), (pfdd_gl.PFDD_FP32, "fp32")):
        s = pfdd_gl.GLSolver((n, n, n), sl["xn"], sl["xb"], sl["four"], m.C44, m.C12, m.C11, m.dslip, 0.01,
                             core_model=pfdd_gl.CORE_SERIES, coef=c_true.reshape(-1), ncoef=4, precision=prec)

        def forward(c, sig, record=False):
            s.set_core_coef(c.reshape(-1), 4)
            s.set_tau(tau_of(sig))
            s.set_xi(x0)
            if record:
                s.tape_record(nsteps, every)
            else:
                s.step(nsteps)
            return s.get_xi()

        target = forward(c_true, sig_mu)
        loss = lambda x: float(np.sum(np.abs(x - target) ** 2))
        forward(c_guess, sig_mu)                       # warm-up
        s.set_xi(x0)
        s.step(0)
        t0 = time.perf_counter(); s.step(nsteps); t_fwd = time.perf_counter() - t0
        s.set_core_coef(c_guess.reshape(-1), 4)
        s.set_tau(tau_of(sig_mu))
        s.set_xi(x0)
        t0 = time.perf_counter(); s.tape_record(nsteps, every); t_rec = time.perf_counter() - t0
        xT = s.get_xi()
        t0 = time.perf_counter()
        dxi0, dtau, dcoef = s.tape_backward(2.0 * (xT - target), 4)
        t_bwd = time.perf_counter() - t0
        dsig = float(np.dot(dtau, tau_of(1.0)))
        res = {"forward_ms_per_step": 1e3 * t_fwd / nsteps, "record_ms_per_step": 1e3 * t_rec / nsteps,
               "backward_ms_per_step": 1e3 * t_bwd / nsteps, "adjoint_cost_vs_forward": t_bwd / t_fwd,
               "loss": loss(xT)}
        if prec == pfdd_gl.PFDD_FP64:
            h, worst, fds, adj, rich = 1e-6, 0.0, [], [], []

            def cfd(dc, ds, hh):
                return (loss(forward(c_guess + hh * dc, sig_mu + hh * ds)) -
                        loss(forward(c_guess - hh * dc, sig_mu - hh * ds))) / (2 * hh)

            for a in range(S + 1):
                for q in range(4 if a < S else 1):
                    dc, ds = np.zeros_like(c_guess), 0.0
                    if a < S:
                        dc[a, q] = 1.0
                    else:
                        ds = 1.0
                    fd, fd2 = cfd(dc, ds, h), cfd(dc, ds, h / 2)
                    g = dcoef[a, q] if a < S else dsig
                    worst = max(worst, abs(fd - g) / max(abs(fd), 1e-30))
                    fds.append(fd)
                    adj.append(g)
                    # Richardson: the O(h^2) truncation of the central difference cancels
                    rich.append((4.0 * fd2 - fd) / 3.0)

            def normwise(x, y):
                x, y = np.asarray(x), np.asarray(y)
                return float(np.linalg.norm(x - y) / np.linalg.norm(x))
            fa, aa = np.asarray(fds), np.asarray(adj)
            res["fd_check"] = {"h": h, "params": "12 GSF coefficients + applied stress",
                               "max_rel_diff_adjoint_vs_central_fd": worst,
                               "normwise_rel_diff": normwise(fds, adj),
                               # per component, scaled by the largest gradient component
                               "max_abs_diff_over_max_grad": float(np.max(np.abs(fa - aa)) / np.max(np.abs(fa))),
                               # rounding floor of a central difference of the loss: eps L / h
                               "fd_rounding_floor_abs": float(np.finfo(float).eps * res["loss"] / h),
                               "min_abs_grad": float(np.min(np.abs(fa))),
                               "max_rel_diff_adjoint_vs_richardson_fd": float(
                                   np.max(np.abs(np.asarray(rich) - aa) / np.abs(np.asarray(rich))))}
            out["grad64"] = (dcoef.copy(), dsig)
        else:
            g64, s64 = out["grad64"]
            res["rel_err_grad_vs_fp64"] = float(max(np.max(np.abs(dcoef - g64)) / np.max(np.abs(g64)),
                                                    abs(dsig - s64) / abs(s64)))
        out[tag] = res
        s.close()
    out.pop("grad64")
    line = {"metric": "adjoint cost relative to a forward GL step, 64^3 x 3, 500 steps, checkpoints every 50",
            "value": out["fp64"]["adjoint_cost_vs_forward"], "unit": "x forward", "n_gpus": 1,
            "steps": nsteps, "warmup": 0, "higher_is_better": False, "dtype": "f64",
            "data": "synthetic (loop r=12 + noise, hidden GSF coefficients +-20%)",
            "config": {"workload": "64^3 x 3 basal systems, 4-term sine-series core, 250 MPa, L2 loss vs target"},
            "fp64": out["fp64"], "fp32": out["fp32"]}
    print(json.dumps(line))
    return 0


def bench_frank_read(args):
    """512 x 512 x 256 skewed grid, 3 systems, pinned Frank-Read source, 500 MPa:
    FP32 and FP64 runs of the same length, steps/s, HBM rate and FP32 error."""
    prob = setup.frank_read_skewed()
    out = {}
    xs = {}
    for prec, tag in ((pfdd_gl.PFDD_FP32, "fp32"), (pfdd_gl.PFDD_FP64, "fp64")):
        s = pfdd_gl.GLSolver(prob["shape"], prob["xn"], prob["xb"], prob["four"], prob["C44"], prob["C12"],
                             prob["C11"], prob["dslip"], prob["cddt"], S12=prob["S12"], S44=prob["S44"],
                             core_model=pfdd_gl.CORE_SINE, coef=prob["core_coef"], xo=prob["xo"],
                             precision=prec)
        s.set_xi(prob["xi0"])
        s.set_tau(prob["tau"])
        # tile shapes and iteration variant for this grid (restores the field)
        tiles = s.autotune()
        s.step(0)
        t0 = time.perf_counter()
        done = 0
        while done < args.steps:
            k = min(500, args.steps - done)
            r = s.step(k)
            done += k
            print(f"frank-read {tag}: {done} steps", file=sys.stderr, flush=True)
        dt = time.perf_counter() - t0
        kbytes, step_bytes = s.traffic()
        xs[tag] = s.get_xi()
        # per-kernel device times (HIP events recorded by the dispatches)
        prof = s.profile(20)
        slots = ["fwd_z", "fwd_y", "x_contract", "inv_y", "inv_z_update"]
        per_kernel = {k: {"ms": prof[k], "GBps": kbytes[q] / (prof[k] * 1e-3) / 1e9}
                      for q, k in enumerate(slots) if kbytes[q] > 0}
        out[tag] = {"steps_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                    "hbm_GBps": step_bytes * args.steps / dt / 1e9, "bytes_per_step": step_bytes,
                    "e_total": r["e_total"], "xiave": r["xiave"], "tiles": tiles, "per_kernel": per_kernel}
        s.close()
    err = float(np.max(np.abs(xs["fp32"] - xs["fp64"])) / np.max(np.abs(xs["fp64"])))
    e_err = abs(out["fp32"]["e_total"] - out["fp64"]["e_total"]) / abs(out["fp64"]["e_total"])
    line = {"metric": "GL steps/s, 512x512x256 skewed grid x 3 systems, Frank-Read source",
            "value": out["fp32"]["steps_per_s"], "unit": "GL steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 0, "ms_per_step": out["fp32"]["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (pinned straight segment of 64 cells)",
            "config": {"workload": "512x512x256, primitive y = [1/2 sqrt3/2 0], basal3 systems, sine core, "
                                   "500 MPa shear on system 0, obstacle mask pins both segment ends"},
            "fp32": out["fp32"], "fp64": out["fp64"], "rel_err_xi_fp32_vs_fp64": err,
            "rel_err_energy_fp32_vs_fp64": e_err}
    print(json.dumps(line))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="workload for rocprofv3 --pmc runs (copy calibration + 3 steps)")
    ap.add_argument("--pmc-summarize", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="turn two rocprofv3 --pmc outputs into pmc_traffic_*.json")
    ap.add_argument("--pmc-out", default=os.path.join(ROOT, "profiles"))
    ap.add_argument("--workload", default="c1", choices=["c1", "ensemble", "depinning", "adjoint", "frank-read"],
                    help="c1: the 128^3 x 12 headline (default); the others print their own JSON line")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"], help="ensemble precision")
    ap.add_argument("--stored-b", action="store_true",
                    help="headline workload with the packed B(k) streamed by the x pass (k_mid)")
    args = ap.parse_args()
    global STORED_B
    STORED_B = args.stored_b
    if args.pmc_probe:
        return pmc_probe(args.size)
    if args.pmc_summarize:
        return pmc_summarize(args.pmc_summarize[0], args.pmc_summarize[1], args.size, args.pmc_out)
    if args.gpus != 1:
        print(f"bench.py: this build runs the GL solver on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU slab decomposition is a design note, not built.", file=sys.stderr)
        return 2

    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "depinning":
        return bench_depinning(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "frank-read":
        return bench_frank_read(args)

    prob = make_workload(args.size)
    out = {}
    for prec, tag in ((pfdd_gl.PFDD_FP64, "fp64"), (pfdd_gl.PFDD_FP32, "fp32")):
        s = solver_for(prob, prec)
        dt, r = time_steps(s, args.steps, args.warmup)
        t0 = time.perf_counter()
        xi_end = s.get_xi()  # device -> host copy of the field (FP64 host layout)
        t_down = time.perf_counter() - t0
        t0 = time.perf_counter()
        s.set_xi(xi_end)     # host -> device
        t_up = time.perf_counter() - t0
        roof = roofline(s)
        roof["tiles"] = s.autotune()  # cached choice of the create-time tuning
        roof["instance"] = kernel_instance(s, roof["kernel"], roof["tiles"])
        # the committed PMC summary is of the default (on-the-fly) contraction
        roof["traffic"] = None if STORED_B else load_pmc_traffic(tag, roof["instance"])
        out[tag] = {"sec": dt, "steps_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                    "host_copy_ms": {"upload": 1e3 * t_up, "download": 1e3 * t_down},
                    "roofline": roof, "e_total": r["e_total"], "xi_end": xi_end}
        s.close()
    # FP32 error against the FP64 GPU run of the same length
    x64, x32 = out["fp64"].pop("xi_end"), out["fp32"].pop("xi_end")
    err_xi = float(np.max(np.abs(x32 - x64)) / np.max(np.abs(x64)))
    err_e = abs(out["fp32"]["e_total"] - out["fp64"]["e_total"]) / abs(out["fp64"]["e_total"])

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(prob, args.cpu_steps)
    # the practical streaming rates the passes are compared with (the
    # roofline peak stays the nominal 8 TB/s): the library's own nontemporal
    # copy and add kernels over arrays of one field each (per precision: the
    # FP32 field is half the size, and part of it stays in the last-level
    # cache between passes).  A failed probe costs only this context field.
    ceil = {}
    for tag, o in out.items():
        nb = int(o["roofline"]["field_bytes"])
        try:
            cp, ad = pfdd_gl.stream_rates(nb)
        except Exception as exc:  # noqa: BLE001
            print(f"bench.py: streaming-rate probe ({tag}) failed: {exc}", file=sys.stderr)
            ceil[tag] = None
            continue
        ceil[tag] = {"copy (1 read + 1 write)": {"GBps": cp, "bytes": 2 * nb},
                     "add (2 reads + 1 write)": {"GBps": ad, "bytes": 3 * nb},
                     "kernels": "pfdd_gl_stream_rates: nontemporal 16-byte loads and stores, one element per thread"}
        rate = {"line": cp, "update": ad}
        for role, k in o["roofline"]["per_kernel"].items():
            k["frac_of_streaming_ceiling"] = k["GBps"] / rate["update" if "update" in role else "line"]

    f64, f32 = out["fp64"], out["fp32"]
    line = {
        "metric": "GL steps/s at 128^3x12 slip systems, FP64 plain vs FP32, + % HBM roofline",
        "value": f64["steps_per_s"],
        "unit": "GL steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded RANMAR loops + noise)",
        "config": {"workload": f"GL step, {args.size}^3 grid x 12 fcc slip systems, sine core, "
                               "400 MPa shear, CD*dt 0.01", "grid": [args.size] * 3,
                   "slip_systems": 12, "precision": "FP64 plain (value); FP32 mode below",
                   "contraction": "packed B(k) streamed" if STORED_B else "B(k) evaluated in the x pass"},
        "roofline": {k: f64["roofline"][k] for k in
                     ("kernel", "instance", "bound", "achieved", "peak", "unit", "frac", "traffic")},
        "cpu_baseline": cpu,
        "fp32": {"value": f32["steps_per_s"], "ms_per_step": f32["ms_per_step"],
                 "speedup_vs_fp64": f32["steps_per_s"] / f64["steps_per_s"],
                 "rel_err_xi_vs_fp64": err_xi, "rel_err_energy_vs_fp64": err_e,
                 "roofline": {k: f32["roofline"][k] for k in
                              ("kernel", "instance", "achieved", "frac", "traffic")}},
        "step_roofline": {
            "fp64": {"bytes": f64["roofline"]["step_bytes"],
                     "frac": f64["roofline"]["step_bytes"] / (f64["ms_per_step"] * 1e-3) / PEAK_HBM},
            "fp32": {"bytes": f32["roofline"]["step_bytes"],
                     "frac": f32["roofline"]["step_bytes"] / (f32["ms_per_step"] * 1e-3) / PEAK_HBM}},
        "per_kernel": {"fp64": f64["roofline"]["per_kernel"], "fp32": f32["roofline"]["per_kernel"]},
        "streaming_ceiling": ceil,
        # the same K steps with the field handed over in host memory (one
        # upload before, one download after: the C ABI's set_xi / get_xi)
        "pcie_inclusive": {
            tag: {"host_copy_ms": o["host_copy_ms"],
                  "steps_per_s": args.steps / (o["sec"] + 1e-3 * sum(o["host_copy_ms"].values()))}
            for tag, o in (("fp64", f64), ("fp32", f32))},
        "autotuned_tiles": {"passes": ["fwd_z", "fwd_y", "inv_y", "inv_z_update",
                                       "iteration (0 five kernels, 1 with pipelined x pass, 2 three-kernel plane)"],
                            "fp64": f64["roofline"]["tiles"], "fp32": f32["roofline"]["tiles"]},
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
