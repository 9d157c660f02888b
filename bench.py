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


def solver_for(prob, precision):
    s = pfdd_gl.GLSolver(prob["shape"], prob["xn"], prob["xb"], prob["four"], prob["C44"],
                         prob["C12"], prob["C11"], prob["dslip"], prob["cddt"], S12=prob["S12"],
                         S44=prob["S44"], core_model=pfdd_gl.CORE_SINE, coef=prob["core_coef"],
                         precision=precision)
    s.set_xi(prob["xi0"])
    s.set_tau(prob["tau"])
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


def roofline(s, prof_steps=20):
    prof = s.profile(prof_steps)  # ms per kernel, HIP events on the solver's stream
    kbytes, step_bytes = s.traffic()
    names = ["fwd_z", "fwd_y", "x_contract", "inv_y", "inv_z_update"]
    times = [prof[k] for k in names]
    dom = int(np.argmax(times))
    achieved = kbytes[dom] / (times[dom] * 1e-3)
    per_kernel = {names[q]: {"ms": times[q], "bytes": kbytes[q],
                             "GBps": kbytes[q] / (times[q] * 1e-3) / 1e9} for q in range(5)}
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
    }


def load_pmc_traffic(tag, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated with a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{tag}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel)
        return k["hbm_bytes_per_launch"] if k else None
    except Exception:
        return None


KERNEL_ROLES = {"k_mid_otf": "x_contract", "k_mid<": "x_contract", "k_update": "inv_z_update"}


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
        s.step(3)
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

    def per_kernel(rows, match):
        v = [val for (d, name), val in rows.items() if match(name)]
        return sum(v) / len(v) if v else None

    is_copy = lambda n: "elementwise" in n and "Mul" in n
    copy_f = per_kernel(fetch, is_copy)
    copy_w = per_kernel(write, is_copy)
    copy_name = next((name for (d, name) in fetch if is_copy(name)), None)
    sf = nbytes / (copy_f * 1024.0) if copy_f else 1.0
    sw = nbytes / (copy_w * 1024.0) if copy_w else 1.0
    prob = make_workload(size)
    out = {}
    for prec, tag, tname in ((pfdd_gl.PFDD_FP64, "fp64", "double"), (pfdd_gl.PFDD_FP32, "fp32", "float")):
        s = solver_for(prob, prec)
        kbytes, _ = s.traffic()
        s.close()
        algo = dict(zip(["fwd_z", "fwd_y", "x_contract", "inv_y", "inv_z_update"], kbytes.tolist()))
        res = {"copy_calibration": {"kernel": copy_name, "bytes_read": nbytes, "bytes_written": nbytes,
                                    "fetch_kib": copy_f, "write_kib": copy_w,
                                    "fetch_scale": sf, "write_scale": sw}, "kernels": {}}
        for sub, role in KERNEL_ROLES.items():
            m = lambda n, sub=sub: sub in n and f"<{tname}," in n
            f, w = per_kernel(fetch, m), per_kernel(write, m)
            if f is None or w is None:
                continue
            res["kernels"][role] = {"fetch_bytes": f * 1024.0 * sf, "write_bytes": w * 1024.0 * sw,
                                    "hbm_bytes_per_launch": f * 1024.0 * sf + w * 1024.0 * sw,
                                    "algorithmic_bytes": algo[role]}
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
    cores = min(16, os.cpu_count() or 1)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="workload for rocprofv3 --pmc runs (copy calibration + 3 steps)")
    ap.add_argument("--pmc-summarize", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="turn two rocprofv3 --pmc outputs into pmc_traffic_*.json")
    ap.add_argument("--pmc-out", default=os.path.join(ROOT, "profiles"))
    args = ap.parse_args()
    if args.pmc_probe:
        return pmc_probe(args.size)
    if args.pmc_summarize:
        return pmc_summarize(args.pmc_summarize[0], args.pmc_summarize[1], args.size, args.pmc_out)
    if args.gpus != 1:
        print(f"bench.py: this build runs the GL solver on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU slab decomposition is a design note, not built.", file=sys.stderr)
        return 2

    prob = make_workload(args.size)
    out = {}
    for prec, tag in ((pfdd_gl.PFDD_FP64, "fp64"), (pfdd_gl.PFDD_FP32, "fp32")):
        s = solver_for(prob, prec)
        dt, r = time_steps(s, args.steps, args.warmup)
        xi_end = s.get_xi() if tag else None
        roof = roofline(s)
        roof["traffic"] = load_pmc_traffic(tag, roof["kernel"])
        out[tag] = {"sec": dt, "steps_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                    "roofline": roof, "e_total": r["e_total"], "xi_end": xi_end}
        s.close()
    # FP32 error against the FP64 GPU run of the same length
    x64, x32 = out["fp64"].pop("xi_end"), out["fp32"].pop("xi_end")
    err_xi = float(np.max(np.abs(x32 - x64)) / np.max(np.abs(x64)))
    err_e = abs(out["fp32"]["e_total"] - out["fp64"]["e_total"]) / abs(out["fp64"]["e_total"])

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(prob, args.cpu_steps)

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
                   "slip_systems": 12, "precision": "FP64 plain (value); FP32 mode below"},
        "roofline": {k: f64["roofline"][k] for k in
                     ("kernel", "bound", "achieved", "peak", "unit", "frac", "traffic")},
        "cpu_baseline": cpu,
        "fp32": {"value": f32["steps_per_s"], "ms_per_step": f32["ms_per_step"],
                 "speedup_vs_fp64": f32["steps_per_s"] / f64["steps_per_s"],
                 "rel_err_xi_vs_fp64": err_xi, "rel_err_energy_vs_fp64": err_e,
                 "roofline": {k: f32["roofline"][k] for k in
                              ("kernel", "achieved", "frac", "traffic")}},
        "step_roofline": {
            "fp64": {"bytes": f64["roofline"]["step_bytes"],
                     "frac": f64["roofline"]["step_bytes"] / (f64["ms_per_step"] * 1e-3) / PEAK_HBM},
            "fp32": {"bytes": f32["roofline"]["step_bytes"],
                     "frac": f32["roofline"]["step_bytes"] / (f32["ms_per_step"] * 1e-3) / PEAK_HBM}},
        "per_kernel": {"fp64": f64["roofline"]["per_kernel"], "fp32": f32["roofline"]["per_kernel"]},
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
