"""Writes c0_ramp_64cubed.npz: the FP64 oracle's run of the 64^3 x 3 loop
under the 0 -> 300 MPa shear ramp (2000 stress states of one GL iteration
each, as solve_style GL max_iter 1 tol 0 with sigma ... delta runs it,
solve_GL.cpp:70-158, sigma += delta after every state :151-155).

The final field has 786432 complex values, too many to commit, so the
fixture keeps size-independent summaries of it that the GPU run must
reproduce: the slip plane k = 32 of every system, per-system sums of Re xi,
Im xi and |xi|^2, a weighted checksum, and 4096 cells drawn with a fixed
seed; plus xinorm / xiave / energies every 50 states and the output-step
averages ave_eps and ave_sigma of the last state (Output::compute after the
last update).  Inputs are regenerated from the seeded recipe in
tests/problems.py; a checksum of the initial field is stored with them.

Run from the repository root (about 3 minutes on 8 cores):
    python tests/golden/make_fullsize_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import problems as P  # noqa: E402  (puts oracle/ on the path)
import oracle  # noqa: E402

SHAPE = (64, 64, 64)
STATES = 2000
DELTA_MU = 3.125e-6   # sigma_xz increment per stress state (mu units): 300 MPa / 48 GPa / 2000
EVERY = 50
NSAMPLE = 4096
PATH = os.path.join(HERE, "c0_ramp_64cubed.npz")


def c0_case():
    """64^3, the three systems of basal3 on the plane n = z, Cu-like material,
    sine core, Re xi_1 = 1 inside r <= 12 of (32, 32) on k = 32, noise 1e-3,
    CD*dt = 0.01; the stress of state s is s * DELTA_MU on xz."""
    return P.make_case(shape=SHAPE, loops=[(0, 2, 32, 32.0, 32.0, 12)], noise=1e-3, sigma_mu=0.0)


def tau_of_state(case, s):
    sig = P.shear_sigma(s * DELTA_MU, "xz")
    g = case["geo"]
    srot = oracle.rotate_stress(g["prim"], g["invprim"], g["norvol"], sig)
    return oracle.resolss(srot, case["xn"], case["xb"], P.DSLIP), srot


def sample_index(S):
    rng = np.random.default_rng(20261203)
    return np.sort(rng.choice(S * int(np.prod(SHAPE)), NSAMPLE, replace=False))


def summarize(xi):
    """Summaries of a field of shape (S, 64, 64, 64)."""
    S = xi.shape[0]
    flat = xi.reshape(-1)
    w = np.arange(1, flat.size + 1, dtype=np.float64)
    return {
        "plane": np.ascontiguousarray(xi[:, :, :, SHAPE[2] // 2]),
        "sums": np.stack([xi.real.sum(axis=(1, 2, 3)), xi.imag.sum(axis=(1, 2, 3)),
                          (np.abs(xi) ** 2).sum(axis=(1, 2, 3))]),
        "checksum": np.array([np.sum(flat.real * w), np.sum(flat.imag * w)]),
        "sample": flat[sample_index(S)],
    }


def main():
    oracle.set_threads(os.cpu_count() or 1)
    case = c0_case()
    BB, eps, mod = P.oracle_B(case)
    st = P.oracle_state(case, BB)
    rows = []
    srot = None
    for s in range(STATES):
        tau, srot = tau_of_state(case, s)
        r = st.step(tau)
        if s % EVERY == EVERY - 1 or s == STATES - 1:
            rows.append([s + 1, r["xinorm"], r["xiave"], r["E_el"], r["E_core"], r["E_stress"], r["E_total"],
                         r["E_core_ref"]])
            print(f"state {s + 1}: xinorm {r['xinorm']:.6e} E_total {r['E_total']:.10e}", flush=True)
    m = case["mat"]
    FF = oracle.fmatrix(SHAPE, m["C44"], m["C12"], mod["mu"], mod["xnu"], eps, case["geo"]["four"])
    d = oracle.diagnostics(st.data_fftw, st.xi, FF, eps, m, mod["mu"], srot)
    summ = summarize(st.xi)
    out = {"xi0_checksum": np.array([np.sum(case["xi0"].real), np.sum(np.abs(case["xi0"]) ** 2)]),
           "scalars": np.array(rows), "ave_eps": d["ave_eps"], "ave_sigma": d["ave_sigma"],
           "sigma_rot_last": srot}
    out.update({"xi/" + k: v for k, v in summ.items()})
    np.savez_compressed(PATH, **out)
    print("wrote", PATH, os.path.getsize(PATH), "bytes")


if __name__ == "__main__":
    main()
