"""Writes gl_fixtures.npz: FP64 oracle results of a few small seeded GL runs.

The reference ships no test vectors, so these fixtures are produced by the
oracle (the C restatement in oracle/, itself pinned by closed forms and
finite differences in tests/test_oracle.py) and committed, so that later
changes to either the oracle or the GPU path are checked against fixed
numbers.  Inputs are regenerated from the seeded recipe in tests/problems.py;
the fixture stores a checksum of the initial field to prove that the recipe
still gives the same input.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import problems as P  # noqa: E402

SKEW = ((1, 0, 0), (0.5, 0.8660254037844386, 0), (0, 0, 1))


def cases():
    """name -> (make_case kwargs, steps)."""
    xo = np.zeros((12, 8, 8, 8))
    xo[0, 2, 3, 4] = xo[0, 6, 3, 4] = 1.0
    return {
        "fcc12_skewed_masked_sine_8cubed": (dict(shape=(8, 8, 8), systems="fcc12", noise=1e-2, sigma_mu=0.004,
                                                 prim_vectors=SKEW, xo=xo,
                                                 loops=[(0, 2, 4, 4.0, 4.0, 3)]), 4),
        "basal3_series_8x8x48": (dict(shape=(8, 8, 48), systems="basal3", noise=1e-2, sigma_mu=0.006,
                                      core="series", ncoef=4), 6),
        "basal3_extended_8cubed": (dict(shape=(8, 8, 8), systems="basal3", noise=1e-2, sigma_mu=0.006,
                                        core="extended"), 5),
    }


def checksum(x):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    w = np.arange(1, x.size + 1, dtype=np.float64).reshape(x.shape)
    return np.array([np.sum(x.real * w), np.sum(x.imag * w), np.sum(np.abs(x) ** 2)])


def run_oracle(kw, steps):
    c = P.make_case(**kw)
    st = P.oracle_state(c)
    r = None
    for _ in range(steps):
        r = st.step(c["tau"])
    return c, st.xi.copy(), r


def main():
    out = {}
    for name, (kw, steps) in cases().items():
        c, xi, r = run_oracle(kw, steps)
        out[name + "/xi0_checksum"] = checksum(c["xi0"])
        out[name + "/xi"] = xi
        out[name + "/scalars"] = np.array([steps, r["xinorm"], r["xiave"], r["E_total"]])
        print(name, xi.shape, r["xinorm"], r["E_total"])
    np.savez_compressed(os.path.join(HERE, "gl_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
