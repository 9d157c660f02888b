"""pfdd_gl_xi_sums_noalloc (no device allocation, fixed device scratch)
against pfdd_gl_xi_sums, through the C ABI.

Two calls in a row on the shared scratch, the second with a smaller input
(fewer systems, smaller grid), each bit-identical to the allocating entry
point and within 1e-12 relative of numpy's sum of the field read back.
"""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


def _check(g, S):
    ref = g.xi_sums()
    got = g.xi_sums_noalloc()
    assert got.shape == (S,)
    np.testing.assert_array_equal(got, ref)
    host = g.get_xi().real.reshape(S, -1).sum(axis=1)
    assert np.max(np.abs(got - host)) <= 1e-12 * max(1.0, np.max(np.abs(host)))
    return got


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_sums_match_allocating_entry_point(precision):
    print(f"noalloc: 32^3 x 12 then 16x16x8 x 3, precision {precision}", flush=True)
    big = P.make_case(shape=(32, 32, 32), systems="fcc12", noise=1e-2)
    small = P.make_case(shape=(16, 16, 8), systems="basal3", noise=5e-2, seed=77)
    g1 = P.gpu_solver(big, precision=precision)
    g2 = P.gpu_solver(small, precision=precision)
    try:
        g1.step(3)
        s1 = _check(g1, 12)
        g2.step(2)
        s2 = _check(g2, 3)
        # the second, smaller call left the first handle's sums unchanged
        np.testing.assert_array_equal(g1.xi_sums_noalloc(), s1)
        assert not np.array_equal(s1[:3], s2)
    finally:
        g1.close()
        g2.close()


def test_noalloc_rejects_handles_beyond_its_scratch():
    import pfdd_gl
    print("noalloc: over-limit batch handle is rejected", flush=True)
    case = P.make_case(shape=(8, 8, 8), systems="one", noise=1e-2)
    m = case["mat"]
    nb = pfdd_gl.NOALLOC_MAX_SYSTEMS + 1
    g = pfdd_gl.GLSolver(case["shape"], case["xn"], case["xb"], case["geo"]["four"], m["C44"], m["C12"],
                         m["C11"], m["dslip"], case["cddt"], batch=nb)
    try:
        with pytest.raises(pfdd_gl.PfddError) as ei:
            g.xi_sums_noalloc()
        assert ei.value.code == 2
        assert "scratch limit" in str(ei.value)
        assert g.xi_sums().shape == (nb,)
    finally:
        g.close()
