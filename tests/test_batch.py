"""Batched runs (pfdd_gl_create_batch) and hipGraph replay (pfdd_gl_graph_*).

A batch member runs exactly the arithmetic of a single-run handle, so its
field must equal, bit for bit, that of a separate handle given the same
inputs; a graph replay must equal the same iterations launched eagerly.
The single-run handles themselves are pinned against the oracle by
test_gpu_parity.py.
"""
import numpy as np
import pytest

import problems as P


def member_cases(shape, n, masked=False):
    cases = []
    nx, ny, nz = shape
    for m in range(n):
        xo = None
        if masked:
            xo = np.zeros((3,) + tuple(shape))
            xo[m % 3, (m + 1) * nx // 5, :, nz // 2] = 1.0
        c = P.make_case(shape=shape, noise=1e-2, sigma_mu=0.002 + 0.002 * m, seed=576 + 17 * m,
                        loops=[(m % 3, 2, nz // 2, nx / 2 - m, ny / 2 + m, max(2, nx // 5))], xo=xo)
        cases.append(c)
    return cases


def batch_solver(cases, precision=0):
    import pfdd_gl
    c0, m = cases[0], cases[0]["mat"]
    xo = None
    if any(c["xo"] is not None for c in cases):
        xo = np.concatenate([c["xo"] if c["xo"] is not None else np.zeros((c["S"],) + c["shape"])
                             for c in cases])
    s = pfdd_gl.GLSolver(c0["shape"], c0["xn"], c0["xb"], c0["geo"]["four"], m["C44"], m["C12"], m["C11"],
                         m["dslip"], c0["cddt"], S12=m["S12"], S44=m["S44"], xo=xo, precision=precision,
                         batch=len(cases), **c0["gpu_core"])
    s.set_xi(np.concatenate([c["xi0"] for c in cases]))
    s.set_tau(np.concatenate([c["tau"] for c in cases]))
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("precision,masked", [(0, False), (1, False), (0, True)])
def test_batch_members_equal_single_runs(precision, masked):
    print(f"case: batch of 3 x (16^3 x 3) vs single handles, precision {precision}, masked {masked}",
          flush=True)
    cases = member_cases((16, 16, 16), 3, masked)
    b = batch_solver(cases, precision)
    b.step(25)
    xb = b.get_xi()
    sums = b.xi_sums()
    b.close()
    for m, c in enumerate(cases):
        s = P.gpu_solver(c, precision=precision)
        s.step(25)
        x = s.get_xi()
        ss = s.xi_sums()
        s.close()
        assert np.array_equal(xb[3 * m:3 * m + 3], x), m
        assert np.allclose(sums[3 * m:3 * m + 3], ss, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_graph_replay_equals_eager_steps():
    print("case: hipGraph over batch 2x16^3, batch 3x32^3 and a single 16^3 handle; 21 + 10 + 6 steps",
          flush=True)
    import pfdd_gl
    ca = member_cases((16, 16, 16), 2)
    cb = member_cases((32, 32, 32), 3)
    cc = member_cases((16, 16, 16), 1)[0]

    def make():
        return [batch_solver(ca), batch_solver(cb, precision=1), P.gpu_solver(cc)]

    eager, graphed = make(), make()
    g = pfdd_gl.GLGraph(graphed, iters=4)
    res = g.run(21)
    for s in eager:
        s.step(21)
    for s, t, r in zip(eager, graphed, res):
        assert r["steps_done"] == 21
        assert np.array_equal(s.get_xi(), t.get_xi())
    # new stresses: device tau for the batches (no new capture), kernel
    # argument for the single handle (new capture)
    new_a = np.concatenate([c["tau"] for c in ca]) * 1.5
    new_c = np.asarray(cc["tau"]) * -2.0
    for grp in (eager, graphed):
        grp[0].set_tau(new_a)
        grp[2].set_tau(new_c)
    g.run(10)
    for s in eager:
        s.step(10)
    for s, t in zip(eager, graphed):
        assert np.array_equal(s.get_xi(), t.get_xi())
    # an odd number of eager steps between graph runs
    for grp in (eager, graphed):
        grp[1].step(1)
    g.run(6)
    for s in eager:
        s.step(6)
    for s, t in zip(eager, graphed):
        assert np.array_equal(s.get_xi(), t.get_xi())
    g.close()
    for s in eager + graphed:
        s.close()


@pytest.mark.gpu
def test_batch_rejects_single_run_only_calls():
    print("case: batch handle error paths", flush=True)
    import pfdd_gl
    cases = member_cases((16, 16, 16), 2)
    b = batch_solver(cases)
    with pytest.raises(pfdd_gl.PfddError, match="UNSUPPORTED"):
        b.diag(np.zeros((3, 3)))
    with pytest.raises(pfdd_gl.PfddError, match="UNSUPPORTED"):
        b.tape_record(2, 1)
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_results(precision):
    print(f"case: autotune 32^3 x 3 precision {precision}: fields equal to the default tiles, "
          "state restored", flush=True)
    c = member_cases((32, 32, 32), 1)[0]
    s0 = P.gpu_solver(c, precision=precision)
    s1 = P.gpu_solver(c, precision=precision)
    s1.step(3)
    x_before = s1.get_xi()
    sig = np.diag([0.0, 0.0, 0.0]) + 1e-3 * np.eye(3)
    # the output-step strain is computed from the field before the last update
    # (the other ping-pong buffer): tuning restores both
    f_before = s1.diag_fields(sig)
    choice = s1.autotune(3)
    # 32-point lines have one radix plan, so the line-pass choices stay 0-2
    assert len(choice) == 5 and all(q in (0, 1, 2) for q in choice[:4]) and choice[4] in (0, 1, 2)
    assert np.array_equal(s1.get_xi(), x_before)  # tuning restores the field
    f_after = s1.diag_fields(sig)
    for k, (a, b) in enumerate(zip(f_before, f_after)):
        assert np.array_equal(a, b), k
    s0.step(3)
    s0.step(17)
    s1.step(17)
    assert np.array_equal(s0.get_xi(), s1.get_xi())
    # the reverse mode runs with the tuned tiles
    s1.tape_record(4, 2)
    lam, dtau, _ = s1.tape_backward(np.ones((3, 32, 32, 32), dtype=np.complex128), 0)
    assert np.all(np.isfinite(lam)) and np.all(np.isfinite(dtau))
    s0.close()
    s1.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,shape,systems", [(0, (32, 32, 32), "basal3"), (1, (32, 32, 32), "basal3"),
                                                     (0, (16, 16, 16), "fcc12"), (1, (48, 32, 16), "basal3"),
                                                     (0, (16, 128, 128), "basal3"), (1, (16, 128, 128), "basal3")])
def test_every_tuning_choice_gives_the_same_field(precision, shape, systems):
    """Tile shapes and iteration variants: the same field bit for bit.  The
    radix-16 plan of the line passes (choices 3-5; lengths 16, 48, 128 here
    have one) rounds differently: within 1e-12 (FP64) / 1e-5 (FP32) of the
    radix-8 plan after 12 iterations, and bit for bit across tile shapes."""
    print(f"case: tuning choices {shape} {systems} precision {precision}", flush=True)
    c = P.make_case(shape=shape, systems=systems, noise=1e-2, sigma_mu=0.004)
    ref = P.gpu_solver(c, precision=precision)
    ref.set_tuning([1, 1, 1, 1, 0])
    ref.step(12)
    x_ref = ref.get_xi()
    ref.close()
    # the three-kernel plane iteration needs square y-z planes that fit in LDS
    plane = shape[1] == shape[2] and shape[1] <= (80 if precision == 0 else 128)
    for choice in ([0, 0, 0, 0, 0], [2, 2, 2, 2, 0], [1, 1, 1, 1, 1], [0, 2, 1, 0, 1]) + (([1, 1, 1, 1, 2],) if plane else ()):
        s = P.gpu_solver(c, precision=precision)
        s.set_tuning(choice)
        s.step(12)
        assert np.array_equal(s.get_xi(), x_ref), choice
        s.close()
    x_p1 = None
    tol = 1e-12 if precision == 0 else 1e-5
    for choice in ([3, 3, 3, 1, 0], [5, 4, 3, 0, 1]):
        s = P.gpu_solver(c, precision=precision)
        s.set_tuning(choice)
        s.step(12)
        x = s.get_xi()
        s.close()
        err = float(np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)))
        print(f"radix-16 plan {choice}: rel. difference from the radix-8 plan {err:.3e}", flush=True)
        assert err < tol, (choice, err)
        if x_p1 is None:
            x_p1 = x
        else:
            assert np.array_equal(x, x_p1), choice


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_step_graph_replay_equals_single_steps(precision):
    print(f"case: pfdd_gl_step graph replay (45 and 20 iterations) vs single iterations, precision {precision}",
          flush=True)
    c = P.make_case(shape=(32, 32, 32), noise=1e-2, sigma_mu=0.004)
    a = P.gpu_solver(c, precision=precision)
    b = P.gpu_solver(c, precision=precision)
    try:
        ra = a.step(45)  # four replays of a 10-iteration graph, then 5 eager iterations
        for _ in range(45):
            rb = b.step(1)
        assert ra["steps_done"] == 45
        assert np.array_equal(a.get_xi(), b.get_xi())
        assert ra["e_total"] == rb["e_total"] and ra["xinorm"] == rb["xinorm"]
        # new kernel arguments (tau) and a new tuning: the graph is captured again
        tau2 = 1.5 * np.asarray(c["tau"])
        a.set_tau(tau2)
        b.set_tau(tau2)
        a.step(20)
        for _ in range(20):
            b.step(1)
        assert np.array_equal(a.get_xi(), b.get_xi())
        a.set_tuning([0, 0, 0, 0, 1])
        a.step(30)
        for _ in range(30):
            b.step(1)
        assert np.array_equal(a.get_xi(), b.get_xi())
        # an odd number of eager iterations switches the ping-pong parity; the
        # next long run replays the graph captured for the other parity
        for n in (3, 20, 1, 20):
            a.step(n)
            for _ in range(n):
                b.step(1)
            assert np.array_equal(a.get_xi(), b.get_xi()), n
    finally:
        a.close()
        b.close()
