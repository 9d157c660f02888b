"""torch.autograd wrapper of the GPU GL solver's reverse mode.

PyTorch is only the autograd harness here: the forward run, the checkpoints,
the recomputation and the adjoint iteration all run in the HIP library
(pfdd_gl_tape_record / pfdd_gl_tape_backward of include/pfdd_gl.h), on the
solver's device.  The tensors that cross this boundary are host (CPU)
tensors: the library links the system HIP runtime, and a torch build that
bundles its own HIP runtime must not initialise a second one in the same
process.  Only the initial field, the final field and their gradients cross
(once per forward / backward call, not per GL iteration).

    xi_T = gl_rollout(solver, xi0, tau, coef, nsteps, every)

xi0: complex128 tensor (S, nx, ny, nz) in the layout of the reference's
fft->xi[0]; tau: float64 (S,) resolved shear stresses in units of mu (e.g.
computed from an applied stress tensor with differentiable torch ops);
coef: float64 (S, ncoef) sine-series core coefficients
(dE = sum_m c_m m pi sin(2 m pi Re xi)).  Gradients flow to all three, in
torch's convention for a real loss (dL/dRe + i dL/dIm).
"""
from __future__ import annotations

import numpy as np
import torch


def _host(t, dtype):
    if t.device.type != "cpu":
        raise ValueError("gl_rollout takes host (CPU) tensors; the solver keeps its fields on its own device")
    return t.detach().to(dtype).contiguous().numpy()


class GLRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi0, tau, coef, solver, nsteps, every):
        ncoef = int(coef.shape[-1])
        solver.set_tau(_host(tau, torch.float64))
        solver.set_core_coef(_host(coef, torch.float64).reshape(-1), ncoef)
        solver.set_xi(_host(xi0, torch.complex128))
        solver.tape_record(int(nsteps), int(every))
        ctx.solver, ctx.ncoef = solver, ncoef
        ctx.tau_dtype, ctx.coef_dtype = tau.dtype, coef.dtype
        return torch.from_numpy(solver.get_xi())

    @staticmethod
    def backward(ctx, g):
        dxi0, dtau, dcoef = ctx.solver.tape_backward(_host(g, torch.complex128), ctx.ncoef)
        return (torch.from_numpy(dxi0), torch.from_numpy(dtau).to(ctx.tau_dtype),
                torch.from_numpy(np.ascontiguousarray(dcoef)).to(ctx.coef_dtype), None, None, None)


def gl_rollout(solver, xi0, tau, coef, nsteps, every=50):
    return GLRollout.apply(xi0, tau, coef, solver, nsteps, every)
