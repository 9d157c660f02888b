"""torch.autograd wrapper of the GPU GL solver's reverse mode.

PyTorch is only the autograd harness here: the forward run, the checkpoints,
the recomputation and the adjoint iteration all run in the HIP library
(pfdd_gl_tape_record / pfdd_gl_tape_backward_device of include/pfdd_gl.h).

    xi_T = gl_rollout(solver, xi0, tau, coef, nsteps, every)

xi0: complex128 CUDA tensor (S, nx, ny, nz) in the layout of the reference's
fft->xi[0]; tau: float64 (S,) resolved shear stresses in units of mu (e.g.
computed from an applied stress tensor with differentiable torch ops);
coef: float64 (S, ncoef) sine-series core coefficients
(dE = sum_m c_m m pi sin(2 m pi Re xi)).  Gradients flow to all three.  The
gradient convention is torch's for a real loss: dL/dRe + i dL/dIm.
"""
from __future__ import annotations

import numpy as np
import torch

import pfdd_gl


class GLRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi0, tau, coef, solver, nsteps, every):
        if not xi0.is_cuda:
            raise ValueError("xi0 must be a CUDA tensor")
        ncoef = int(coef.shape[-1])
        solver.set_tau(tau.detach().double().cpu().numpy())
        solver.set_core_coef(coef.detach().double().cpu().numpy().reshape(-1), ncoef)
        fp32 = solver.precision == pfdd_gl.PFDD_FP32
        x = xi0.detach().to(torch.complex64 if fp32 else torch.complex128).contiguous()
        torch.cuda.current_stream().synchronize()
        solver.set_xi_device(x.data_ptr())
        solver.tape_record(int(nsteps), int(every))
        out = torch.empty_like(x)
        solver.get_xi_device(out.data_ptr())
        ctx.solver, ctx.ncoef = solver, ncoef
        ctx.tau_dtype, ctx.coef_dtype = tau.dtype, coef.dtype
        return out.to(torch.complex128)

    @staticmethod
    def backward(ctx, g):
        g = g.detach().to(torch.complex128).contiguous()
        dxi0 = torch.empty_like(g)
        torch.cuda.current_stream().synchronize()
        dtau, dcoef = ctx.solver.tape_backward_device(g.data_ptr(), dxi0.data_ptr(), ctx.ncoef)
        dev = g.device
        return (dxi0, torch.as_tensor(dtau, device=dev, dtype=ctx.tau_dtype),
                torch.as_tensor(np.ascontiguousarray(dcoef), device=dev, dtype=ctx.coef_dtype),
                None, None, None)


def gl_rollout(solver, xi0, tau, coef, nsteps, every=50):
    return GLRollout.apply(xi0, tau, coef, solver, nsteps, every)
