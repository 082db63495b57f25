"""UCC energy as a differentiable torch function.

Forward: E(t) from the compiled program (one sweep of exponentials and one
grouped expectation pass per X mask).  Backward: the hand-written adjoint
gradient (``qfe_program_gradient``): sigma = H psi, then a reverse sweep that
un-applies each exponential on psi and sigma with one extra state, giving
dE/dt_k = 2 Re <sigma_k| K_k |psi_k> -- the recursion of the reference's
``UCCVQE.measure_gradient`` (src/qforte/abc/uccvqeabc.py:122-251).
PyTorch is used only for this binding; all arithmetic runs in the engine.
"""

import numpy as np
import torch


class UCCEnergy(torch.autograd.Function):
    """``UCCEnergy.apply(t, program)`` -> 0-d float64 tensor."""

    @staticmethod
    def forward(ctx, t, program):
        if t.dtype != torch.float64:
            raise TypeError("UCCEnergy: amplitudes must be float64")
        vals = t.detach().cpu().numpy().astype(np.float64).tolist()
        ctx.program = program
        ctx.vals = vals
        return torch.tensor(program.energy(vals), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        _, g = ctx.program.gradient(ctx.vals)
        grad = torch.tensor(g, dtype=torch.float64) * grad_out
        return grad, None
