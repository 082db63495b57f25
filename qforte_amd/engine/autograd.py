"""UCC energy as a differentiable torch function.

Forward: E(t) from a compiled program.  Backward: the hand-written adjoint
gradient: sigma = H psi, then a reverse sweep that un-applies each generator
on psi and sigma, giving dE/dt_k = 2 Re <sigma_k| K_k |psi_k> -- the
recursion of the reference's ``UCCVQE.measure_gradient``
(src/qforte/abc/uccvqeabc.py:122-251).  The program is a full-space
``UCCProgram`` (``qfe_program_gradient``) or a ``SectorProgram`` over a
determinant-sector template (``qfe_template_gradient``).  PyTorch is used only
for this binding; all arithmetic runs in the engine.
"""

import numpy as np
import torch


class SectorProgram:
    """A ``SectorTemplate`` bound to one request's Hamiltonian inputs, with
    the ``energy(t)`` / ``gradient(t)`` interface of ``UCCProgram``."""

    def __init__(self, template, inputs=()):
        self.template = template
        self.inputs = [float(v) for v in inputs]

    @property
    def nparam(self):
        return self.template.nparam

    def energy(self, t):
        return self.template.energy(list(t), self.inputs)

    def gradient(self, t):
        return self.template.gradient(list(t), self.inputs)


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
