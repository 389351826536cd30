"""torch.autograd bridge for the batched adjoint gradients.

`AnsatzEnergy.apply(params, ansatz)` maps a [batch, n_params] float64
tensor (any device) to the [batch] energies <psi_b|H|psi_b> of a
`mettleq_amd.ansatz.LayeredAnsatz`. When the parameters need a gradient,
the forward call runs the HIP engine once for energies and the full
Jacobian (adjoint method, one graph launch) and backward contracts it with
the incoming gradient; otherwise it runs the forward pass alone. Torch
carries the tensors;
all statevector work is the engine's.
"""
from __future__ import annotations

import numpy as np
import torch

from .ansatz import LayeredAnsatz


class AnsatzEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params: torch.Tensor, ansatz: LayeredAnsatz):
        p = params.detach().to("cpu", torch.float64).contiguous().numpy()
        squeeze = p.ndim == 1
        if not ctx.needs_input_grad[0]:
            e = ansatz.evaluate(np.atleast_2d(p), grads=False)
            out = torch.from_numpy(e).to(params.device, params.dtype)
            return out[0] if squeeze else out
        e, g = ansatz.evaluate(np.atleast_2d(p), grads=True)
        jac = torch.from_numpy(g).to(params.device, params.dtype)
        out = torch.from_numpy(e).to(params.device, params.dtype)
        if squeeze:
            jac = jac[0]
            out = out[0]
        ctx.save_for_backward(jac)
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (jac,) = ctx.saved_tensors
        if jac.dim() == 1:
            return grad_out * jac, None
        return grad_out[:, None] * jac, None


def ansatz_energy(params: torch.Tensor, ansatz: LayeredAnsatz) -> torch.Tensor:
    return AnsatzEnergy.apply(params, ansatz)
