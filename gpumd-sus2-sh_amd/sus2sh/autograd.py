"""torch.autograd binding of the engine: total energy -> (positions, model parameters).

The forward pass runs the HIP engine once with the parameter adjoint enabled
(sus2sh_compute_param_grad) and keeps forces and dE/dtheta; the backward pass returns
dE/dpositions = -F and dE/dtheta scaled by the incoming gradient.  This is a test binding
for gradient checks and training-style use; MD uses sus2sh_compute directly.
"""
import numpy as np
import torch

from .potential import SUS2_SH, box_struct


class _EnergyFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, theta, engine, box, types):
        pot = engine.pot
        n = int(types.numel())
        theta_np = theta.detach().cpu().numpy().astype(np.float64)
        if engine.theta is None or not np.array_equal(theta_np, engine.theta):
            pot.set_params(theta_np)
            engine.theta = theta_np.copy()
        dev = positions.device
        e = torch.zeros(n, dtype=torch.float64, device=dev)
        f = torch.zeros(3, n, dtype=torch.float64, device=dev)
        v = torch.zeros(9, n, dtype=torch.float64, device=dev)
        g = torch.zeros(len(theta_np), dtype=torch.float64, device=dev)
        pot.compute(box, types, positions.detach().contiguous(), e, f, v, param_grad=g)
        ctx.save_for_backward(f, g)
        ctx.theta_device = theta.device
        return e.sum()

    @staticmethod
    def backward(ctx, grad_e):
        f, g = ctx.saved_tensors
        return -grad_e * f, (grad_e * g).to(ctx.theta_device), None, None, None


class EnergyModel(torch.nn.Module):
    """E(positions, theta) for one model file and one cell; theta starts at the file values."""

    def __init__(self, model_path, cell, num_atoms, symbols, options=(), pbc=(1, 1, 1)):
        super().__init__()
        self.pot = SUS2_SH(model_path, num_atoms, list(symbols) + list(options))
        self.box = box_struct(cell, pbc)
        self.theta = None
        self.params = torch.nn.Parameter(torch.as_tensor(self.pot.get_params(), dtype=torch.float64))

    def forward(self, positions, types):
        return _EnergyFn.apply(positions, self.params, self, self.box, types)
