"""torch.autograd.Function around the GPU PISO rollout (test harness only).

forward(coef, viscosity) runs nsteps PISO steps from rest with inlet values
g = sum_k coef_k basis_k and returns the final velocity; backward runs the
library's hand-written adjoint sweep (cfsa_ns_adjoint_backward)."""
import numpy as np
import torch


class PisoRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coef, viscosity, solver, basis, nsteps):
        solver.setViscosity(float(viscosity))
        g = np.tensordot(coef.detach().cpu().numpy(), basis, axes=1)
        u = solver.adjointForward(int(nsteps), g)
        ctx.solver = solver
        ctx.basis = basis
        return torch.from_numpy(u)

    @staticmethod
    def backward(ctx, grad_u):
        dg, dmu = ctx.solver.adjointBackward(grad_u.detach().cpu().numpy())
        dcoef = np.tensordot(ctx.basis, dg, axes=([1, 2], [0, 1]))
        return (torch.from_numpy(dcoef), torch.tensor(dmu, dtype=torch.float64), None, None,
                None)


def outlet_energy(u, outlet_mask):
    return 0.5 * (u[outlet_mask] ** 2).sum()
