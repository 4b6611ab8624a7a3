"""torch.autograd.Function over the discrete adjoint, for checking gradients.

J(mu_w, Ka) = 1/2 sum_fluid (rho^N - target)^2 after N steps of the FP64 step
from a fixed initial density; the backward pass returns the adjoint gradients
computed by cpmp_adjoint (reverse sweep with checkpoints).  The forward value
and gradients both come from the HIP library; torch only carries them.
"""
import numpy as np
import torch

from . import Lattice, make_params


class WallObjective(torch.autograd.Function):
    """apply(mu_lo, ka, setup) -> J

    mu_lo: (nx*ny,) float64 tensor, bottom wall chemical potential field
    ka:    0-d float64 tensor, kappa of mu = K^2 mu_eos - Ka lap(rho)
    setup: dict with dims, params (make_params keywords), rho0, target, nsteps,
           checkpoint_every, mu_hi (top wall field or scalar)
    """

    @staticmethod
    def forward(ctx, mu_lo, ka, setup):
        kw = dict(setup["params"])
        kw["Ka"] = float(ka.item())
        lat = Lattice(make_params(*setup["dims"], **kw))
        try:
            lat.set_wall_mu(mu_lo.detach().cpu().numpy().astype(np.float64), setup.get("mu_hi", 0.0))
            lat.init_from_rho(setup["rho0"])
            J, glo, _, gka = lat.adjoint(setup["nsteps"], setup["target"], setup.get("checkpoint_every", 100))
        finally:
            lat.close()
        ctx.save_for_backward(torch.from_numpy(glo), torch.tensor(gka, dtype=torch.float64))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        glo, gka = ctx.saved_tensors
        return grad_out * glo, grad_out * gka, None


def wall_objective(mu_lo, ka, setup):
    return WallObjective.apply(mu_lo, ka, setup)
