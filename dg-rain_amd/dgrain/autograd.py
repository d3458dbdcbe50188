"""torch.autograd binding of the gauge-misfit gradient.

GaugeProblem fixes a mesh, solver parameters, an initial state, gauge
elements, observations and the run length; gauge_loss(problem, z, n) returns
the misfit as a torch scalar whose backward is the hand-written reverse of
the RK step on the GPU (libdgrain's dgr_adjoint_gauges): one call computes the
loss and both gradients, which backward then scales by the incoming gradient.

z and n are float64 tensors of nodal depth and Manning n (any device; the
library takes them as host arrays)."""
from dataclasses import dataclass

import numpy as np
import torch

from . import native


@dataclass
class GaugeProblem:
    mesh: object
    params: object            # native.Params
    zeta: np.ndarray
    qx: np.ndarray
    qy: np.ndarray
    gauge_el: np.ndarray      # member-local element ids
    obs: np.ndarray           # [nsteps // sample_every, ngauges]
    nsteps: int
    sample_every: int = 10
    checkpoint_every: int = 50

    def solver(self, z=None, n=None):
        s = native.Solver(self.mesh, self.params)
        s.set_nodal(z=z, n=n)
        s.set_state(self.zeta, self.qx, self.qy)
        return s

    def loss(self, z=None, n=None):
        """The misfit by a plain forward run."""
        s = self.solver(z, n)
        h = s.gauge_run(self.nsteps, self.sample_every, self.gauge_el)
        s.close()
        return float(((h[0] - self.obs) ** 2).sum())

    def loss_and_grads(self, z=None, n=None):
        s = self.solver(z, n)
        loss, gz, gn = s.adjoint_gauges(self.nsteps, self.sample_every, self.checkpoint_every, self.gauge_el,
                                        self.obs[None])
        s.close()
        return float(loss[0]), gz, gn


class _GaugeLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, n, problem):
        zn = z.detach().to("cpu", torch.float64).numpy()
        nn_ = n.detach().to("cpu", torch.float64).numpy()
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            loss, gz, gn = problem.loss_and_grads(zn, nn_)
            ctx.save_for_backward(torch.from_numpy(gz).to(z.device), torch.from_numpy(gn).to(n.device))
        else:
            loss = problem.loss(zn, nn_)
        return torch.tensor(loss, dtype=torch.float64, device=z.device)

    @staticmethod
    def backward(ctx, grad_out):
        gz, gn = ctx.saved_tensors
        return grad_out * gz, grad_out * gn, None


def gauge_loss(problem, z, n):
    """The gauge misfit of `problem` with nodal depth z and Manning n."""
    return _GaugeLoss.apply(z, n, problem)
