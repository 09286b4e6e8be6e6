"""torch.autograd.Function over the settling-velocity adjoint
(lbm_adjoint_velocity in include/lbm_amd.h): the loss is the mean vertical
velocity of one disk over the last steps of a run, the inputs are
theta = (tau_y, mu0, rho_p).  The forward runs the simulation and the
adjoint together on the GPU; backward returns the stored gradient."""
import torch

from .solver import D2Q9


class SettlingVelocity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, build, nsteps, window, checkpoint):
        """theta: float64 tensor (3,); build(tau_y, mu0, rho_p) -> (cfg, part)."""
        tau_y, mu0, rho_p = (float(v) for v in theta.detach().cpu().tolist())
        cfg, part = build(tau_y, mu0, rho_p)
        sim = D2Q9(cfg)
        try:
            sim.initialize()
            sim.set_particles(part)
            loss, grad, _ = sim.adjoint_velocity(nsteps, window, checkpoint)
        finally:
            sim.close()
        ctx.save_for_backward(torch.tensor(grad, dtype=theta.dtype, device=theta.device))
        return theta.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None, None


def settling_velocity(theta, build, nsteps, window, checkpoint):
    return SettlingVelocity.apply(theta, build, nsteps, window, checkpoint)
