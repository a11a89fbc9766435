"""Gradients of ensemble SAR and final <m_z> with respect to (K, alpha, H, r).

`EnsembleTrajectory` is a torch.autograd.Function around the device passes
of `core.Adjoint` (include/magpy_amd.h, mp_adjoint_*): its forward returns
the sampled ensemble magnetisation <M_z>(t_s) (A/m) and the final mean
reduced m_z of an ensemble of independent particles, recording (when theta
needs a gradient) the hand-written adjoint of the Heun step over the same
Wiener increments as linear maps between sample steps; its backward applies
them to the incoming gradients.  `SARObjective` builds the
workload (one lognormal ensemble, sine field, the reference's sampling) and
evaluates the hysteresis energy the way magpy/results.py:167-217 does,
E = -mu0 * trapz(H, <M_z>) over the last field period, SAR = f * E, in torch
so that the field's own dependence on H is differentiated by autograd.

The reference has no derivatives; its forward model is the one of
simulation::full_dynamics for one-particle members (lib/simulation.cpp:
476-624), with the field phase taken in SI units at each step so that it
does not move with the parameters.
"""
import numpy as np
import torch

from . import core

MU0 = 1.25663706e-6


class EnsembleTrajectory(torch.autograd.Function):
    """(theta [4] float64, runner: core.Adjoint) -> (ens_mz [S], final_mz).

    backward needs the runner's most recent forward pass: a later forward on
    the same runner (another theta) makes it raise instead of returning the
    gradient of the wrong trajectory."""

    @staticmethod
    def forward(ctx, theta, runner):
        # the gradient is recorded only when theta needs one
        ens, fin, status = runner.forward(theta.detach().cpu().to(torch.float64).numpy(),
                                          record_gradient=bool(ctx.needs_input_grad[0]))
        if status:
            raise FloatingPointError("non-finite state in the forward pass (status %d)" % status)
        ctx.runner = runner
        ctx.pass_id = runner.forward_passes
        return (torch.from_numpy(np.array(ens, dtype=np.float64)),
                torch.tensor(fin, dtype=torch.float64))

    @staticmethod
    def backward(ctx, g_ens, g_fin):
        runner = ctx.runner
        if runner.forward_passes != ctx.pass_id:
            raise RuntimeError("the runner ran another forward pass since this one; "
                               "the adjoint maps recorded by this pass are gone")
        g = np.zeros(runner.n_samples) if g_ens is None else \
            g_ens.detach().cpu().to(torch.float64).numpy()
        gf = 0.0 if g_fin is None else float(g_fin)
        grad = runner.reverse(g, gf)
        return torch.from_numpy(grad), None


def hysteresis_energy(ens_mz, field, time, field_frequency):
    """-mu0 * trapz(H, M) over t >= t_end - 1/f (magpy/results.py:188-217)."""
    mask = time >= time[-1] - 1.0 / field_frequency
    return -MU0 * torch.trapezoid(field[mask], ens_mz[mask])


class SARObjective:
    """An ensemble of independent particles under a sine field, with the
    reference's sample grid, whose SAR and final <m_z> are differentiable in
    theta = (K [J/m^3], alpha, H [A/m], median radius [m]).

    radius_shape [P]: particle radius / median radius (lognormal factors);
    axes, m0 [P,3]; seeds [P]; Ms, T; field_frequency; time_step; end_time;
    max_samples: the sample grid of lib/simulation.cpp:342-406 (in SI units,
    fixed for all theta).
    """

    def __init__(self, radius_shape, axes, m0, seeds, magnetisation, temperature,
                 field_frequency, time_step, end_time, max_samples, checkpoint_interval=1000,
                 precision="fp32", device=0):
        k, n_steps = core.sample_schedule(time_step, end_time, max_samples)
        self.sample_steps = k
        self.n_steps = int(n_steps)
        self.time = torch.from_numpy(np.arange(max_samples) * (end_time / (max_samples - 1)))
        self.field_frequency = float(field_frequency)
        self.magnetisation = float(magnetisation)
        self.runner = core.Adjoint(radius_shape, axes, m0, seeds, magnetisation, temperature,
                                   field_frequency, time_step, self.n_steps, k,
                                   checkpoint_interval, precision, 0, device)

    def __call__(self, theta):
        """theta: float64 tensor [4] (requires_grad for gradients) ->
        dict(sar [W/m^3], energy [J/m^3], final_mz, ens_mz [S], field [S])."""
        ens, fin = EnsembleTrajectory.apply(theta, self.runner)
        field = theta[2] * torch.sin(2 * np.pi * self.field_frequency * self.time)
        energy = hysteresis_energy(ens, field, self.time, self.field_frequency)
        return {"sar": self.field_frequency * energy, "energy": energy, "final_mz": fin,
                "ens_mz": ens, "field": field}
