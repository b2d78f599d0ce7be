"""torch.autograd.Function over the C ABI's adjoint (a thin harness).

``ConductanceScalingLoss.apply(scalings, problem)`` returns the voltage-
mismatch loss of a monodomain run whose per-node cell parameters are
``scalings * default`` for the parameters named in ``problem.param_ids``;
its backward is the gradient ``cmx_adjoint`` computes on the device.  All
arithmetic runs in libcmx.so; torch only carries the tensors.
"""
import numpy as np
import torch

from . import cmx


def smooth_random_field(shape, h, ell, rng, lo=0.6, hi=1.4):
    """Seeded Gaussian random field on a grid of spacing h (cm): white noise
    filtered with a Gaussian of correlation length ell (cm), rescaled to
    [lo, hi].  Ground-truth conductance scalings for inverse problems."""
    noise = rng.standard_normal(shape)
    ky = np.fft.fftfreq(shape[0], d=h)
    kx = np.fft.fftfreq(shape[1], d=h)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.exp(-2.0 * (np.pi * ell) ** 2 * k2)))
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


class AdjointProblem:
    """What a loss evaluation needs besides the scalings: the context, the
    state it starts from, the run length and the observations."""

    def __init__(self, mono, initial_state, t0, n_steps, obs_nodes, obs_every, obs_values, param_ids,
                 checkpoint_every=100):
        self.mono = mono
        self.initial_state = np.array(initial_state, dtype=np.float64)
        self.t0, self.n_steps = float(t0), int(n_steps)
        self.obs_nodes = np.asarray(obs_nodes, dtype=np.int32)
        self.obs_every = int(obs_every)
        self.obs_values = np.asarray(obs_values, dtype=np.float64)
        self.param_ids = list(param_ids)
        self.checkpoint_every = int(checkpoint_every)
        self.defaults = np.array([d for _, d in cmx.param_info(mono.model)])

    def params_for(self, scalings):
        """[n_params, n] parameter planes: defaults, with the listed ones scaled."""
        p = np.repeat(self.defaults[:, None], self.mono.n, axis=1)
        for k, q in enumerate(self.param_ids):
            p[q] = scalings[k] * self.defaults[q]
        return p

    def evaluate(self, scalings):
        """(loss, dL/dscalings [len(param_ids), n]) at the given scalings."""
        self.mono.set_state(self.initial_state)
        self.mono.set_params(self.params_for(scalings))
        loss, gp, _ = self.mono.adjoint(self.t0, self.n_steps, self.obs_nodes, self.obs_every, self.obs_values,
                                        self.checkpoint_every)
        g = np.stack([gp[q] * self.defaults[q] for q in self.param_ids])
        return loss, g


class ConductanceScalingLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scalings, problem):
        s = scalings.detach().cpu().numpy().astype(np.float64)
        loss, g = problem.evaluate(s)
        ctx.save_for_backward(torch.from_numpy(g).to(scalings.device, scalings.dtype))
        return torch.tensor(loss, dtype=scalings.dtype, device=scalings.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None
