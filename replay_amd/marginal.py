"""Differentiable marginal-probability SEIR over a replay context.

The deterministic individual-level form of the replay (the C ABI's
replay_marginal_*): per node i and parameter set b, probabilities of S, E, I,
R advance over the same duration-weighted contact slices as the stochastic
step,

    lam   = 1 - exp(-beta_b * phi_i * E_i / 3600),  E_i = sum_j A_k[i,j] iota_j
    s'    = s (1 - lam)
    e'    = e (1 - a_b) + s lam,        a_b = 1 - exp(-sigma_b * dt)
    iota' = iota (1 - c_b) + e a_b,     c_b = 1 - exp(-gamma_b * dt)
    r'    = r + iota c_b,               dt = step in hours

from s = 1 - p0, iota = p0.  The curves are the expected S, E, I, R counts per
iteration and parameter set.  beta is per hour of contact (a stochastic
infect_prob p corresponds to beta = -log(1 - p)); sigma and gamma are rates
per hour (1 / mean exposed and infectious durations).

MarginalSEIR is a torch.autograd.Function whose forward and backward run the
HIP kernels of libreplay_hip (fp32, deterministic): the backward is the
hand-written adjoint (transposed-slice product plus the transition adjoints)
with checkpoints every `checkpoint_every` iterations recomputed on the fly.
PyTorch only carries the device buffers and the stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import ReplayError, Simulation, load_library


def _dev_f32(t: torch.Tensor, shape, name: str) -> torch.Tensor:
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a device tensor")
    t = t.detach().to(torch.float32).contiguous()
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


class MarginalEngine:
    """Buffers and checkpoints of the marginal form for one replay context."""

    def __init__(self, sim: Simulation, n_sets: int, checkpoint_every: int = 32):
        L = load_library()
        self.sim = sim
        info = sim.info()
        self.n_nodes = info["n_nodes"]
        self.iterations = info["iterations"]
        self.n_sets = n_sets
        h = C.c_void_p()
        rc = L.replay_marginal_create(sim._ctx, n_sets, checkpoint_every, C.byref(h))
        sim._check(rc)
        self._h = h

    def close(self) -> None:
        if self._h:
            load_library().replay_marginal_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise ReplayError(rc, load_library().replay_last_error(self.sim._ctx).decode())

    def forward(self, beta, sigma, gamma, phi) -> torch.Tensor:
        B, N, T = self.n_sets, self.n_nodes, self.iterations
        beta = _dev_f32(beta, (B,), "beta")
        sigma = _dev_f32(sigma, (B,), "sigma")
        gamma = _dev_f32(gamma, (B,), "gamma")
        phi = _dev_f32(phi, (N,), "phi")
        curves = torch.empty((T + 1, B, 4), dtype=torch.float32, device=beta.device)
        stream = torch.cuda.current_stream(beta.device).cuda_stream
        self._check(load_library().replay_marginal_forward(
            self._h, beta.data_ptr(), sigma.data_ptr(), gamma.data_ptr(), phi.data_ptr(),
            curves.data_ptr(), stream))
        return curves

    def backward(self, grad_curves: torch.Tensor):
        B, N, T = self.n_sets, self.n_nodes, self.iterations
        g = _dev_f32(grad_curves, (T + 1, B, 4), "grad_curves")
        dev = g.device
        gb = torch.empty(B, dtype=torch.float32, device=dev)
        gs = torch.empty(B, dtype=torch.float32, device=dev)
        gg = torch.empty(B, dtype=torch.float32, device=dev)
        gp = torch.empty(N, dtype=torch.float32, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        self._check(load_library().replay_marginal_backward(
            self._h, g.data_ptr(), gb.data_ptr(), gs.data_ptr(), gg.data_ptr(), gp.data_ptr(),
            stream))
        return gb, gs, gg, gp


class MarginalSEIR(torch.autograd.Function):
    """curves = MarginalSEIR.apply(engine, beta, sigma, gamma, phi)."""

    @staticmethod
    def forward(ctx, engine: MarginalEngine, beta, sigma, gamma, phi):
        ctx.engine = engine
        return engine.forward(beta, sigma, gamma, phi)

    @staticmethod
    def backward(ctx, grad_curves):
        gb, gs, gg, gp = ctx.engine.backward(grad_curves)
        return None, gb, gs, gg, gp


def marginal_seir(engine: MarginalEngine, beta, sigma, gamma, phi) -> torch.Tensor:
    """Expected (S, E, I, R) counts, shape (iterations + 1, n_sets, 4)."""
    return MarginalSEIR.apply(engine, beta, sigma, gamma, phi)
