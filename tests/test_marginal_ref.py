"""Pins the FP64 reference of the marginal-probability SEIR form (CPU).

The reference is the checker of the HIP forward and adjoint, so it is pinned
here by what the mathematics fixes: closed forms with no transmission,
conservation of probability mass, a hand-computed two-node step, and central
finite differences of its own autograd gradients."""
import math

import numpy as np
import pytest
import torch

from oracle import marginal_ref


def _params(B, seed=0):
    g = torch.Generator().manual_seed(seed)
    beta = (0.5 + torch.rand(B, generator=g, dtype=torch.float64)) * 0.8
    sigma = (0.5 + torch.rand(B, generator=g, dtype=torch.float64)) / 2.0
    gamma = (0.5 + torch.rand(B, generator=g, dtype=torch.float64)) / 5.0
    return beta, sigma, gamma


def _random_A(T, N, seed=1, density=0.2):
    rng = np.random.default_rng(seed)
    A = np.zeros((T, N, N))
    for k in range(T):
        m = rng.uniform(size=(N, N)) < density
        w = rng.integers(60, 3600, size=(N, N)) * m
        w = np.triu(w, 1)
        A[k] = w + w.T
    return torch.from_numpy(A.astype(np.float64))


def test_no_transmission_closed_form():
    T, N, B = 12, 7, 3
    A = torch.zeros(T, N, N, dtype=torch.float64)
    beta, sigma, gamma = _params(B)
    p0, dt = 0.2, 0.5
    C = marginal_ref.simulate(A, beta, sigma, gamma, torch.ones(N, dtype=torch.float64), p0, dt)
    k = torch.arange(T + 1, dtype=torch.float64)[:, None]
    keep = torch.exp(-gamma * dt)[None, :] ** k  # (1 - c)^k
    torch.testing.assert_close(C[:, :, 2], N * p0 * keep, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(C[:, :, 3], N * p0 * (1 - keep), rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(C[:, :, 0], torch.full_like(C[:, :, 0], N * (1 - p0)))
    assert torch.all(C[:, :, 1] == 0)


def test_mass_conserved():
    T, N, B = 15, 9, 4
    A = _random_A(T, N)
    beta, sigma, gamma = _params(B, 3)
    phi = torch.linspace(0.5, 1.5, N, dtype=torch.float64)
    C = marginal_ref.simulate(A, beta, sigma, gamma, phi, 0.1, 1 / 3)
    torch.testing.assert_close(C.sum(-1), torch.full((T + 1, B), float(N), dtype=torch.float64),
                               rtol=1e-12, atol=1e-10)


def test_two_node_step_by_hand():
    # node 0 infectious with certainty, node 1 susceptible, one contact of 1800 s
    A = torch.tensor([[[0.0, 1800.0], [1800.0, 0.0]]], dtype=torch.float64)
    beta = torch.tensor([0.6], dtype=torch.float64)
    sigma = torch.tensor([0.25], dtype=torch.float64)
    gamma = torch.tensor([0.1], dtype=torch.float64)
    phi = torch.tensor([1.0, 2.0], dtype=torch.float64)
    p0 = torch.tensor([1.0, 0.0], dtype=torch.float64)
    C = marginal_ref.simulate(A, beta, sigma, gamma, phi, p0, 1.0)
    lam = 1 - math.exp(-0.6 * 2.0 * 1800 / 3600)
    c = 1 - math.exp(-0.1)
    expect = [1 - lam, lam, 1 - c, c]  # (S, E, I, R) summed over the two nodes
    np.testing.assert_allclose(C[1, 0].numpy(), expect, rtol=1e-14)


@pytest.mark.parametrize("seed", [0, 1])
def test_autograd_matches_central_differences(seed):
    T, N, B = 10, 8, 3
    A = _random_A(T, N, seed=seed + 5)
    beta, sigma, gamma = _params(B, seed)
    phi = torch.linspace(0.7, 1.3, N, dtype=torch.float64)
    target = torch.rand(T + 1, dtype=torch.float64, generator=torch.Generator().manual_seed(9))

    def loss(b, s, g, f):
        C = marginal_ref.simulate(A, b, s, g, f, 0.15, 0.5)
        return ((C[:, :, 2] - target[:, None]) ** 2).mean() + (C[:, :, 1] ** 2).mean()

    xs = [t.clone().requires_grad_(True) for t in (beta, sigma, gamma, phi)]
    L = loss(*xs)
    grads = torch.autograd.grad(L, xs)
    h = 1e-6
    for which in range(4):
        for idx in range(min(3, xs[which].numel())):
            plus = [t.detach().clone() for t in xs]
            minus = [t.detach().clone() for t in xs]
            plus[which][idx] += h
            minus[which][idx] -= h
            fd = (loss(*plus) - loss(*minus)).item() / (2 * h)
            assert abs(fd - grads[which][idx].item()) <= 1e-6 * max(1.0, abs(fd)), (which, idx)


def test_sparse_reference_equals_dense(oracle):
    """The sparse-slice FP64 reference (used at full size) gives the dense
    one's curves and autograd gradients on random graphs, to fp64 rounding."""
    from test_gpu_parity import _random_case
    for seed in (3, 4):
        n, a, b, s, e = _random_case(seed, n=40, e=600, span=20 * 1800, long_frac=0.4)
        g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
        p = oracle.make_params(M=1, step_size=1800.0, iterations=20, static_network_duration=3600)
        dense = marginal_ref.dense_slices(g, p, 20)
        sparse = marginal_ref.sparse_slices(g, p, 20)
        rng = np.random.default_rng(seed)
        xs = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in
              (rng.uniform(0.5, 2, 5), rng.uniform(0.1, 0.5, 5), rng.uniform(0.05, 0.3, 5),
               rng.uniform(0.5, 1.5, n))]
        W = torch.tensor(rng.normal(size=(21, 5, 4)))
        Cd = marginal_ref.simulate(dense, *xs, 0.1, 0.5)
        gd = torch.autograd.grad((Cd * W).sum(), xs)
        Cs = marginal_ref.simulate_sparse(sparse, *xs, 0.1, 0.5)
        gs = torch.autograd.grad((Cs * W).sum(), xs)
        assert float((Cd - Cs).abs().max()) < 1e-10
        for x, y in zip(gd, gs):
            assert float((x - y).abs().max()) <= 1e-10 * max(1.0, float(x.abs().max()))
