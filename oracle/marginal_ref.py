"""FP64 reference of the marginal-probability SEIR form (TEST INFRASTRUCTURE).

Only tests/ use this module: it is the checker for the HIP forward and
adjoint kernels of replay_amd.marginal, never part of the product.  The
contact weights of each iteration come from the oracle's restatement of the
reference CSR build (oracle_py.OracleGraph.step_csr, which follows
tool/gpu_cpu_temporal_sim.cu:238-350 and the window arithmetic of
cu:860-959), so both sides use the same duration-weighted slices; the
recurrence itself is the model documented in replay_amd/marginal.py and
written here in plain torch float64 so that autograd gives the exact
gradients (up to fp64 rounding) and central differences can pin them.
"""
from __future__ import annotations

import numpy as np
import torch


def dense_slices(graph, params, iterations: int) -> torch.Tensor:
    """A[k] = duration-weighted adjacency of iteration k, shape (T, N, N)."""
    N = graph.n_nodes
    A = np.zeros((iterations, N, N), np.float64)
    for k in range(iterations):
        rowptr, col, val, _ = graph.step_csr(params, k)
        for i in range(N):
            for q in range(rowptr[i], rowptr[i + 1]):
                A[k, i, col[q]] += float(val[q])
    return torch.from_numpy(A)


def simulate(A: torch.Tensor, beta, sigma, gamma, phi, p0, dt_h: float) -> torch.Tensor:
    """Expected (S, E, I, R) counts per iteration and parameter set.

    A: (T, N, N); beta, sigma, gamma: (B,); phi: (N,); p0: scalar or (N,).
    Returns (T + 1, B, 4) in the dtype of the parameters.
    """
    T, N, _ = A.shape
    B = beta.shape[0]
    dt = beta.dtype
    p0 = torch.as_tensor(p0, dtype=dt).expand(N)[:, None].expand(N, B)
    s = 1.0 - p0
    e = torch.zeros(N, B, dtype=dt)
    io = p0.clone()
    r = torch.zeros(N, B, dtype=dt)
    a = -torch.expm1(-sigma * dt_h)
    c = -torch.expm1(-gamma * dt_h)
    kappa = (beta / 3600.0)[None, :] * phi[:, None]
    rows = [torch.stack([s.sum(0), e.sum(0), io.sum(0), r.sum(0)], dim=1)]
    for k in range(T):
        E = A[k].to(dt) @ io
        lam = -torch.expm1(-kappa * E)
        s, e, io, r = s * (1.0 - lam), e * (1.0 - a) + s * lam, io * (1.0 - c) + e * a, r + io * c
        rows.append(torch.stack([s.sum(0), e.sum(0), io.sum(0), r.sum(0)], dim=1))
    return torch.stack(rows, dim=0)


def sparse_slices(graph, params, iterations: int) -> list:
    """The same slices as sparse (N, N) float64 COO tensors, one per
    iteration, for graphs too large for dense_slices (duplicate CSR entries
    are summed, as dense_slices sums them)."""
    N = graph.n_nodes
    out = []
    for k in range(iterations):
        rowptr, col, val, _ = graph.step_csr(params, k)
        rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(rowptr).astype(np.int64))
        idx = torch.from_numpy(np.stack([rows, col.astype(np.int64)]))
        A = torch.sparse_coo_tensor(idx, torch.from_numpy(val.astype(np.float64)), (N, N))
        out.append(A.coalesce())
    return out


def simulate_sparse(slices: list, beta, sigma, gamma, phi, p0, dt_h: float) -> torch.Tensor:
    """simulate() with the exposure product taken over sparse slices
    (torch.sparse.mm, which carries autograd to the dense operand)."""
    N = slices[0].shape[0] if slices else phi.shape[0]
    B = beta.shape[0]
    dt = beta.dtype
    p0 = torch.as_tensor(p0, dtype=dt).expand(N)[:, None].expand(N, B)
    s = 1.0 - p0
    e = torch.zeros(N, B, dtype=dt)
    io = p0.clone()
    r = torch.zeros(N, B, dtype=dt)
    a = -torch.expm1(-sigma * dt_h)
    c = -torch.expm1(-gamma * dt_h)
    kappa = (beta / 3600.0)[None, :] * phi[:, None]
    rows = [torch.stack([s.sum(0), e.sum(0), io.sum(0), r.sum(0)], dim=1)]
    for A in slices:
        E = torch.sparse.mm(A.to(dt), io)
        lam = -torch.expm1(-kappa * E)
        s, e, io, r = s * (1.0 - lam), e * (1.0 - a) + s * lam, io * (1.0 - c) + e * a, r + io * c
        rows.append(torch.stack([s.sum(0), e.sum(0), io.sum(0), r.sum(0)], dim=1))
    return torch.stack(rows, dim=0)
