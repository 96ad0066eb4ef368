"""HIP forward and hand-written adjoint of the marginal-probability SEIR form
against the FP64 torch reference (oracle/marginal_ref.py), through the C ABI
and the torch.autograd.Function.  Needs an MI355X.

Tolerances (fp32 kernels against fp64): curves within 1e-5 of the node
count; gradients within 1e-4 relative (norm over each parameter group)."""
import numpy as np
import pytest
import torch

from oracle import marginal_ref
from replay_amd import Network, Params, Simulation
from replay_amd.marginal import MarginalEngine, marginal_seir

from test_gpu_parity import _random_case, _to_oracle_params

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _case(oracle, seed, n=48, T=36, step=1200.0, D=1200):
    n, a, b, s, e = _random_case(seed, n=n, e=n * 10, span=T * int(step), long_frac=0.3)
    p = Params(M=64, step_size=step, static_network_duration=D, iterations=T,
               initial_infected=0.08, seed=seed)
    sim = Simulation.from_contacts(n, a, b, s, e, p)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    A = marginal_ref.dense_slices(g, _to_oracle_params(oracle, p), T)
    return sim, A, p


def _params(B, N, seed):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.3, 2.0, B)
    sigma = rng.uniform(0.05, 0.5, B)
    gamma = rng.uniform(0.02, 0.3, B)
    phi = rng.uniform(0.5, 1.5, N)
    return [torch.tensor(x, dtype=torch.float64) for x in (beta, sigma, gamma, phi)]


def _ref_grads(A, params, p0, dt_h, W):
    xs = [t.clone().requires_grad_(True) for t in params]
    C = marginal_ref.simulate(A, *xs, p0, dt_h)
    L = (C * W).sum()
    return C.detach(), torch.autograd.grad(L, xs)


def _rel(a, b):
    return float(torch.linalg.norm(a - b) / max(float(torch.linalg.norm(b)), 1e-30))


@pytest.mark.parametrize("seed,B", [(0, 5), (1, 64), (2, 70)])
def test_forward_and_adjoint_match_fp64(oracle, seed, B):
    sim, A, p = _case(oracle, seed)
    N, T = A.shape[1], A.shape[0]
    params = _params(B, N, seed)
    W = torch.tensor(np.random.default_rng(seed + 7).normal(size=(T + 1, B, 4)))
    C_ref, g_ref = _ref_grads(A, params, p.initial_infected, p.step_size / 3600.0, W)
    with sim:
        eng = MarginalEngine(sim, B, checkpoint_every=8)
        dev = [t.to(torch.float32).to(DEV) for t in params]
        C = eng.forward(*dev)
        g = eng.backward(W.to(torch.float32).to(DEV))
        torch.cuda.synchronize()
        eng.close()
    C = C.cpu().double()
    assert float((C - C_ref).abs().max()) <= 1e-5 * N
    for name, got, ref in zip(("beta", "sigma", "gamma", "phi"), g, g_ref):
        assert _rel(got.cpu().double(), ref) <= 1e-4, name


def test_checkpoint_interval_does_not_change_results(oracle):
    sim, A, p = _case(oracle, 3, T=30)
    N, T = A.shape[1], A.shape[0]
    B = 9
    params = [t.to(torch.float32).to(DEV) for t in _params(B, N, 3)]
    W = torch.randn(T + 1, B, 4, generator=torch.Generator().manual_seed(1)).to(DEV)
    outs = []
    with sim:
        for every in (1, 7, 30, 1000):
            eng = MarginalEngine(sim, B, checkpoint_every=every)
            C = eng.forward(*params)
            g = eng.backward(W)
            torch.cuda.synchronize()
            outs.append((C.cpu(), [x.cpu() for x in g]))
            eng.close()
    for C, g in outs[1:]:
        assert torch.equal(C, outs[0][0])
        for x, y in zip(g, outs[0][1]):
            assert torch.equal(x, y)


def test_autograd_function_mse_loss(oracle):
    sim, A, p = _case(oracle, 4, T=24)
    N, T = A.shape[1], A.shape[0]
    B = 6
    base = _params(B, N, 4)
    target = marginal_ref.simulate(A, *_params(1, N, 40)[:3], base[3], p.initial_infected,
                                   p.step_size / 3600.0)[:, 0, 2].to(torch.float32)
    with sim:
        eng = MarginalEngine(sim, B, checkpoint_every=5)
        xs = [t.to(torch.float32).to(DEV).requires_grad_(True) for t in base]
        C = marginal_seir(eng, *xs)
        loss = ((C[:, :, 2] - target.to(DEV)[:, None]) ** 2).mean()
        loss.backward()
        torch.cuda.synchronize()
        eng.close()
    # the same gradients from the FP64 reference
    rs = [t.clone().requires_grad_(True) for t in base]
    Cr = marginal_ref.simulate(A, *rs, p.initial_infected, p.step_size / 3600.0)
    lr = ((Cr[:, :, 2] - target.double()[:, None]) ** 2).mean()
    lr.backward()
    assert abs(loss.item() - lr.item()) <= 1e-5 * max(1.0, lr.item())
    for x, r in zip(xs, rs):
        assert _rel(x.grad.cpu().double(), r.grad) <= 1e-4


def test_batched_context_per_network_seeds(oracle):
    nets, As, p0s = [], [], []
    for k, n in enumerate((30, 45)):
        _, a, b, s, e = _random_case(50 + k, n=n, e=n * 8, span=20 * 1200, long_frac=0.3)
        nets.append(Network(n_nodes=n, src=a, tgt=b, start=s, end=e,
                            initial_infected=0.05 * (k + 1), infect_prob=0.2,
                            exposed_duration=3600, infectious_duration=7200,
                            resistant_duration=0, stream=k))
        g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
        pk = Params(M=64, step_size=1200.0, static_network_duration=1200, iterations=20,
                    initial_infected=0.05 * (k + 1))
        As.append(marginal_ref.dense_slices(g, _to_oracle_params(oracle, pk), 20))
        p0s.append(torch.full((n,), 0.05 * (k + 1), dtype=torch.float64))
    T, N = 20, 75
    A = torch.zeros(T, N, N, dtype=torch.float64)
    A[:, :30, :30] = As[0]
    A[:, 30:, 30:] = As[1]
    p0 = torch.cat(p0s)
    shared = Params(M=64, step_size=1200.0, static_network_duration=1200, iterations=T)
    B = 4
    params = _params(B, N, 8)
    W = torch.tensor(np.random.default_rng(3).normal(size=(T + 1, B, 4)))
    C_ref, g_ref = _ref_grads(A, params, p0, 1200 / 3600.0, W)
    with Simulation.from_networks(nets, shared) as sim:
        eng = MarginalEngine(sim, B, checkpoint_every=6)
        C = eng.forward(*[t.to(torch.float32).to(DEV) for t in params])
        g = eng.backward(W.to(torch.float32).to(DEV))
        torch.cuda.synchronize()
        eng.close()
    assert float((C.cpu().double() - C_ref).abs().max()) <= 1e-5 * N
    for got, ref in zip(g, g_ref):
        assert _rel(got.cpu().double(), ref) <= 1e-4


def test_calibration_steps_reduce_loss(oracle):
    """Gradient descent on log-parameters fits a target curve made at other
    parameters (the calibration use of the adjoint)."""
    sim, A, p = _case(oracle, 6, n=60, T=40)
    N = A.shape[1]
    B = 4
    phi = torch.ones(N, device=DEV)
    with sim:
        one = MarginalEngine(sim, 1, checkpoint_every=10)
        truth = [torch.tensor([v], dtype=torch.float32, device=DEV) for v in (1.2, 0.3, 0.1)]
        target = one.forward(*truth, phi)[:, 0, 2].clone()
        eng = MarginalEngine(sim, B, checkpoint_every=10)
        logs = [torch.log(torch.tensor(v, dtype=torch.float32, device=DEV)).requires_grad_(True)
                for v in ([0.4, 0.6, 2.5, 3.0], [0.15, 0.6, 0.2, 0.5], [0.2, 0.05, 0.3, 0.15])]
        opt = torch.optim.Adam(logs, lr=0.05)
        losses = []
        for _ in range(60):
            opt.zero_grad()
            C = marginal_seir(eng, *[torch.exp(x) for x in logs], phi)
            loss = ((C[:, :, 2] - target[:, None]) ** 2).mean(0)  # per set
            loss.sum().backward()
            opt.step()
            losses.append(loss.detach().cpu())
        torch.cuda.synchronize()
        eng.close()
        one.close()
    assert torch.all(losses[-1] < 0.2 * losses[0])


def test_config4_full_size_directional_derivatives():
    """Full workload size: the adjoint's directional derivatives against
    central differences of the HIP forward (fp32, step 1e-3 relative)."""
    from replay_amd import workloads
    w = workloads.config4_marginal()
    g = w.graph
    with Simulation.from_contacts(g.n_nodes, g.src, g.tgt, g.start, g.end, g.params) as sim:
        N = sim.info()["n_nodes"]
        B = int(w.beta.size)
        eng = MarginalEngine(sim, B, checkpoint_every=32)
        one = MarginalEngine(sim, 1, checkpoint_every=32)
        phi = torch.ones(N, device=DEV)
        truth = [torch.tensor([v], dtype=torch.float32, device=DEV) for v in w.truth]
        target = one.forward(*truth, phi)[:, 0, 2].double()
        base = [torch.tensor(v, device=DEV) for v in (w.beta, w.sigma, w.gamma)] + [phi]

        def loss_of(xs):
            C = eng.forward(*xs)
            return float(((C[:, :, 2].double() - target[:, None]) ** 2).mean())

        xs = [x.clone().requires_grad_(True) for x in base]
        C = marginal_seir(eng, *xs)
        L = ((C[:, :, 2] - target.float()[:, None]) ** 2).mean()
        L.backward()
        grads = [x.grad.double() for x in xs]
        gen = torch.Generator(device="cpu").manual_seed(5)
        for k in range(4):
            d = torch.randn(base[k].shape, generator=gen).to(DEV)
            h = 1e-3
            plus = [x.clone() for x in base]
            minus = [x.clone() for x in base]
            plus[k] = base[k] * (1 + h * d)
            minus[k] = base[k] * (1 - h * d)
            fd = (loss_of(plus) - loss_of(minus)) / (2 * h)
            ad = float((grads[k] * base[k].double() * d.double()).sum())
            assert abs(fd - ad) <= 2e-2 * abs(ad) + 1e-9 * abs(loss_of(base)), (k, fd, ad)
        torch.cuda.synchronize()
        eng.close()
        one.close()


def test_config4_full_size_gradients_match_fp64(oracle):
    """The 20k-node block network of the marginal workload over its first 64
    iterations (20-minute steps), 64 parameter sets: curves within 1e-5 of
    the node count and the adjoint's gradients (beta, sigma, gamma and the
    20,000 susceptibilities) within 1e-4 relative of FP64 autograd over the
    oracle's sparse slices."""
    from dataclasses import replace
    from replay_amd import workloads
    w = workloads.config4_marginal()
    g = w.graph
    T = 64
    p = replace(g.params, iterations=T)
    og = oracle.OracleGraph.from_contacts(g.n_nodes, g.src, g.tgt, g.start, g.end)
    slices = marginal_ref.sparse_slices(og, _to_oracle_params(oracle, p), T)
    N, B = g.n_nodes, int(w.beta.size)
    rng = np.random.default_rng(11)
    params = [torch.tensor(x, dtype=torch.float64) for x in
              (w.beta, w.sigma, w.gamma, rng.uniform(0.5, 1.5, N))]
    W = torch.tensor(rng.normal(size=(T + 1, B, 4)))
    xs = [t.clone().requires_grad_(True) for t in params]
    C_ref = marginal_ref.simulate_sparse(slices, *xs, p.initial_infected, p.step_size / 3600.0)
    g_ref = torch.autograd.grad((C_ref * W).sum(), xs)
    print(f"fp64 reference done: N={N} B={B} T={T} nnz(k=0)={slices[0]._nnz()}", flush=True)
    with Simulation.from_contacts(g.n_nodes, g.src, g.tgt, g.start, g.end, p) as sim:
        eng = MarginalEngine(sim, B, checkpoint_every=32)
        C = eng.forward(*[x.to(torch.float32).to(DEV) for x in params])
        grads = eng.backward(W.to(torch.float32).to(DEV))
        torch.cuda.synchronize()
        eng.close()
    assert float((C.cpu().double() - C_ref.detach()).abs().max()) <= 1e-5 * N
    for name, got, ref in zip(("beta", "sigma", "gamma", "phi"), grads, g_ref):
        rel = _rel(got.cpu().double(), ref)
        assert rel <= 1e-4, (name, rel)


def test_marginal_error_paths(oracle):
    import ctypes as C
    from replay_amd import ReplayError, load_library
    sim, A, p = _case(oracle, 9, T=8)
    L = load_library()
    with sim:
        h = C.c_void_p()
        assert L.replay_marginal_create(sim._ctx, 0, 4, C.byref(h)) != 0       # no sets
        assert L.replay_marginal_create(sim._ctx, 3, 0, C.byref(h)) != 0       # no interval
        assert L.replay_marginal_create(None, 3, 4, C.byref(h)) != 0           # no context
        eng = MarginalEngine(sim, 3, checkpoint_every=4)
        g = torch.zeros(9, 3, 4, device=DEV)
        with pytest.raises(ReplayError):                                        # before forward
            eng.backward(g)
        with pytest.raises(ValueError):
            eng.forward(torch.ones(2, device=DEV), torch.ones(3, device=DEV),
                        torch.ones(3, device=DEV), torch.ones(A.shape[1], device=DEV))
        eng.close()
