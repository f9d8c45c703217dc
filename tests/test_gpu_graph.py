"""hipGraph capture of rollouts (torch.cuda.graph): the ragged INC training
step on the one-call path (inc_ragged_inc_fwd / _adj: the network as ragged
convolutions, the ragged solver step, the elementwise cotangent of the masked
MSE and the adjoint), the KS ETDRK4 rollout with its adjoint, and the KS
INC training step on its one-call path (inc_ks_inc_fwd / _adj).

Inside a captured region the library makes no stream operation other than
kernel launches: its occupancy and device-attribute queries are answered
from caches filled by the eager warm-up, workspaces come from torch's
allocator (the graph's pool), zeroing is done by kernels, nothing
synchronises.  The masked-MSE cotangent is elementwise; the loss value, a
reduction, stays outside the graph.  Each graph is replayed several times
with eager work and synchronisations between the replays, and every replay
must reproduce the eager result exactly (same kernels, same launch shapes,
fixed summation order).
"""
import math

import pytest
import torch

import inc_amd
from inc_amd import KSSimParams, KSSolverTorch, RaggedBurgers
from inc_amd.batched import stable_dt
from inc_amd.correction import CorrectionCNN

pytestmark = pytest.mark.gpu


def _field(n, seed):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(n, dtype=torch.float64)
    for k in range(1, 9):
        u += torch.randn(1, generator=gen, dtype=torch.float64) / k * torch.sin(
            k * x + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
    return u / u.abs().max()


def _capture(fn):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = fn()
    return graph, out


@pytest.mark.parametrize("keep", [False, True, "auto"])
def test_ragged_inc_training_step_graph_replays_match_eager(cuda, keep):
    # keep: the adjoint reads the forward's stored activations (True; "auto"
    # decides inside the capture) or evaluates the network again (False)
    count, steps = 40, 12
    gen = torch.Generator().manual_seed(21)
    choices = [64, 128, 256, 512, 1024]
    ns = [choices[int(i)] for i in torch.randint(0, 5, (count,), generator=gen)]
    nus = torch.exp(torch.empty(count, dtype=torch.float64).uniform_(math.log(0.003), math.log(0.03),
                                                                      generator=gen)).tolist()
    dts = [stable_dt(n, nu) for n, nu in zip(ns, nus)]
    hz = torch.randint(3, steps + 1, (count,), generator=gen).tolist()
    rb = RaggedBurgers(ns, nus, dts, horizon=hz, team=128, keep_activations=keep)
    u0 = rb.pack([_field(n, 500 + b) for b, n in enumerate(ns)])
    cnn = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(3), scale=0.01).cuda()
    params = list(cnn.parameters())
    with torch.no_grad():
        target = RaggedBurgers(ns, [1.2 * v for v in nus], dts, horizon=hz, team=128).rollout(u0, steps)
    mask = rb.valid_mask(steps)
    msum = float(mask.sum())

    def train_step():
        traj = rb.rollout(u0, steps, model=cnn)
        torch.autograd.backward(traj, (2.0 / msum) * (traj - target) * mask)
        return traj

    for p in params:
        p.grad = None
    traj_e = train_step().detach().clone()
    g_e = [p.grad.clone() for p in params]
    for p in params:
        p.grad = None
    graph, traj_g = _capture(train_step)
    for rep in range(3):
        for p in params:
            p.grad.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(traj_g, traj_e), rep
        for a, b in zip(params, g_e):
            assert torch.equal(a.grad, b), rep
        # eager work between replays (the graph's buffers must not be disturbed)
        with torch.no_grad():
            _ = rb.rollout(u0 * 0.5, steps, model=cnn).sum().item()


def test_ks_rollout_and_adjoint_graph_replays_match_eager(cuda):
    b, n, h, Lk = 64, 256, 50, 64.0
    sv = KSSolverTorch(KSSimParams([Lk], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    Ld = torch.tensor([Lk], dtype=torch.float64, device=cuda)
    gen = torch.Generator().manual_seed(4)
    x = torch.arange(n, dtype=torch.float64) * (Lk / n)
    u0 = sum(0.5 * torch.randn(b, 1, generator=gen, dtype=torch.float64) * torch.cos(2 * math.pi * k * x / Lk)
             for k in range(1, 11)).float().to(cuda)
    cot = (1e-3 * torch.randn(b, h, n, generator=gen)).to(cuda)
    u = u0.clone().requires_grad_(True)

    def step():
        u.grad = None
        traj, _ = inc_amd.simulate_rollout(None, sv, u, h, None, L=Ld)
        torch.autograd.backward(traj, cot)
        return traj

    traj_e = step().detach().clone()
    g_e = u.grad.clone()
    graph, traj_g = _capture(step)
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(traj_g, traj_e), rep
        assert torch.equal(u.grad, g_e), rep
        with torch.no_grad():
            _ = sv.rollout(u0 * 0.5, 5, Ld).sum().item()


@pytest.mark.parametrize("keep", [False, True])
def test_ks_inc_training_step_graph_replays_match_eager(cuda, keep):
    # the KS INC one-call path (inc_ks_inc_fwd / _adj) with the network's
    # parameter gradients; keep: the forward's stored activations in the
    # graph's pool (True) or the network evaluated again in the adjoint
    b, n, h, Lk = 16, 128, 8, 64.0
    sv = KSSolverTorch(KSSimParams([Lk], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    sv.keep_activations = keep
    Ld = torch.tensor([Lk], dtype=torch.float64, device=cuda)
    gen = torch.Generator().manual_seed(8)
    x = torch.arange(n, dtype=torch.float64) * (Lk / n)
    u0 = sum(0.5 * torch.randn(b, 1, generator=gen, dtype=torch.float64) * torch.cos(2 * math.pi * k * x / Lk)
             for k in range(1, 9)).float().to(cuda)
    target = (0.9 * u0[:, None, :]).expand(b, h, n).contiguous()
    net = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(9), scale=0.01).cuda()
    params = list(net.parameters())

    def train_step():
        traj, _ = inc_amd.simulate_rollout(net, sv, u0, h, "INC", L=Ld)
        torch.autograd.backward(traj, (2.0 / traj.numel()) * (traj - target))
        return traj

    for p in params:
        p.grad = None
    traj_e = train_step().detach().clone()
    g_e = [p.grad.clone() for p in params]
    for p in params:
        p.grad = None
    graph, traj_g = _capture(train_step)
    for rep in range(3):
        for p in params:
            p.grad.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(traj_g, traj_e), rep
        for a, g in zip(params, g_e):
            assert torch.equal(a.grad, g), rep
        with torch.no_grad():
            _ = sv.rollout(u0 * 0.5, 3, Ld).sum().item()
