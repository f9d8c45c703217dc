"""Checkpoint-and-recompute in the fused rollouts (``checkpoint_every=k``):
the forward keeps no stage states, and the backward recomputes them one
segment of k steps at a time from the trajectory and runs that segment's
adjoint.  Gradients must match the adjoint over the stored stages of the
whole rollout.

Burgers keeps its state in physical space, so a segment recomputed from the
stored output repeats the forward's arithmetic: without a time-dependent
source the state cotangent is bit-identical (the sums over all stages, the
cotangents of s and nu, round differently); with the 'learning' forcing the
segment's clock t0 + k0 dt rounds differently, 1e-12 relative in float64.
KS keeps its state in spectral space across steps and restarts a segment
from the physical output, whose transform rounds differently: 1e-11
relative in float64, 2e-5 in float32 (40 steps of the chaotic KS map).
"""
import math

import pytest
import torch

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch, KSSimParams, KSSolverTorch
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _smooth(b, n, seed, kmax=6):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        u += torch.randn(b, 1, generator=gen, dtype=torch.float64) / k * torch.sin(
            k * x + 2 * math.pi * torch.rand(b, 1, generator=gen, dtype=torch.float64))
    return u / u.abs().amax(-1, keepdim=True)


def _ks_u0(b, n, L, seed):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    return sum(0.5 * torch.randn(b, 1, generator=gen, dtype=torch.float64) * torch.cos(
        2 * math.pi * k * x / L + 2 * math.pi * torch.rand(b, 1, generator=gen, dtype=torch.float64))
        for k in range(1, 11))


def _ks_grads(dtype, k, b=16, n=128, H=40, L=64.0, with_s=True, doublestep=False):
    sv = KSSolverTorch(KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = doublestep
    Ld = torch.tensor([L], dtype=torch.float64, device="cuda")
    u = _ks_u0(b, n, L, 3).to("cuda", dtype).requires_grad_(True)
    s = (1e-3 * torch.randn(b, n, generator=torch.Generator().manual_seed(4), dtype=torch.float64)).to(
        "cuda", dtype).requires_grad_(with_s)
    w = torch.randn(b, H, n, generator=torch.Generator().manual_seed(5), dtype=torch.float64).to("cuda", dtype)
    sv.correction_s = s
    traj = sv.rollout(u, H, Ld, checkpoint_every=k)
    (traj * w).sum().backward()
    return traj.detach(), u.grad, (s.grad if with_s else None)


@pytest.mark.parametrize("k", [1, 7, 8, 39])
def test_ks_checkpointed_fused_rollout_fp64(cuda, k):
    traj0, gu0, gs0 = _ks_grads(torch.float64, None)
    traj, gu, gs = _ks_grads(torch.float64, k)
    assert torch.equal(traj, traj0)  # the forward is the same launch
    assert rel_l2(gu, gu0) < 1e-11
    assert rel_l2(gs, gs0) < 1e-11


def test_ks_checkpointed_fused_rollout_fp32_and_doublestep(cuda):
    _, gu0, gs0 = _ks_grads(torch.float32, None)
    _, gu, gs = _ks_grads(torch.float32, 8)
    assert rel_l2(gu, gu0) < 2e-5 and rel_l2(gs, gs0) < 2e-5
    _, gu0, _ = _ks_grads(torch.float64, None, doublestep=True, with_s=False)
    _, gu, _ = _ks_grads(torch.float64, 5, doublestep=True, with_s=False)
    assert rel_l2(gu, gu0) < 1e-11


def _burgers_grads(k, dtype=torch.float64, forcing=False, t_tensor=False, b=8, n=64, H=30):
    gen = torch.Generator().manual_seed(7)
    nu = torch.tensor([0.005 * (i + 1) for i in range(b)], dtype=dtype, device="cuda")[:, None]
    nu.requires_grad_(True)
    J = 3
    kw = dict(equation_type="learning") if forcing else {}
    p = BGSimParams(dt=0.01, resolution=n, nu=nu, batch_size=b, time_scheme="SSPRK3", device="cuda", **kw)
    if forcing:
        p.A = (0.5 * torch.randn(b, J, generator=gen, dtype=torch.float64)).to("cuda", dtype)
        p.omega = (torch.rand(b, J, generator=gen, dtype=torch.float64) * 2).to("cuda", dtype)
        p.phi = (torch.rand(b, J, generator=gen, dtype=torch.float64) * 2 * math.pi).to("cuda", dtype)
        p.l = torch.randint(1, 4, (b, J), generator=gen).to("cuda", dtype)
    sv = BurgersSolverTorch(p)
    if t_tensor:
        sv.t = torch.linspace(0.0, 0.7, b, dtype=dtype, device="cuda")[:, None]
    u = _smooth(b, n, 8).to("cuda", dtype).requires_grad_(True)
    s = (0.05 * torch.randn(b, n, generator=gen, dtype=torch.float64)).to("cuda", dtype).requires_grad_(True)
    w = torch.randn(b, H, n, generator=gen, dtype=torch.float64).to("cuda", dtype)
    sv.correction_s = s
    traj = sv.rollout(u, H, checkpoint_every=k)
    (traj * w).sum().backward()
    return traj.detach(), u.grad, s.grad, nu.grad


@pytest.mark.parametrize("k", [1, 4, 7, 29])
def test_burgers_checkpointed_fused_rollout_is_exact(cuda, k):
    # homogeneous source: the recomputed stages repeat the forward bit for bit
    traj0, gu0, gs0, gn0 = _burgers_grads(None)
    traj, gu, gs, gn = _burgers_grads(k)
    assert torch.equal(traj, traj0)
    assert torch.equal(gu, gu0)
    # g_s and g_nu are sums over every stage: per segment, then over the
    # segments, which rounds differently from one running sum
    assert rel_l2(gs, gs0) < 1e-14 and rel_l2(gn, gn0) < 1e-14


@pytest.mark.parametrize("t_tensor", [False, True])
def test_burgers_checkpointed_with_forcing_clock(cuda, t_tensor):
    _, gu0, gs0, gn0 = _burgers_grads(None, forcing=True, t_tensor=t_tensor)
    _, gu, gs, gn = _burgers_grads(6, forcing=True, t_tensor=t_tensor)
    assert rel_l2(gu, gu0) < 1e-12 and rel_l2(gs, gs0) < 1e-12 and rel_l2(gn, gn0) < 1e-12


def test_burgers_checkpointed_fp32_and_simulate_rollout(cuda):
    _, gu0, gs0, _ = _burgers_grads(None, dtype=torch.float32)
    _, gu, gs, _ = _burgers_grads(5, dtype=torch.float32)
    assert torch.equal(gu, gu0) and rel_l2(gs, gs0) < 1e-6
    # simulate_rollout passes checkpoint_every to the fused rollout
    b, n, H = 4, 64, 12
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b, time_scheme="SSPRK3",
                                        device="cuda"))
    u = _smooth(b, n, 9).to(cuda).requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(None, sv, u, H, None, checkpoint_every=5)
    traj.square().sum().backward()
    g_ck = u.grad.clone()
    u.grad = None
    traj, _ = inc_amd.simulate_rollout(None, sv, u, H, None)
    traj.square().sum().backward()
    assert torch.equal(g_ck, u.grad)


def test_checkpointed_rollout_peak_memory(cuda):
    # KS, B=256, N=256, 200 steps: the full adjoint keeps 4 stage states per
    # step (210 MB in float32); every 20 steps keeps 21 MB of them
    b, n, H, L = 256, 256, 200, 64.0
    sv = KSSolverTorch(KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    Ld = torch.tensor([L], dtype=torch.float64, device=cuda)
    u0 = _ks_u0(b, n, L, 11).float().to(cuda)
    peaks = {}
    for k in (None, 20):
        u = u0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        traj = sv.rollout(u, H, Ld, checkpoint_every=k)
        traj.sum().backward()
        torch.cuda.synchronize()
        peaks[k] = torch.cuda.max_memory_allocated() - base
        del traj
    stages_full = b * H * 4 * n * 4
    assert peaks[None] >= stages_full
    assert peaks[20] < peaks[None] - 0.8 * stages_full, peaks
