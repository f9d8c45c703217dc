"""The KS INC rollout in one library call (inc_ks_inc_fwd / _adj, reached
from simulate_rollout(model, ks_solver, u, steps, 'INC', L) with a
CorrectionCNN): against the step-by-step path (the same network and solver
kernels driven from Python, fused=False) and, at the KS INC training
workload's shape, against the FP64 oracle with a CPU float64 copy of the
network.

Bars: float64 1e-10 relative against the step loop; float32 1e-5 (states)
and 1e-4 (gradients) against the step loop and the oracle.
"""
import math

import pytest
import torch

import inc_amd
from inc_amd import KSSimParams, KSSolverTorch
from inc_amd.correction import CorrectionCNN
from oracle import KSOracle, ks_params
from oracle.rollout import simulate_rollout as oracle_rollout
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _u0(b, n, Ls, seed, kmax=10):
    gen = torch.Generator().manual_seed(seed)
    out = []
    for i in range(b):
        L = Ls[i % len(Ls)]
        x = torch.arange(n, dtype=torch.float64) * (L / n)
        u = torch.zeros(n, dtype=torch.float64)
        for k in range(1, kmax + 1):
            u += 0.5 * torch.randn(1, generator=gen, dtype=torch.float64) * torch.cos(
                2 * math.pi * k * x / L + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
        out.append(u)
    return torch.stack(out)


def _run(dtype, fused, scheme="ETDRK4", doublestep=False, Ls=(64.0,), b=6, n=128, H=12, use_corr=True,
         keep="auto", depth=3, width=16, k=5):
    Ld = torch.tensor([Ls[i % len(Ls)] for i in range(b)] if len(Ls) > 1 else list(Ls), dtype=torch.float64,
                      device="cuda")
    sv = KSSolverTorch(KSSimParams(list(Ls), resolution=n, dt=0.25, time_scheme=scheme, device="cuda"))
    sv.doublestep = doublestep
    sv.keep_activations = keep
    net = CorrectionCNN(depth, width, k, generator=torch.Generator().manual_seed(1), scale=0.01).to("cuda", dtype)
    u = _u0(b, n, Ls, 2).to("cuda", dtype).requires_grad_(True)
    w = torch.randn(b, H, n, generator=torch.Generator().manual_seed(3), dtype=torch.float64).to("cuda", dtype)
    wc = torch.randn(b, H, n, generator=torch.Generator().manual_seed(4), dtype=torch.float64).to("cuda", dtype)
    traj, corr = inc_amd.simulate_rollout(net, sv, u, H, "INC", L=Ld, fused=fused)
    loss = (traj * w).sum() + (100.0 * (corr * wc).sum() if use_corr else 0.0)
    loss.backward()
    grads = [u.grad] + [p.grad for p in net.parameters()]
    return traj.detach(), corr.detach(), grads, sv


@pytest.mark.parametrize("scheme,doublestep", [("ETDRK4", False), ("ETRK2", True), ("ETD1", False)])
def test_ks_inc_one_call_matches_step_loop_fp64(cuda, scheme, doublestep):
    t0, c0, g0, _ = _run(torch.float64, False, scheme, doublestep)
    t1, c1, g1, sv = _run(torch.float64, True, scheme, doublestep)
    assert rel_l2(t1, t0) < 1e-12 and rel_l2(c1, c0) < 1e-12
    for a, b in zip(g1, g0):
        assert rel_l2(a, b) < 1e-10
    assert sv.correction_s is not None and torch.equal(sv.correction_s, c1[:, -1])


def test_ks_inc_one_call_two_domain_sizes_and_fp32(cuda):
    t0, c0, g0, _ = _run(torch.float64, False, Ls=(64.0, 40.0), use_corr=False)
    t1, c1, g1, _ = _run(torch.float64, True, Ls=(64.0, 40.0), use_corr=False)
    assert rel_l2(t1, t0) < 1e-12
    for a, b in zip(g1, g0):
        assert rel_l2(a, b) < 1e-10
    t2, _, g2, _ = _run(torch.float32, True)
    t3, _, g3, _ = _run(torch.float32, False)
    assert rel_l2(t2, t3) < 1e-5
    for a, b in zip(g2, g3):
        assert rel_l2(a, b) < 1e-4


@pytest.mark.parametrize("dtype,n,shape", [(torch.float64, 128, (3, 16, 5)), (torch.float32, 256, (6, 48, 7)),
                                           (torch.float32, 96, (4, 32, 3))])
def test_ks_inc_stored_activations_match_recompute(cuda, dtype, n, shape):
    # the forward's stored hidden activations and GELU derivatives are the
    # very values the adjoint's repeat of the network computes (same kernels,
    # same inputs), so both adjoints agree bit for bit
    depth, width, k = shape
    t0, c0, g0, _ = _run(dtype, True, n=n, b=5, H=7, keep=False, depth=depth, width=width, k=k)
    t1, c1, g1, _ = _run(dtype, True, n=n, b=5, H=7, keep=True, depth=depth, width=width, k=k)
    assert torch.equal(t0, t1) and torch.equal(c0, c1)
    for i, (a, b) in enumerate(zip(g1, g0)):
        assert torch.equal(a, b), i


def test_activation_store_policy(cuda):
    net = CorrectionCNN(5, 24, 5).cuda()
    like = torch.zeros(3, device="cuda")
    assert net.activation_store(4, 100, like, False) is None
    a = net.activation_store(4, 100, like, True)
    assert a.shape == (4, 4, 2, 2400) and a.dtype == like.dtype and a.is_cuda
    assert net.activation_store(4, 100, like, "auto").shape == a.shape     # small: kept
    free, total = torch.cuda.mem_get_info()
    assert net.activation_store(1, total, like, "auto") is None               # over half the free memory
    # a buffer the caching allocator holds unused counts as free: the same
    # decision on the next training step
    big = torch.empty(int(0.6 * free) // 4, device="cuda")
    del big
    n_pts = int(0.3 * free) // (4 * 2 * 24 * 4)    # 0.3 of the free memory: above half the driver's 0.4
    try:
        assert net.activation_store(1, n_pts, like, "auto") is not None
    finally:
        torch.cuda.empty_cache()
    with pytest.raises(ValueError, match="keep_activations"):
        net.activation_store(4, 100, like, "yes")


def test_ks_inc_one_call_without_gradient(cuda):
    b, n, H = 4, 64, 8
    sv = KSSolverTorch(KSSimParams([32.0], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    net = CorrectionCNN(3, 8, 3, generator=torch.Generator().manual_seed(5), scale=0.01).cuda()
    u = _u0(b, n, (32.0,), 6).float().cuda()
    Ld = torch.tensor([32.0], dtype=torch.float64, device=cuda)
    with torch.no_grad():
        t1, c1 = inc_amd.simulate_rollout(net, sv, u, H, "INC", L=Ld)
        t0, c0 = inc_amd.simulate_rollout(net, sv, u, H, "INC", L=Ld, fused=False)
    assert rel_l2(t1, t0) < 1e-6 and rel_l2(c1, c0) < 1e-6


def test_ks_inc_one_call_at_the_training_shape_against_oracle(cuda):
    # KS INC training workload: N=512, L=100, dt=0.25, 64 ETDRK4 steps,
    # CorrectionCNN 6 x 48, k=7, output x 0.01; 4 of its trajectories
    b, n, L, H = 4, 512, 100.0, 64
    gen = torch.Generator().manual_seed(1003)
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u0 = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, 17):
        u0 += torch.randn(b, 1, generator=gen, dtype=torch.float64) * 0.5 * torch.cos(
            2 * math.pi * k * x / L + 2 * math.pi * torch.rand(b, 1, generator=gen, dtype=torch.float64))
    target = 0.3 * torch.randn(b, H, n, generator=torch.Generator().manual_seed(11), dtype=torch.float64)
    cnn = CorrectionCNN(6, 48, 7, generator=torch.Generator().manual_seed(3), backend="torch", scale=0.01).double()
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
    so.doublestep = False
    uo = u0.clone().requires_grad_(True)
    ref, _ = oracle_rollout(lambda v, L_: cnn(v), so, uo, H, "INC", L=so.domain_size)
    ((ref - target) ** 2).mean().backward()
    net = CorrectionCNN(6, 48, 7, scale=0.01).to(cuda)
    net.load_state_dict(cnn.state_dict())
    sv = KSSolverTorch(KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    Ld = torch.tensor([L], dtype=torch.float64, device=cuda)
    u = u0.float().to(cuda).requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(net, sv, u, H, "INC", L=Ld)
    ((traj - target.float().to(cuda)) ** 2).mean().backward()
    assert rel_l2(u.grad, uo.grad) < 1e-4
    for p, q in zip(net.parameters(), cnn.parameters()):
        assert rel_l2(p.grad, q.grad) < 1e-4
    assert rel_l2(traj[:, 0].detach(), ref[:, 0].detach()) < 1e-5
