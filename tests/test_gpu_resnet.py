"""ResNet1D, the reference's correction network (models/models_1d/resnet.py:22-52),
on this library's convolution kernels (ReLU fused into the convolutions):
against the same module run by torch (backend='torch', torch.nn.Conv1d
circular convolutions and ReLU) in float64 and float32, forward and the
gradients of the state and of every parameter; and as the INC correction
of a Burgers rollout on the step-by-step path, against the FP64 oracle
driven by a CPU float64 copy of the network.

Bars: float64 1e-12 relative; float32 1e-5 (outputs) and 1e-4 (gradients).
"""
import math

import pytest
import torch

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch, ResNet1D
from oracle import BurgersOracle, burgers_params
from oracle.rollout import simulate_rollout as oracle_rollout
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

def _pair(dtype, seed=0, **kw):
    torch.manual_seed(seed)
    hip = ResNet1D(**kw).to("cuda", dtype)
    ref = ResNet1D(backend="torch", **kw).to("cuda", dtype)
    ref.load_state_dict(hip.state_dict())
    return hip, ref


def _fields(b, n, seed):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, 7):
        u += torch.randn(b, 1, generator=gen, dtype=torch.float64) / k * torch.sin(
            k * x + 2 * math.pi * torch.rand(b, 1, generator=gen, dtype=torch.float64))
    return u


@pytest.mark.parametrize("dtype,tol_f,tol_g", [(torch.float64, 1e-12, 1e-12), (torch.float32, 1e-5, 1e-4)])
@pytest.mark.parametrize("L", ["float", "tensor"])
def test_resnet_hip_matches_torch(cuda, dtype, tol_f, tol_g, L):
    b, n = 8, 128
    hip, ref = _pair(dtype)
    Lv = 2 * math.pi if L == "float" else torch.linspace(4.0, 9.0, b, dtype=torch.float64, device=cuda)
    u0 = _fields(b, n, 1).to(cuda, dtype)
    w = torch.randn(b, n, generator=torch.Generator().manual_seed(2), dtype=torch.float64).to(cuda, dtype)
    outs, grads = [], []
    for m in (hip, ref):
        u = u0.clone().requires_grad_(True)
        y = m(u, Lv)
        (y * w).sum().backward()
        outs.append(y.detach())
        grads.append([u.grad] + [p.grad for p in m.parameters()])
    assert rel_l2(outs[0], outs[1]) < tol_f
    for a, c in zip(*grads):
        assert rel_l2(a, c) < tol_g
    with torch.no_grad():
        assert rel_l2(hip(u0, Lv), outs[1]) < tol_f


def test_resnet_as_burgers_inc_correction_against_oracle(cuda):
    b, n, H = 4, 64, 6
    torch.manual_seed(5)
    cpu = ResNet1D(L=2 * math.pi, init_features=16, num_blocks=2, backend="torch").double()
    with torch.no_grad():
        cpu.final.weight.mul_(0.05)
    net = ResNet1D(L=2 * math.pi, init_features=16, num_blocks=2).cuda().double()
    net.load_state_dict(cpu.state_dict())
    u0 = _fields(b, n, 7)
    target = 0.3 * torch.randn(b, H, n, generator=torch.Generator().manual_seed(8), dtype=torch.float64)
    so = BurgersOracle(burgers_params(n, 0.01, nu=0.01, batch_size=b, time_scheme="SSPRK3"))
    uo = u0.clone().requires_grad_(True)
    ref, ref_c = oracle_rollout(cpu, so, uo, H, "INC")
    ((ref - target) ** 2).mean().backward()
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b, time_scheme="SSPRK3",
                                        device="cuda"))
    u = u0.to(cuda).requires_grad_(True)
    traj, corr = inc_amd.simulate_rollout(net, sv, u, H, "INC")
    ((traj - target.to(cuda)) ** 2).mean().backward()
    assert rel_l2(traj.detach(), ref.detach()) < 1e-12 and rel_l2(corr.detach(), ref_c.detach()) < 1e-12
    assert rel_l2(u.grad, uo.grad) < 1e-10
    for p, q in zip(net.parameters(), cpu.parameters()):
        assert rel_l2(p.grad, q.grad) < 1e-10
