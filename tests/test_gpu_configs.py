"""The INC training workloads at their own shapes, against the FP64 oracle.

* Burgers INC (B = 512, N = 128, nu = 0.01, dt = 0.01, 16 SSP-RK3 steps,
  CorrectionCNN 4 x 32, k = 5): the whole batch through the fused training
  path (network inside the rollout and adjoint kernels) against the oracle
  solver driven by a CPU float64 copy of the network, both differentiated.
* KS INC (L = 100, N = 512, dt = 0.25, 64 ETDRK4 steps, checkpoint every 8,
  CorrectionCNN 6 x 48, k = 7, output x 0.01) on 4 trajectories of the batch:
  teacher-forced float32 steps against the oracle, the float32 gradients of
  u0 and of every parameter against oracle autograd, and the float64 HIP path
  against FP64 central differences in 32 random directions.

The loss is the MSE to a random target (not a nearby trajectory, so its
cotangent has no cancellation).  Bars (relative L2): one step 1e-5,
gradients 1e-4; central differences 1e-6 of the directional derivative
(+1e-11 absolute for the loss's rounding over 2h).
"""
import math

import pytest
import torch

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch, KSSimParams, KSSolverTorch
from inc_amd.correction import CorrectionCNN
from oracle import BurgersOracle, KSOracle, burgers_params, ks_params
from oracle.rollout import simulate_rollout as oracle_rollout
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _sine_states(b, n, seed):
    out = []
    for i in range(b):
        gen = torch.Generator().manual_seed(seed + i)
        x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
        u = torch.zeros(n, dtype=torch.float64)
        for k in range(1, 9):
            u += torch.randn(1, generator=gen, dtype=torch.float64) / k * torch.sin(
                k * x + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
        out.append(u / u.abs().max())
    return torch.stack(out)


def _cos_modes(b, n, L, seed, kmax, sd=0.5):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        u += torch.randn(b, 1, generator=gen, dtype=torch.float64) * sd * torch.cos(
            2 * math.pi * k * x / L + 2 * math.pi * torch.rand(b, 1, generator=gen, dtype=torch.float64))
    return u


def test_burgers_inc_training_full_batch_against_oracle(cuda):
    b, n, H = 512, 128, 16
    u0 = _sine_states(b, n, 2000)
    target = 0.3 * torch.randn(b, H, n, generator=torch.Generator().manual_seed(9), dtype=torch.float64)
    cnn = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2), backend="torch", scale=0.1).double()
    so = BurgersOracle(burgers_params(n, 0.01, nu=0.01, batch_size=b, time_scheme="SSPRK3"))
    uo = u0.clone().requires_grad_(True)
    ref, _ = oracle_rollout(cnn, so, uo, H, "INC")
    ((ref - target) ** 2).mean().backward()
    net = CorrectionCNN(4, 32, 5, scale=0.1).to(cuda)
    net.load_state_dict(cnn.state_dict())
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b, time_scheme="SSPRK3",
                                        device="cuda"))
    u = u0.float().to(cuda).requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(net, sv, u, H, "INC")
    ((traj - target.float().to(cuda)) ** 2).mean().backward()
    assert rel_l2(traj.detach(), ref.detach()) < 1e-5
    assert rel_l2(u.grad, uo.grad) < 1e-4
    for p, q in zip(net.parameters(), cnn.parameters()):
        assert rel_l2(p.grad, q.grad) < 1e-4


def _ks_case(dtype, b=4):
    n, L, H = 512, 100.0, 64
    u0 = _cos_modes(b, n, L, 1003, 16)
    cnn = CorrectionCNN(6, 48, 7, generator=torch.Generator().manual_seed(3), backend="torch", scale=0.01).double()
    net = CorrectionCNN(6, 48, 7, scale=0.01).to(cuda_dev(), dtype)
    net.load_state_dict(cnn.state_dict())
    sv = KSSolverTorch(KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    return n, L, H, u0, cnn, net, sv


def cuda_dev():
    return torch.device("cuda")


def test_ks_inc_at_its_shape_against_oracle(cuda):
    n, L, H, u0, cnn, net, sv = _ks_case(torch.float32)
    b = u0.shape[0]
    target = 0.3 * torch.randn(b, H, n, generator=torch.Generator().manual_seed(11), dtype=torch.float64)
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
    so.doublestep = False
    uo = u0.clone().requires_grad_(True)
    ref, _ = oracle_rollout(lambda v, L_: cnn(v), so, uo, H, "INC", L=so.domain_size)
    ((ref - target) ** 2).mean().backward()
    Ld = torch.tensor([L], dtype=torch.float64, device=cuda)
    u = u0.float().to(cuda).requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(lambda v, L_: net(v), sv, u, H, "INC", L=Ld, checkpoint_every=8)
    ((traj - target.float().to(cuda)) ** 2).mean().backward()
    assert rel_l2(u.grad, uo.grad) < 1e-4
    for p, q in zip(net.parameters(), cnn.parameters()):
        assert rel_l2(p.grad, q.grad) < 1e-4
    # teacher-forced float32 steps from the oracle's states with the oracle's correction
    ref = ref.detach()
    with torch.no_grad():
        for k in (0, 7, 8, 31, 63):
            prev = u0 if k == 0 else ref[:, k - 1]
            so.correction_s = cnn(prev)
            want = so.step(prev, so.domain_size)
            sv.correction_s = so.correction_s.float().to(cuda)
            assert rel_l2(sv.step(prev.float().to(cuda), Ld), want) < 1e-5, k
    sv.correction_s = None


def test_ks_inc_at_its_shape_central_differences_fp64(cuda):
    n, L, H, u0, cnn, net, sv = _ks_case(torch.float64, b=2)
    b = u0.shape[0]
    target = 0.3 * torch.randn(b, H, n, generator=torch.Generator().manual_seed(12), dtype=torch.float64)
    target = target.to(cuda)
    Ld = torch.tensor([L], dtype=torch.float64, device=cuda)

    def loss_fn(u, checkpoint_every=None):
        traj, _ = inc_amd.simulate_rollout(lambda v, L_: net(v), sv, u, H, "INC", L=Ld,
                                           checkpoint_every=checkpoint_every)
        return ((traj - target) ** 2).mean()

    u = u0.to(cuda).requires_grad_(True)
    net.zero_grad(set_to_none=True)
    loss_fn(u, checkpoint_every=8).backward()
    gen = torch.Generator().manual_seed(13)
    # unit directions, h = 1e-4: the loss (~1.8) carries ~1e-16 relative
    # rounding through 64 chaotic steps, which over 2h must stay below the
    # bar; the truncation error at this h measured below 3e-7 of the
    # derivative (h = 1e-6 gave up to 2.6e-5 from rounding alone)
    h = 1e-4
    with torch.no_grad():
        for _ in range(32):
            d = torch.randn(b, n, generator=gen, dtype=torch.float64).to(cuda)
            d /= d.norm()
            fd = (loss_fn(u0.to(cuda) + h * d) - loss_fn(u0.to(cuda) - h * d)) / (2 * h)
            an = (u.grad * d).sum()
            assert abs(float(fd - an)) <= 1e-6 * abs(float(an)) + 1e-11, (float(fd), float(an))
