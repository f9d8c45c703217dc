"""Gradients with respect to the viscosity nu: the reference's autograd
through calc_linear's nu * (u @ D2_T) (solver_1d.py:144-146) when nu is a
tensor that requires grad.  Per-trajectory [B, 1] and one-element nu,
through the fused rollout, the step loop, the per-trajectory adaptive
rollout, the fused-MSE rollout and the right-hand-side pieces, against FP64
oracle autograd and central differences of the float64 HIP path.

Bars: float64 1e-10 relative to the oracle; central differences 1e-7 of the
directional derivative; float32 1e-4 relative to the oracle.
"""
import math

import pytest
import torch

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch
from inc_amd.correction import CorrectionCNN
from oracle import BurgersOracle, burgers_params
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


def _oracle_rollout(u0, nu, H, **kw):
    """Oracle trajectory [B, H, N] (autograd graph kept) for nu of any shape."""
    b, n = u0.shape
    so = BurgersOracle(burgers_params(n, kw.pop("dt", 0.01), nu=nu, batch_size=b, **kw))
    cur, outs = u0, []
    for _ in range(H):
        cur = so.step(cur)
        outs.append(cur)
    return torch.stack(outs, 1)


def _gpu_solver(n, b, nu, dtype=torch.float64, dt=0.01, **kw):
    return BurgersSolverTorch(BGSimParams(dt=dt, resolution=n, nu=nu, batch_size=b, device="cuda", **kw))


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK3"])
@pytest.mark.parametrize("conv", ["WENO", "FDM"])
def test_nu_grad_per_trajectory_fp64_matches_oracle(cuda, scheme, conv):
    n, b, H = 64, 5, 6
    u0 = _smooth(b, n, 1)
    nu = torch.tensor([0.004, 0.01, 0.02, 0.03, 0.05], dtype=torch.float64)[:, None]
    w = torch.randn(b, H, n, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
    nu_o = nu.clone().requires_grad_(True)
    uo = u0.clone().requires_grad_(True)
    (_oracle_rollout(uo, nu_o, H, time_scheme=scheme, convection_term=conv) * w).sum().backward()
    nu_g = nu.to(cuda).requires_grad_(True)
    sv = _gpu_solver(n, b, nu_g, time_scheme=scheme, convection_term=conv)
    u = u0.to(cuda).requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(None, sv, u, H, None)
    (traj * w.to(cuda)).sum().backward()
    assert nu_g.grad.shape == nu.shape
    assert rel_l2(nu_g.grad, nu_o.grad) < 1e-10
    assert rel_l2(u.grad, uo.grad) < 1e-10
    # the step loop accumulates the same gradient step by step
    nu_s = nu.to(cuda).requires_grad_(True)
    sv2 = _gpu_solver(n, b, nu_s, time_scheme=scheme, convection_term=conv)
    cur, outs = u0.to(cuda), []
    for _ in range(H):
        cur = sv2.step(cur)
        outs.append(cur)
    (torch.stack(outs, 1) * w.to(cuda)).sum().backward()
    assert rel_l2(nu_s.grad, nu_o.grad) < 1e-10


def test_nu_grad_scalar_and_fp32(cuda):
    n, b, H = 128, 4, 8
    u0 = _smooth(b, n, 3)
    w = torch.randn(b, H, n, generator=torch.Generator().manual_seed(4), dtype=torch.float64)
    nu_o = torch.tensor(0.015, dtype=torch.float64, requires_grad=True)
    (_oracle_rollout(u0, nu_o, H, time_scheme="SSPRK3") * w).sum().backward()
    for dtype, tol in ((torch.float64, 1e-10), (torch.float32, 1e-4)):
        nu_g = torch.tensor(0.015, dtype=dtype, device=cuda, requires_grad=True)
        sv = _gpu_solver(n, b, nu_g, time_scheme="SSPRK3")
        traj, _ = inc_amd.simulate_rollout(None, sv, u0.to(cuda, dtype), H, None)
        (traj * w.to(cuda, dtype)).sum().backward()
        assert nu_g.grad.shape == () and nu_g.grad.dtype == dtype
        assert abs(float(nu_g.grad) - float(nu_o.grad)) <= tol * abs(float(nu_o.grad)), dtype


def test_nu_grad_central_differences_fp64(cuda):
    n, b, H = 64, 3, 10
    u0 = _smooth(b, n, 5).to(cuda)
    nu0 = torch.tensor([[0.006], [0.012], [0.025]], dtype=torch.float64, device=cuda)
    w = torch.randn(b, H, n, generator=torch.Generator().manual_seed(6), dtype=torch.float64).to(cuda)

    def loss(nu):
        sv = _gpu_solver(n, b, nu, time_scheme="SSPRK3")
        traj, _ = inc_amd.simulate_rollout(None, sv, u0, H, None)
        return (traj * w).sum()

    nu = nu0.clone().requires_grad_(True)
    loss(nu).backward()
    gen = torch.Generator().manual_seed(7)
    h = 1e-6
    with torch.no_grad():
        for _ in range(4):
            d = torch.randn(b, 1, generator=gen, dtype=torch.float64).to(cuda)
            fd = (loss(nu0 + h * d) - loss(nu0 - h * d)) / (2 * h)
            an = (nu.grad * d).sum()
            assert abs(float(fd - an)) <= 1e-7 * abs(float(an)) + 1e-9, (float(fd), float(an))


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK3"])
def test_nu_grad_adaptive_per_trajectory_matches_oracle(cuda, scheme):
    # per-trajectory sub-stepping; the oracle runs one trajectory at a time
    # (its batch-wide sub-steps are then that trajectory's)
    n, b, dt, cfl, H = 64, 4, 0.05, 0.3, 3
    amps = torch.tensor([0.5, 1.0, 2.0, 3.0], dtype=torch.float64)[:, None]
    u0 = amps * _smooth(b, n, 8)
    nus = [0.005, 0.01, 0.02, 0.04]
    w = torch.randn(b, H, n, generator=torch.Generator().manual_seed(9), dtype=torch.float64)
    ref = []
    for i in range(b):
        nu_i = torch.tensor([[nus[i]]], dtype=torch.float64, requires_grad=True)
        tr = _oracle_rollout(u0[i:i + 1], nu_i, H, dt=dt, adaptive_CFL=cfl, time_scheme=scheme)
        (tr * w[i:i + 1]).sum().backward()
        ref.append(nu_i.grad.reshape(1))
    ref = torch.cat(ref)[:, None]
    nu_g = torch.tensor(nus, dtype=torch.float64, device=cuda)[:, None].requires_grad_(True)
    sv = _gpu_solver(n, b, nu_g, dt=dt, adaptive_CFL=cfl, adaptive_mode="trajectory", max_substeps=64,
                     time_scheme=scheme)
    traj, _ = inc_amd.simulate_rollout(None, sv, u0.to(cuda), H, None)
    (traj * w.to(cuda)).sum().backward()
    assert rel_l2(nu_g.grad, ref) < 1e-10


def test_nu_grad_rhs_pieces(cuda):
    n, b = 96, 3
    u0 = _smooth(b, n, 10)
    g = torch.randn(b, n, generator=torch.Generator().manual_seed(11), dtype=torch.float64)
    nu_o = torch.tensor([[0.01], [0.02], [0.03]], dtype=torch.float64, requires_grad=True)
    so = BurgersOracle(burgers_params(n, 0.01, nu=nu_o, batch_size=b))
    (so.calc_linear(u0) * g).sum().backward()
    nu_g = nu_o.detach().to(cuda).requires_grad_(True)
    sv = _gpu_solver(n, b, nu_g)
    (sv.calc_linear(u0.to(cuda)) * g.to(cuda)).sum().backward()
    assert rel_l2(nu_g.grad, nu_o.grad) < 1e-12
    # pieces without diffusion do not depend on nu
    nu_g.grad = None
    (sv.calc_nonlinear_weno(u0.to(cuda)) * g.to(cuda)).sum().backward()
    assert nu_g.grad is not None and float(nu_g.grad.abs().max()) == 0.0


def test_nu_grad_fused_mse_matches_unfused(cuda):
    n, b, H = 64, 4, 5
    u0 = _smooth(b, n, 12).to(cuda)
    target = (0.3 * torch.randn(b, H, n, generator=torch.Generator().manual_seed(13),
                                dtype=torch.float64)).to(cuda)
    nu0 = torch.tensor([[0.01], [0.02], [0.015], [0.03]], dtype=torch.float64, device=cuda)
    nu_a = nu0.clone().requires_grad_(True)
    loss_a, _ = inc_amd.simulate_rollout_mse(_gpu_solver(n, b, nu_a, time_scheme="SSPRK3"), u0, target)
    loss_a.backward()
    nu_b = nu0.clone().requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(None, _gpu_solver(n, b, nu_b, time_scheme="SSPRK3"), u0, H, None)
    ((traj - target) ** 2).mean().backward()
    assert rel_l2(nu_a.grad, nu_b.grad) < 1e-12


def test_nu_grad_in_kernel_inc_raises(cuda):
    n, b = 64, 2
    nu = torch.tensor(0.01, dtype=torch.float32, device=cuda, requires_grad=True)
    sv = _gpu_solver(n, b, nu, time_scheme="SSPRK3")
    net = CorrectionCNN(3, 32, 5, generator=torch.Generator().manual_seed(1)).cuda()
    u = _smooth(b, n, 14).float().to(cuda)
    with pytest.raises(ValueError, match="nu"):
        sv.rollout_inc(net, u, 3)
    with torch.no_grad():
        traj, _ = sv.rollout_inc(net, u, 3)
    assert bool(torch.isfinite(traj).all())
