"""GPU parity of the INC training rollout with the correction network's
transpose inside the adjoint kernel: the gradient of a loss on
simulate_rollout(model, solver, u, steps, 'INC') (trainer/trainer_1d.py:154-183,
the loss of CorrectionTrainer.process_batch, :699-727) with respect to the
initial state and every weight and bias of a CorrectionCNN.

Checked against the per-step path (one network evaluation and one solver
launch per step, autograd through the convolution kernels; itself pinned
to the oracle by the other GPU tests) on the same inputs, and against the
FP64 oracle solver driven by a CPU float64 copy of the network, both
differentiated by torch autograd.  The loss weights the trajectory and the
corrections, so both cotangents reach the kernel.  Shapes cover 2, 3 and 4
layers, kernel sizes 5 and 7, one and two position tiles per wavefront,
resolutions that are not multiples of 16, the smallest resolutions (8, 16),
Euler and SSP-RK3, and the training workload's batch (512 x 128, 16 steps).
The kernel runs float32; tolerances (relative L2): trajectory and
corrections 1e-5; gradients 2e-5 against the per-step float32 path
(measured on MI355X: 9e-8 .. 2.1e-6) and 5e-5 against the FP64 oracle.
"""
import copy
import math

import pytest
import torch

from inc_amd import BGSimParams, BurgersSolverTorch, simulate_rollout
from inc_amd.correction import CorrectionCNN
from inc_amd.solver_1d import inc_train_supported
from oracle import BurgersOracle, burgers_params
from oracle.rollout import simulate_rollout as oracle_rollout
from tests.conftest import rel_l2
from tests.test_gpu_burgers import _forcing, make_pair, smooth

pytestmark = pytest.mark.gpu


def _net(depth, k, seed, device, scale=1.0):
    gen = torch.Generator().manual_seed(seed)
    net = CorrectionCNN(depth, 32, k, generator=gen, scale=scale)
    with torch.no_grad():
        for cv in net.convs:
            cv.bias.copy_(0.1 * torch.randn(cv.bias.shape, generator=gen))
    return net.to(device=device, dtype=torch.float32)


def _solver(n, b, scheme, dtype=torch.float32, nu=0.02):
    dx = 2 * math.pi / n
    p = BGSimParams(dt=0.3 * min(dx, dx * dx / nu), resolution=n, nu=nu, batch_size=b, time_scheme=scheme,
                    device="cuda")
    return BurgersSolverTorch(p), p


def _loss(res, corr, target, wc):
    return ((res - target) ** 2).mean() + (corr * wc).mean()


def _run(net, u0, n, b, H, scheme, target, wc, fused):
    sv, _ = _solver(n, b, scheme)
    u = u0.clone().requires_grad_(True)
    net.zero_grad(set_to_none=True)
    res, corr = simulate_rollout(net, sv, u, H, "INC", fused=fused)
    _loss(res, corr, target, wc).backward()
    grads = [p.grad.detach().clone() for p in net.parameters()]
    return res.detach(), corr.detach(), u.grad.detach(), grads, sv


# (n, batch, kernel, depth, scheme, horizon)
# n = 8 and 16 are the smallest resolutions (the wrapped halo rows of the
# activations then cover a quarter or more of the grid)
CASES = [(128, 6, 5, 4, "SSPRK3", 7), (100, 3, 7, 3, "SSPRK3", 6), (64, 5, 5, 2, "Euler", 5),
         (48, 4, 7, 4, "Euler", 6), (37, 2, 5, 3, "SSPRK3", 4), (128, 512, 5, 4, "SSPRK3", 16),
         (8, 1, 5, 3, "SSPRK3", 3), (16, 2, 7, 4, "Euler", 2)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n{}_b{}_k{}_d{}_{}_h{}".format(*c))
def test_inc_training_rollout_matches_per_step_path(cuda, case):
    n, b, k, depth, scheme, H = case
    gen = torch.Generator().manual_seed(n + 7 * depth)
    u0 = smooth(b, n, gen).to(cuda, torch.float32)
    target = (0.9 * smooth(b * H, n, gen)).reshape(b, H, n).to(cuda, torch.float32)
    wc = torch.randn(b, H, n, generator=gen).to(cuda, torch.float32)
    net = _net(depth, k, 3 + n, cuda)
    assert inc_train_supported(net, n, torch.float32)
    r1, c1, gu1, gw1, sv1 = _run(net, u0, n, b, H, scheme, target, wc, fused=True)
    r2, c2, gu2, gw2, sv2 = _run(net, u0, n, b, H, scheme, target, wc, fused=False)
    assert rel_l2(r1, r2) < 1e-5
    assert rel_l2(c1, c2) < 1e-5
    assert rel_l2(gu1, gu2) < 2e-5
    for i, (g1, g2) in enumerate(zip(gw1, gw2)):
        assert rel_l2(g1, g2) < 2e-5, f"parameter {i} ({tuple(g1.shape)})"
    assert abs(sv1.t - sv2.t) < 1e-12
    assert rel_l2(sv1.correction_s.detach(), sv2.correction_s.detach()) < 1e-5


@pytest.mark.parametrize("scale", [1.0, 0.1])
def test_inc_training_rollout_matches_fp64_oracle(cuda, scale):
    # scale: the network's output scale (CorrectionCNN(scale=...)), applied in the kernels
    n, b, H, k, depth = 96, 3, 6, 5, 4
    gen = torch.Generator().manual_seed(21)
    u0 = smooth(b, n, gen)
    target = (0.9 * smooth(b * H, n, gen)).reshape(b, H, n)
    wc = torch.randn(b, H, n, generator=gen)
    net = _net(depth, k, 5, cuda, scale)
    cpu_net = copy.deepcopy(net).cpu().double()
    for cv in cpu_net.convs:
        cv.backend = "torch"
    _, p = _solver(n, b, "SSPRK3")
    oracle = BurgersOracle(burgers_params(n, p.dt, nu=p.nu, batch_size=b, time_scheme="SSPRK3"))
    uo = u0.clone().requires_grad_(True)
    res, corr = oracle_rollout(cpu_net, oracle, uo, H, "INC")
    _loss(res, corr, target, wc).backward()
    r, c, gu, gw, _ = _run(net, u0.to(cuda, torch.float32), n, b, H, "SSPRK3", target.to(cuda, torch.float32),
                           wc.to(cuda, torch.float32), fused=True)
    assert rel_l2(r.double().cpu(), res.detach()) < 1e-5
    assert rel_l2(c.double().cpu(), corr.detach()) < 1e-5
    assert rel_l2(gu.double().cpu(), uo.grad) < 5e-5
    for i, (g, pr) in enumerate(zip(gw, cpu_net.parameters())):
        assert rel_l2(g.double().cpu(), pr.grad) < 5e-5, f"parameter {i} ({tuple(pr.shape)})"


@pytest.mark.parametrize("H", [1, 5])
def test_inc_training_rollout_learning_forcing_per_trajectory_nu(cuda, H):
    # the trainer's solver (trainer_1d.py:338-358): 'learning' forcing and one
    # viscosity per trajectory, against the per-step path and the FP64 oracle
    n, b = 128, 4
    gen = torch.Generator().manual_seed(40 + H)
    nu = torch.exp(torch.empty(b, 1, dtype=torch.float64).uniform_(math.log(0.003), math.log(0.03), generator=gen))
    forcing = _forcing(b, gen)
    dx = 2 * math.pi / n
    oracle, gpu = make_pair(n, b, dt=0.3 * min(dx, dx * dx / 0.03), nu=nu, scheme="SSPRK3", eq="learning",
                            forcing=forcing)
    u0 = smooth(b, n, gen)
    target = (0.9 * smooth(b * H, n, gen)).reshape(b, H, n)
    wc = torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    net = _net(4, 5, 9, cuda)
    outs = []
    for fused in (True, False):
        sv = gpu(torch.float32)
        u = u0.to(cuda, torch.float32).requires_grad_(True)
        net.zero_grad(set_to_none=True)
        res, corr = simulate_rollout(net, sv, u, H, "INC", fused=fused)
        _loss(res, corr, target.to(cuda, torch.float32), wc.to(cuda, torch.float32)).backward()
        outs.append([res.detach(), corr.detach(), u.grad.detach()] + [p.grad.detach().clone() for p in net.parameters()])
    for i, (g1, g2) in enumerate(zip(*outs)):
        assert rel_l2(g1, g2) < 2e-5, f"output {i}"
    cpu_net = copy.deepcopy(net).cpu().double()
    for cv in cpu_net.convs:
        cv.backend = "torch"
    uo = u0.clone().requires_grad_(True)
    res, corr = oracle_rollout(cpu_net, oracle, uo, H, "INC")
    _loss(res, corr, target, wc).backward()
    assert rel_l2(outs[0][0].double().cpu(), res.detach()) < 1e-5
    assert rel_l2(outs[0][1].double().cpu(), corr.detach()) < 1e-5
    assert rel_l2(outs[0][2].double().cpu(), uo.grad) < 5e-5
    for i, (g, pr) in enumerate(zip(outs[0][3:], cpu_net.parameters())):
        assert rel_l2(g.double().cpu(), pr.grad) < 5e-5, f"parameter {i} ({tuple(pr.shape)})"


def test_inc_training_rollout_gradient_of_corrections_only(cuda):
    # a loss on the corrections alone: only g_corr reaches the kernel (scaled output)
    n, b, H = 64, 3, 5
    gen = torch.Generator().manual_seed(4)
    u0 = smooth(b, n, gen).to(cuda, torch.float32)
    net = _net(3, 5, 8, cuda, scale=0.05)
    out = []
    for fused in (True, False):
        sv, _ = _solver(n, b, "SSPRK3")
        u = u0.clone().requires_grad_(True)
        net.zero_grad(set_to_none=True)
        _, corr = simulate_rollout(net, sv, u, H, "INC", fused=fused)
        (corr ** 2).sum().backward()
        out.append([u.grad.clone()] + [p.grad.clone() for p in net.parameters()])
    for g1, g2 in zip(*out):
        assert rel_l2(g1, g2) < 2e-5


def test_inc_training_shapes_outside_the_kernel_take_the_per_step_path(cuda):
    # n = 256 exceeds the training adjoint's LDS: the rollout falls back to the per-step path
    n, b, H = 256, 2, 3
    net = _net(4, 5, 1, cuda)
    assert not inc_train_supported(net, n, torch.float32)
    gen = torch.Generator().manual_seed(2)
    u0 = smooth(b, n, gen).to(cuda, torch.float32)
    sv, _ = _solver(n, b, "SSPRK3")
    u = u0.clone().requires_grad_(True)
    res, corr = simulate_rollout(net, sv, u, H, "INC")
    (res.sum() + corr.sum()).backward()
    assert torch.isfinite(u.grad).all()
    assert res.grad_fn is not None and "IncRollout" not in type(res.grad_fn).__name__
