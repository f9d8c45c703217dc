"""GPU parity of the rollout with its MSE loss fused into the kernels
(simulate_rollout_mse: the trainer's simulate_rollout followed by its
MSELoss, trainer/trainer_1d.py:154-183 and :723).

Each case is checked against the same rollout with torch's MSELoss and
autograd (the unfused path, itself pinned to the oracle by the other GPU
tests) and, for float64, against the FP64 oracle's rollout + MSE + autograd.
Shapes cover every layout of the per-trajectory loss partials: teams narrower
than a wavefront, one wavefront, two and four wavefronts, an unpaired KS
trajectory (odd batch) and unpacked KS trajectories (mixed domain sizes).
Tolerances: float64 1e-11 relative; float32 1e-5 (loss and gradients against
the float32 unfused path).
"""
import math

import pytest
import torch
from torch.nn import functional as F

import inc_amd
from inc_amd import simulate_rollout_mse
from oracle import BurgersOracle, KSOracle, burgers_params, ks_params
from oracle.rollout import simulate_rollout as oracle_rollout
from tests.conftest import rel_l2
from tests.test_gpu_burgers import make_pair, smooth
from tests.test_gpu_ks import gpu_solver, modes

pytestmark = pytest.mark.gpu


def _tol(dtype):
    return 1e-11 if dtype == torch.float64 else 1e-5


def _burgers_case(n, b, H, seed, nu=0.02):
    """Inputs and a GPU solver factory; dt at 0.3 of both the advective
    (max|u| = 1) and the diffusive limits of the grid."""
    gen = torch.Generator().manual_seed(seed)
    u0 = smooth(b, n, gen)
    target = (smooth(b * H, n, gen) * 0.8).reshape(b, H, n)
    s = 0.1 * smooth(b, n, gen)
    w = torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    dx = 2 * math.pi / n
    _, gpu = make_pair(n, b, dt=0.3 * min(dx, dx * dx / nu), nu=nu, scheme="SSPRK3")
    return u0, target, s, w, gpu


# (n, batch, team): 16-thread teams (four per wavefront, 4 points each),
# 64-thread teams, 256-thread teams (four wavefronts); 0 = library choice
BURGERS = [(40, 5, 16), (128, 3, 64), (1024, 2, 256), (200, 4, 0)]


@pytest.mark.parametrize("shape", BURGERS, ids=lambda s: "n{}_b{}_t{}".format(*s))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_burgers_rollout_mse_matches_unfused(cuda, shape, dtype):
    n, b, team = shape
    H = 6
    u0, target, s, w, gpu = _burgers_case(n, b, H, 100 + n)
    c = lambda t: t.to(cuda, dtype)
    # fused: loss in the forward kernel, 2 (u - target) / M seeded in the adjoint, plus an extra
    # trajectory cotangent w and the correction's gradient
    sv1 = gpu(dtype, team=team)
    u1, s1 = c(u0).requires_grad_(True), c(s).requires_grad_(True)
    sv1.correction_s = s1
    loss1, traj1 = sv1.rollout_mse(u1, c(target))
    (3.0 * loss1 + (traj1 * c(w)).sum()).backward()
    # unfused: the same rollout, torch's MSELoss
    sv2 = gpu(dtype, team=team)
    u2, s2 = c(u0).requires_grad_(True), c(s).requires_grad_(True)
    sv2.correction_s = s2
    traj2 = sv2.rollout(u2, H)
    loss2 = F.mse_loss(traj2, c(target))
    (3.0 * loss2 + (traj2 * c(w)).sum()).backward()
    tol = _tol(dtype)
    assert abs(loss1.item() - loss2.item()) <= tol * abs(loss2.item())
    assert rel_l2(traj1, traj2) < tol
    assert rel_l2(u1.grad, u2.grad) < tol
    assert rel_l2(s1.grad, s2.grad) < tol
    assert abs(sv1.t - sv2.t) < 1e-12


def test_burgers_rollout_mse_fp64_matches_oracle(cuda):
    n, b, H = 64, 3, 5
    gen = torch.Generator().manual_seed(71)
    u0 = smooth(b, n, gen)
    target = smooth(b * H, n, gen).reshape(b, H, n)
    oracle, gpu = make_pair(n, b, dt=0.01, nu=0.01, scheme="SSPRK3")
    uo = u0.clone().requires_grad_(True)
    res, _ = oracle_rollout(None, oracle, uo, H, None)
    lo = F.mse_loss(res, target)
    lo.backward()
    sv = gpu(torch.float64)
    ug = u0.to(cuda).requires_grad_(True)
    lg, traj = simulate_rollout_mse(sv, ug, target.to(cuda))
    lg.backward()
    assert abs(lg.item() - lo.item()) <= 1e-11 * lo.item()
    assert rel_l2(traj, res.detach()) < 1e-11
    assert rel_l2(ug.grad, uo.grad) < 1e-11


def test_burgers_rollout_mse_loss_only_and_target_gradient(cuda):
    # the trajectory output unused (no extra cotangent passed), and a target
    # that requires grad: d loss / d target = -2 (u - target) / M
    n, b, H = 96, 2, 4
    u0, target, _, _, gpu = _burgers_case(n, b, H, 5)
    tg = target.to(cuda).requires_grad_(True)
    loss, traj = gpu(torch.float64).rollout_mse(u0.to(cuda).requires_grad_(True), tg)
    loss.backward()
    expect = -2.0 * (traj.detach() - tg.detach()) / traj.numel()
    assert rel_l2(tg.grad, expect) < 1e-12


def test_rollout_mse_rejects_a_target_of_the_wrong_shape(cuda):
    _, gpu = make_pair(64, 2, dt=0.01, nu=0.01)
    with pytest.raises(ValueError, match="target"):
        gpu(torch.float64).rollout_mse(torch.zeros(2, 64, device=cuda, dtype=torch.float64),
                                       torch.zeros(2, 3, 32, device=cuda, dtype=torch.float64))


# (n, batch, radix, domain sizes, scheme, double step): 8-lane teams (eight
# per wavefront), two- and four-wavefront teams, an odd batch (one unpaired
# trajectory), mixed domain sizes (every trajectory unpacked), ETD1 double step
KS = [(16, 3, 2, [22.0], "ETDRK4", False), (256, 5, 2, [64.0], "ETDRK4", False),
      (512, 2, 2, [100.0], "ETDRK4", False), (128, 3, 4, [22.0, 30.0, 41.0], "ETRK2", False),
      (64, 4, 0, [22.0], "ETD1", True)]


@pytest.mark.parametrize("case", KS, ids=lambda c: f"n{c[0]}_b{c[1]}_r{c[2]}_{c[4]}_{len(c[3])}L")
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ks_rollout_mse_matches_unfused(cuda, case, dtype):
    n, b, radix, Ls, scheme, double = case
    H = 5
    gen = torch.Generator().manual_seed(300 + n)
    Lb = (Ls * b)[:b] if len(Ls) > 1 else Ls
    u0 = modes(b, n, Lb[0], gen, kmax=6)
    target = modes(b * H, n, Lb[0], gen, kmax=6).reshape(b, H, n)
    s = 0.05 * modes(b, n, Lb[0], gen, kmax=4)
    w = torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    Ld = torch.tensor(Lb, device=cuda, dtype=torch.float64)
    c = lambda t: t.to(cuda, dtype)
    outs = []
    for fused in (True, False):
        sv = gpu_solver(Lb, n, 0.1, scheme, dtype=dtype, radix=radix)
        sv.doublestep = double
        ug, sg = c(u0).requires_grad_(True), c(s).requires_grad_(True)
        sv.correction_s = sg
        if fused:
            loss, traj = sv.rollout_mse(ug, c(target), Ld)
        else:
            traj = sv.rollout(ug, H, Ld)
            loss = F.mse_loss(traj, c(target))
        (2.0 * loss + (traj * c(w)).sum()).backward()
        outs.append((loss.item(), traj.detach(), ug.grad, sg.grad))
    (l1, t1, gu1, gs1), (l2, t2, gu2, gs2) = outs
    tol = _tol(dtype)
    assert abs(l1 - l2) <= tol * abs(l2)
    assert rel_l2(t1, t2) < tol
    assert rel_l2(gu1, gu2) < tol
    assert rel_l2(gs1, gs2) < tol


def test_ks_rollout_mse_fp64_matches_oracle(cuda):
    n, b, H, L = 64, 3, 6, 22.0
    gen = torch.Generator().manual_seed(77)
    u0 = modes(b, n, L, gen, kmax=6)
    target = modes(b * H, n, L, gen, kmax=6).reshape(b, H, n)
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.1, time_scheme="ETDRK4"))
    so.doublestep = False
    uo = u0.clone().requires_grad_(True)
    res, _ = oracle_rollout(None, so, uo, H, None, L=so.domain_size)
    lo = F.mse_loss(res, target)
    lo.backward()
    sv = gpu_solver([L], n, 0.1, "ETDRK4")
    sv.doublestep = False
    ug = u0.to(cuda).requires_grad_(True)
    lg, traj = simulate_rollout_mse(sv, ug, target.to(cuda), L=torch.tensor([L], device=cuda))
    lg.backward()
    assert abs(lg.item() - lo.item()) <= 1e-11 * lo.item()
    assert rel_l2(traj, res.detach()) < 1e-11
    assert rel_l2(ug.grad, uo.grad) < 1e-11


@pytest.mark.parametrize("case", [(128, 8, 5, 4, "SSPRK3", 12), (64, 5, 7, 3, "Euler", 6), (96, 3, 5, 2, "SSPRK3", 5)],
                         ids=lambda c: "n{}_b{}_k{}_d{}_{}_h{}".format(*c))
def test_inc_training_rollout_with_fused_mse(cuda, case):
    # simulate_rollout(model, solver, u, H, 'INC') + MSELoss(results, target)
    # (trainer_1d.py:154-183, :716-723) with the network inside the kernels and
    # the loss fused into them, against the same rollout with torch's MSE and
    # autograd (same kernels, the cotangent formed in torch), and against the
    # FP64 oracle driven by a CPU float64 copy of the network
    from inc_amd.correction import CorrectionCNN
    from oracle.rollout import simulate_rollout as oracle_rollout
    n, b, k, depth, scheme, H = case
    gen = torch.Generator().manual_seed(n + b)
    u0 = smooth(b, n, gen)
    target = 0.3 * torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    cpu = CorrectionCNN(depth, 32, k, generator=torch.Generator().manual_seed(7), backend="torch", scale=0.1).double()
    dx = 2 * math.pi / n
    dt = 0.3 * min(dx, dx * dx / 0.02)
    net = CorrectionCNN(depth, 32, k, scale=0.1).to(cuda)
    net.load_state_dict(cpu.state_dict())

    def solver():
        return inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=dt, resolution=n, nu=0.02, batch_size=b,
                                                              time_scheme=scheme, device="cuda"))
    tg = target.float().to(cuda)
    u = u0.float().to(cuda).requires_grad_(True)
    net.zero_grad(set_to_none=True)
    loss, traj, corr = inc_amd.simulate_rollout_mse(solver(), u, tg, model=net)
    loss.backward()
    g_fused = [u.grad.clone()] + [p.grad.clone() for p in net.parameters()]
    u2 = u0.float().to(cuda).requires_grad_(True)
    net.zero_grad(set_to_none=True)
    res, corr2 = inc_amd.simulate_rollout(net, solver(), u2, H, "INC")
    loss2 = torch.nn.functional.mse_loss(res, tg)
    loss2.backward()
    g_torch = [u2.grad.clone()] + [p.grad.clone() for p in net.parameters()]
    assert torch.equal(traj, res.detach()) and torch.equal(corr, corr2.detach())
    assert rel_l2(loss.detach().reshape(1), loss2.detach().reshape(1)) < 1e-6
    for a, c in zip(g_fused, g_torch):
        assert rel_l2(a, c) < 1e-5
    so = BurgersOracle(burgers_params(n, dt, nu=0.02, batch_size=b, time_scheme=scheme))
    uo = u0.clone().requires_grad_(True)
    ref, _ = oracle_rollout(cpu, so, uo, H, "INC")
    lo = ((ref - target) ** 2).mean()
    lo.backward()
    assert rel_l2(loss.detach().reshape(1), lo.detach().reshape(1)) < 1e-5
    assert rel_l2(g_fused[0], uo.grad) < 1e-4
    for a, p in zip(g_fused[1:], cpu.parameters()):
        assert rel_l2(a, p.grad) < 1e-4
