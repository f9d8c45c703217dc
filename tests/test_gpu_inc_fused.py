"""GPU parity of the INC rollout with the correction network evaluated inside
the Burgers rollout kernel (simulate_rollout(model, solver, u, steps, 'INC'),
trainer/trainer_1d.py:154-183, under torch.no_grad with a CorrectionCNN).

Checked against the per-step path (one network evaluation and one solver
launch per step, itself pinned to the oracle by the other GPU tests) on the
same inputs, and in float64 against the FP64 oracle solver driven by a CPU
torch copy of the network.  Shapes cover one and two points per thread, a
resolution that is not a multiple of the team, widths 32 and 48, kernel
sizes 5 and 7, 2..4 layers, Euler and SSP-RK3.  Tolerances: float64 1e-11
relative; float32 1e-5.
"""
import copy
import math

import pytest
import torch

from inc_amd import BGSimParams, BurgersSolverTorch, simulate_rollout
from inc_amd.correction import CorrectionCNN
from oracle import BurgersOracle, burgers_params
from oracle.rollout import simulate_rollout as oracle_rollout
from tests.conftest import rel_l2
from tests.test_gpu_burgers import smooth

pytestmark = pytest.mark.gpu


def _net(depth, width, k, seed, dtype, device):
    gen = torch.Generator().manual_seed(seed)
    net = CorrectionCNN(depth, width, k, generator=gen)
    with torch.no_grad():
        for cv in net.convs:
            cv.bias.copy_(0.1 * torch.randn(cv.bias.shape, generator=gen))
    return net.to(device=device, dtype=dtype)


def _solver(n, b, scheme, dtype, nu=0.02):
    dx = 2 * math.pi / n
    p = BGSimParams(dt=0.3 * min(dx, dx * dx / nu), resolution=n, nu=nu, batch_size=b, time_scheme=scheme,
                    device="cuda")
    return BurgersSolverTorch(p), p


# (n, batch, width, kernel, depth, scheme).  float32 runs the hidden layers
# on the matrix cores when there are any and their weights fit in LDS: four
# 16-position tiles per wavefront (n = 128, 100, 64), two points per thread
# (n = 256 at a batch that fills the SIMDs); depth 2 (no hidden layer) and
# float64 run the vector-ALU network.
CASES = [(128, 6, 32, 5, 4, "SSPRK3"), (100, 3, 48, 7, 3, "SSPRK3"), (512, 2, 32, 5, 2, "Euler"),
         (64, 5, 48, 5, 4, "Euler"), (256, 2048, 32, 5, 3, "SSPRK3")]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n{}_b{}_c{}_k{}_d{}_{}".format(*c))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_inc_rollout_in_kernel_matches_per_step_path(cuda, case, dtype):
    n, b, width, k, depth, scheme = case
    if dtype == torch.float64 and n > 256:
        pytest.skip("the in-kernel path runs float64 up to n = 256")
    H = 7 if b < 1000 else 3
    gen = torch.Generator().manual_seed(n + width)
    u0 = smooth(b, n, gen).to(cuda, dtype)
    net = _net(depth, width, k, 11 + n, dtype, cuda)
    sv1, _ = _solver(n, b, scheme, dtype)
    with torch.no_grad():
        res1, corr1 = simulate_rollout(net, sv1, u0, H, "INC")
    # the per-step path: gradient enabled on the initial state
    sv2, _ = _solver(n, b, scheme, dtype)
    res2, corr2 = simulate_rollout(net, sv2, u0.clone().requires_grad_(True), H, "INC", fused=False)
    tol = 1e-11 if dtype == torch.float64 else 1e-5
    assert res1.shape == (b, H, n) and corr1.shape == (b, H, n)
    assert rel_l2(res1, res2.detach()) < tol
    assert rel_l2(corr1, corr2.detach()) < tol
    assert abs(sv1.t - sv2.t) < 1e-12


@pytest.mark.parametrize("scale", [1.0, 0.1])
def test_inc_rollout_in_kernel_fp64_matches_oracle(cuda, scale):
    # scale: the network's output scale (CorrectionCNN(scale=...)), applied in the kernel
    n, b, H = 96, 3, 6
    gen = torch.Generator().manual_seed(5)
    u0 = smooth(b, n, gen)
    net = _net(4, 32, 5, 17, torch.float64, cuda)
    net.scale = scale
    cpu_net = copy.deepcopy(net).cpu()
    for cv in cpu_net.convs:
        cv.backend = "torch"
    sv, p = _solver(n, b, "SSPRK3", torch.float64)
    oracle = BurgersOracle(burgers_params(n, p.dt, nu=p.nu, batch_size=b, time_scheme="SSPRK3"))
    with torch.no_grad():
        res, corr = simulate_rollout(net, sv, u0.to(cuda), H, "INC")
        ref, ref_corr = oracle_rollout(cpu_net, oracle, u0, H, "INC")
    assert rel_l2(res, ref) < 1e-11
    assert rel_l2(corr, ref_corr) < 1e-11


def test_inc_rollout_shapes_outside_the_kernel_take_the_per_step_path(cuda):
    # width 24 is not one the in-kernel network runs: same results through the per-step path
    n, b, H = 64, 2, 4
    gen = torch.Generator().manual_seed(9)
    u0 = smooth(b, n, gen).to(cuda, torch.float64)
    net = _net(3, 24, 5, 3, torch.float64, cuda)
    assert net.fused_desc(torch.float64) is None
    sv1, _ = _solver(n, b, "SSPRK3", torch.float64)
    with torch.no_grad():
        res1, _ = simulate_rollout(net, sv1, u0, H, "INC")
    sv2, _ = _solver(n, b, "SSPRK3", torch.float64)
    res2, _ = simulate_rollout(net, sv2, u0.clone().requires_grad_(True), H, "INC", fused=False)
    assert rel_l2(res1, res2.detach()) < 1e-12


@pytest.mark.parametrize("team", [128, 256])
def test_inc_rollout_in_kernel_explicit_team(cuda, team):
    # an explicit team (sim_params.team) changes the block shape and the matrix-core tiling (one
    # wavefront holds 16 * PW positions): same rollout as the automatic choice, with and without gradient
    n, b, H = 128, 4, 5
    gen = torch.Generator().manual_seed(31)
    u0 = smooth(b, n, gen).to(cuda, torch.float32)
    net = _net(4, 32, 5, 7, torch.float32, cuda)
    outs = []
    for t in (0, team):
        sv, p = _solver(n, b, "SSPRK3", torch.float32)
        sv.team = t
        with torch.no_grad():
            res, corr = simulate_rollout(net, sv, u0, H, "INC")
        sv2, _ = _solver(n, b, "SSPRK3", torch.float32)
        sv2.team = t
        u = u0.clone().requires_grad_(True)
        net.zero_grad(set_to_none=True)
        r2, c2 = simulate_rollout(net, sv2, u, H, "INC")
        (r2.square().mean() + c2.mean()).backward()
        outs.append([res, corr, r2.detach(), u.grad] + [q.grad.clone() for q in net.parameters()])
    for i, (a, c) in enumerate(zip(*outs)):
        assert rel_l2(a, c) < 1e-6, i
