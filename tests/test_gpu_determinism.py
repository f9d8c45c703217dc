"""Deterministic replay: every path of the library computed twice from the
same inputs gives bitwise-equal outputs and gradients.  The kernels use no
atomics; reductions (the weight gradients' per-block partial sums, the
loss partials, the team reductions) run in a fixed order whatever the
launch shape's scheduling, so a rerun cannot differ in the last bit.  The
cases cover multi-job blocks (matrix-core convolutions and their
asynchronous weight-gradient tiles over many jobs per block), ragged row
tiles, packed FFT pairs and the mixed-radix transform."""
import math

import pytest
import torch

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch, KSSimParams, KSSolverTorch, RaggedBurgers
from inc_amd.batched import stable_dt
from inc_amd.correction import CorrectionCNN, UNet1D, _conv_bwd_data, _conv_bwd_weight, _conv_fwd
import inc_amd._lib as L

pytestmark = pytest.mark.gpu


def _flat(xs):
    out = []
    for x in xs:
        if isinstance(x, (tuple, list)):
            out += _flat(x)
        elif x is not None:
            out.append(x)
    return out


def twice(fn):
    a, b = _flat(fn()), _flat(fn())
    assert len(a) == len(b) and len(a) > 0
    for i, (x, y) in enumerate(zip(a, b)):
        assert torch.equal(x, y), f"rerun differs in result {i}"
    return a


def test_convolutions_rerun_bitwise(cuda):
    for (bsz, c, n, k) in ((256, 48, 512, 7), (512, 32, 256, 5), (64, 96, 128, 3)):
        gen = torch.Generator().manual_seed(bsz + c)
        x = torch.randn(bsz, c, n, generator=gen).to(cuda)
        w = (0.1 * torch.randn(c, c, k, generator=gen)).to(cuda)
        b = torch.randn(c, generator=gen).to(cuda)
        g = torch.randn(bsz, c, n, generator=gen).to(cuda)
        m = torch.rand(bsz, c, n, generator=gen).to(cuda)
        twice(lambda: [_conv_fwd(x, w, b, act=L.INC_ACT_GELU), _conv_bwd_data(g, w, c, m),
                       *_conv_bwd_weight(g, x, w, True)])


def _burgers_run(cuda, scheme):
    b, n = 64, 256
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u0 = (torch.sin(x)[None] * torch.linspace(0.5, 1.5, b, dtype=torch.float64)[:, None]).float().to(cuda)
    w = torch.randn(b, 20, n, generator=torch.Generator().manual_seed(3)).to(cuda)
    sv = BurgersSolverTorch(BGSimParams(dt=0.2 * 2 * math.pi / n, resolution=n, nu=0.01, batch_size=b,
                                        time_scheme=scheme, device="cuda"))

    def run():
        u = u0.clone().requires_grad_(True)
        traj = sv.rollout(u, 20)
        (traj * w).sum().backward()
        return [traj.detach(), u.grad]
    return run


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK3"])
def test_burgers_rollout_rerun_bitwise(cuda, scheme):
    twice(_burgers_run(cuda, scheme))


@pytest.mark.parametrize("n", [256, 384])
def test_ks_rollout_rerun_bitwise(cuda, n):
    # 256: register transform, packed pairs; 384: mixed-radix transform
    b, H, Ln = 9, 12, 64.0 * n / 256
    gen = torch.Generator().manual_seed(n)
    u0 = (0.3 * torch.randn(b, n, generator=gen)).to(cuda)
    w = torch.randn(b, H, n, generator=gen).to(cuda)
    sv = KSSolverTorch(KSSimParams([Ln], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    Ld = torch.tensor([Ln], device=cuda)

    def run():
        u = u0.clone().requires_grad_(True)
        traj = sv.rollout(u, H, Ld)
        (traj * w).sum().backward()
        return [traj.detach(), u.grad]
    twice(run)


def test_burgers_inc_training_rollout_rerun_bitwise(cuda):
    # the network inside the rollout and adjoint kernels, loss fused
    b, n, H = 64, 128, 8
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u0 = (torch.sin(x)[None] * torch.linspace(0.5, 1.5, b, dtype=torch.float64)[:, None]).float().to(cuda)
    target = (0.9 * u0[:, None, :]).expand(b, H, n).contiguous()
    sv = BurgersSolverTorch(BGSimParams(dt=0.2 * 2 * math.pi / n, resolution=n, nu=0.01, batch_size=b,
                                        time_scheme="SSPRK3", device="cuda"))
    net = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(1), scale=0.01).to(cuda)

    def run():
        net.zero_grad(set_to_none=True)
        u = u0.clone().requires_grad_(True)
        traj, _ = inc_amd.simulate_rollout(net, sv, u, H, "INC")
        ((traj - target) ** 2).mean().backward()
        return [traj.detach(), u.grad] + [p.grad.clone() for p in net.parameters()]
    twice(run)


def test_ks_inc_one_call_rerun_bitwise(cuda):
    b, n, H, Ln = 16, 512, 6, 100.0
    gen = torch.Generator().manual_seed(7)
    u0 = (0.3 * torch.randn(b, n, generator=gen)).to(cuda)
    target = (0.3 * torch.randn(b, H, n, generator=gen)).to(cuda)
    sv = KSSolverTorch(KSSimParams([Ln], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    net = CorrectionCNN(4, 48, 7, generator=torch.Generator().manual_seed(2), scale=0.01).to(cuda)
    Ld = torch.tensor([Ln], device=cuda)

    def run():
        net.zero_grad(set_to_none=True)
        u = u0.clone().requires_grad_(True)
        traj, _ = inc_amd.simulate_rollout(net, sv, u, H, "INC", L=Ld)
        ((traj - target) ** 2).mean().backward()
        return [traj.detach(), u.grad] + [p.grad.clone() for p in net.parameters()]
    twice(run)


def test_ragged_inc_one_call_rerun_bitwise(cuda):
    # uneven batch: ragged row tiles, trajectories leaving at their horizons
    gen = torch.Generator().manual_seed(11)
    count, steps = 40, 10
    ns = [[64, 128, 256, 512, 1024][int(i)] for i in torch.randint(0, 5, (count,), generator=gen)]
    nus = torch.exp(torch.empty(count, dtype=torch.float64).uniform_(math.log(0.003), math.log(0.03),
                                                                      generator=gen)).tolist()
    dts = [stable_dt(n, nu) for n, nu in zip(ns, nus)]
    hz = torch.randint(2, steps + 1, (count,), generator=gen).tolist()
    rb = RaggedBurgers(ns, nus, dts, horizon=hz)
    fields = []
    for i, n in enumerate(ns):
        x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
        fields.append(torch.sin(x + 0.1 * i))
    u0 = rb.pack(fields)
    net = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(4), scale=0.01).to(cuda)
    mask = rb.valid_mask(steps)

    def run():
        net.zero_grad(set_to_none=True)
        u = u0.clone().requires_grad_(True)
        traj = rb.rollout(u, steps, model=net)
        ((traj ** 2) * mask).sum().backward()
        return [traj.detach(), u.grad] + [p.grad.clone() for p in net.parameters()]
    twice(run)


def test_unet1d_rerun_bitwise(cuda):
    torch.manual_seed(5)
    net = UNet1D().to(cuda)
    gen = torch.Generator().manual_seed(6)
    x = torch.randn(128, 128, generator=gen).to(cuda)
    g = torch.randn(128, 128, generator=gen).to(cuda)

    def run():
        net.zero_grad(set_to_none=True)
        xx = x.clone().requires_grad_(True)
        y = net(xx)
        (y * g).sum().backward()
        return [y.detach(), xx.grad] + [p.grad.clone() for p in net.parameters()]
    twice(run)
