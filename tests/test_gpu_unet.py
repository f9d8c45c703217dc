"""UNet1D, the reference's default correction network (models/models_1d/unet.py:5-53,
scripts/Train_1D.py:82), on this library's kernels: the 5-tap convolutions with
ReLU fused, the 2-to-1 max pooling, and the kernel-2 stride-2 transposed
convolution with its skip concatenation.

Against torch itself: max pooling and the upsampling concatenation bit for bit
(values and gradients, ties and NaNs included) against torch.nn.MaxPool1d and
torch.cat(ConvTranspose1d, skip) on the GPU; the whole network against the same
module run by torch (backend='torch') in float64 and float32, forward and the
gradients of the state and of every parameter; and as the INC correction of a
Burgers rollout, against the FP64 oracle driven by a CPU float64 copy of the
network.

Bars: float64 1e-12 relative; float32 1e-5 (outputs) and 1e-4 (gradients).
At B = 512 float32 gradients are held per layer (test_unet_at_batch_512 says
why).
"""
import math

import pytest
import torch
from torch import nn

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch, UNet1D
from inc_amd.correction import _conv_act, max_pool2, upconv2_cat
from oracle import BurgersOracle, burgers_params
from oracle.rollout import simulate_rollout as oracle_rollout
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _fields(b, n, seed):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, 7):
        u += torch.randn(b, 1, generator=gen, dtype=torch.float64) / k * torch.sin(
            k * x + 2 * math.pi * torch.rand(b, 1, generator=gen, dtype=torch.float64))
    return u


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_max_pool_matches_torch_bitwise_with_ties_and_nans(cuda, dtype):
    gen = torch.Generator().manual_seed(3)
    x = torch.randint(-3, 4, (5, 7, 96), generator=gen).to(dtype)  # many exact ties
    x[0, 0, 0], x[0, 0, 3], x[0, 0, 4], x[0, 0, 5] = float("nan"), float("nan"), float("nan"), float("nan")
    x[1, 2, 10] = float("inf")
    x[1, 2, 13] = float("-inf")
    x = x.to(cuda)
    g = torch.randn(5, 7, 48, generator=gen, dtype=torch.float64).to(cuda, dtype)
    outs = []
    for f in (max_pool2, nn.MaxPool1d(2, 2)):
        xi = x.clone().requires_grad_(True)
        y = f(xi)
        y.backward(g)
        outs.append((y.detach(), xi.grad))
    (y0, g0), (y1, g1) = outs
    assert torch.equal(y0.isnan(), y1.isnan())
    assert torch.equal(torch.nan_to_num(y0), torch.nan_to_num(y1))
    assert torch.equal(g0, g1)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_upconv_cat_matches_conv_transpose_and_cat(cuda, dtype):
    torch.manual_seed(4)
    up = nn.ConvTranspose1d(24, 12, kernel_size=2, stride=2).to(cuda, dtype)
    x0 = torch.randn(6, 24, 40, dtype=torch.float64).to(cuda, dtype)
    s0 = torch.randn(6, 9, 80, dtype=torch.float64).to(cuda, dtype)
    g = torch.randn(6, 21, 80, dtype=torch.float64).to(cuda, dtype)
    res = []
    for f in (lambda x, s: upconv2_cat(x, up, s), lambda x, s: torch.cat((up(x), s), 1)):
        up.zero_grad()
        x, s = x0.clone().requires_grad_(True), s0.clone().requires_grad_(True)
        y = f(x, s)
        y.backward(g)
        res.append([y.detach(), x.grad, s.grad, up.weight.grad.clone(), up.bias.grad.clone()])
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    for a, b in zip(*res):
        assert rel_l2(a, b) < tol
    assert torch.equal(res[0][2], res[1][2])  # the skip gradient is a copy


def _pair(dtype, seed=0, **kw):
    torch.manual_seed(seed)
    hip = UNet1D(**kw).to("cuda", dtype)
    ref = UNet1D(backend="torch", **kw).to("cuda", dtype)
    ref.load_state_dict(hip.state_dict())
    return hip, ref


@pytest.mark.parametrize("dtype,tol_f,tol_g", [(torch.float64, 1e-12, 1e-12), (torch.float32, 1e-5, 1e-4)])
@pytest.mark.parametrize("L", ["float", "tensor"])
@pytest.mark.parametrize("b,n,f", [(8, 128, 32), (3, 64, 8)])
def test_unet_hip_matches_torch(cuda, dtype, tol_f, tol_g, L, b, n, f):
    hip, ref = _pair(dtype, init_features=f)
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


@pytest.mark.parametrize("ci,co,n", [(2, 32, 128), (32, 32, 128), (32, 64, 64), (64, 64, 64), (64, 128, 32),
                                      (128, 128, 32), (128, 64, 64), (64, 32, 128)])
def test_unet_layers_at_batch_512_float32(cuda, ci, co, n):
    """Every 5-tap convolution shape of UNet1D at its default width, B = 512,
    float32 against torch in float64: several matrix-core tiles per block,
    and the layers of more than 64 channels (weight gradients in channel
    blocks).  Linear in each operand, so the bars hold without the near-tie
    caveat of the whole network."""
    torch.manual_seed(ci * 1000 + co)
    conv = nn.Conv1d(ci, co, 5, padding=2, padding_mode="circular").to(cuda, torch.float64)
    x0 = torch.randn(512, ci, n, dtype=torch.float64, device=cuda)
    g = torch.randn(512, co, n, dtype=torch.float64, device=cuda)
    res = []
    for hip, dt in ((True, torch.float32), (False, torch.float64)):
        m = nn.Conv1d(ci, co, 5, padding=2, padding_mode="circular").to(cuda, dt)
        m.load_state_dict(conv.state_dict())
        x = x0.to(dt).clone().requires_grad_(True)
        y = _conv_act(x, m, inc_amd._lib.INC_ACT_NONE) if hip else m(x)
        y.backward(g.to(dt))
        res.append([y.detach(), x.grad, m.weight.grad, m.bias.grad])
    assert rel_l2(res[0][0], res[1][0]) < 1e-5
    for a, c in zip(res[0][1:], res[1][1:]):
        assert rel_l2(a, c) < 1e-5


@pytest.mark.parametrize("ci,co,n", [(128, 64, 32), (64, 32, 64)])
def test_unet_upconv_at_batch_512_float32(cuda, ci, co, n):
    torch.manual_seed(ci + co)
    up = nn.ConvTranspose1d(ci, co, kernel_size=2, stride=2).to(cuda, torch.float64)
    x0 = torch.randn(512, ci, n, dtype=torch.float64, device=cuda)
    s0 = torch.randn(512, co, 2 * n, dtype=torch.float64, device=cuda)
    g = torch.randn(512, 2 * co, 2 * n, dtype=torch.float64, device=cuda)
    res = []
    for hip, dt in ((True, torch.float32), (False, torch.float64)):
        m = nn.ConvTranspose1d(ci, co, kernel_size=2, stride=2).to(cuda, dt)
        m.load_state_dict(up.state_dict())
        x, sk = x0.to(dt).clone().requires_grad_(True), s0.to(dt).clone().requires_grad_(True)
        y = upconv2_cat(x, m, sk) if hip else torch.cat((m(x), sk), 1)
        y.backward(g.to(dt))
        res.append([y.detach(), x.grad, sk.grad, m.weight.grad, m.bias.grad])
    for a, c in zip(*res):
        assert rel_l2(a, c) < 1e-5


def test_unet_at_batch_512(cuda):
    """The whole network at B = 512, N = 128, default width: float64 forward
    and every gradient against torch's float64 run (1e-12), and the float32
    output against it (1e-5).  Float32 gradients of the whole network are
    held at the small batches above and per layer here: at this size a few
    ReLU or max-pool near-ties are taken differently by float32 and float64
    (torch's own float32 run shows the same), which moves the gradient
    behind them by O(1)."""
    torch.manual_seed(6)
    hip = UNet1D().to(cuda, torch.float64)
    ref = UNet1D(backend="torch").to(cuda, torch.float64)
    ref.load_state_dict(hip.state_dict())
    hip32 = UNet1D().to(cuda)
    hip32.load_state_dict(hip.state_dict())
    u0 = _fields(512, 128, 9).to(cuda)
    w = torch.randn(512, 128, generator=torch.Generator().manual_seed(10), dtype=torch.float64).to(cuda)
    outs, grads = [], []
    for m in (hip, ref):
        u = u0.clone().requires_grad_(True)
        y = m(u, 2 * math.pi)
        (y * w).sum().backward()
        outs.append(y.detach())
        grads.append([u.grad] + [p.grad for p in m.parameters()])
    assert rel_l2(outs[0], outs[1]) < 1e-12
    for a, c in zip(*grads):
        assert rel_l2(a, c) < 1e-12
    with torch.no_grad():
        assert rel_l2(hip32(u0.float(), 2 * math.pi), outs[1]) < 1e-5


def test_unet_as_burgers_inc_correction_against_oracle(cuda):
    b, n, H = 4, 64, 6
    torch.manual_seed(5)
    cpu = UNet1D(L=2 * math.pi, init_features=8, backend="torch").double()
    with torch.no_grad():
        cpu.conv.weight.mul_(0.05)
    net = UNet1D(L=2 * math.pi, init_features=8).cuda().double()
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


@pytest.mark.parametrize("net_cls", ["UNet1D", "ResNet1D"])
def test_reference_nets_as_ks_inc_correction_against_oracle(cuda, net_cls):
    """The reference's KS training configuration (trainer_1d.py:361-378: ETD1,
    the double step, per-trajectory domain sizes passed to the network as L,
    64 points): the reference's correction networks on the library's kernels
    as the INC correction of a KS rollout, against the FP64 oracle driven by
    a CPU float64 copy of the network."""
    from inc_amd import KSSimParams, KSSolverTorch, ResNet1D
    from oracle import KSOracle, ks_params
    cls = {"UNet1D": UNet1D, "ResNet1D": ResNet1D}[net_cls]
    b, n, H = 4, 64, 5
    Ls = [22.0, 30.0, 40.0, 64.0]
    torch.manual_seed(12)
    cpu = cls(init_features=8, backend="torch").double()
    with torch.no_grad():
        (cpu.conv if net_cls == "UNet1D" else cpu.final).weight.mul_(0.05)
    net = cls(init_features=8).to(cuda, torch.float64)
    net.load_state_dict(cpu.state_dict())
    gen = torch.Generator().manual_seed(13)
    u0 = torch.zeros(b, n, dtype=torch.float64)
    for i, L in enumerate(Ls):
        x = torch.arange(n, dtype=torch.float64) * (L / n)
        for k in range(1, 5):
            u0[i] += 0.5 * torch.randn(1, generator=gen, dtype=torch.float64) * torch.cos(
                2 * math.pi * k * x / L + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
    target = 0.3 * torch.randn(b, H, n, generator=torch.Generator().manual_seed(14), dtype=torch.float64)
    so = KSOracle(ks_params(Ls, resolution=n, dt=0.1, time_scheme="ETD1"))
    uo = u0.clone().requires_grad_(True)
    ref, ref_c = oracle_rollout(cpu, so, uo, H, "INC", L=so.domain_size)
    ((ref - target) ** 2).mean().backward()
    sv = KSSolverTorch(KSSimParams(Ls, resolution=n, dt=0.1, time_scheme="ETD1", device="cuda"))
    Ld = torch.tensor(Ls, dtype=torch.float64, device=cuda)
    u = u0.to(cuda).requires_grad_(True)
    traj, corr = inc_amd.simulate_rollout(net, sv, u, H, "INC", L=Ld)
    ((traj - target.to(cuda)) ** 2).mean().backward()
    assert rel_l2(traj.detach(), ref.detach()) < 1e-11 and rel_l2(corr.detach(), ref_c.detach()) < 1e-11
    assert rel_l2(u.grad, uo.grad) < 1e-10
    for p, q in zip(net.parameters(), cpu.parameters()):
        assert rel_l2(p.grad, q.grad) < 1e-10
