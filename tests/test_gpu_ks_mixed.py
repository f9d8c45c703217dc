"""KS at resolutions that are not powers of two in [16, 1024] (the reference
takes any N through torch.fft, solvers/solver_1d.py:208-210): the
mixed-radix transform through LDS (and, in float32 at P * 2^k, the radix-P
register step before power-of-two register transforms) against the FP64
oracle, per scheme, for
odd, prime, mixed-radix and small n, packed pairs and unpaired or mixed-L
batches, the fused rollout + MSE, the INC rollout with a correction
network, and resolution independence against the power-of-two kernels.

Bars: float64 1e-11 (states) / 1e-10 (gradients) relative L2 against the
oracle, float32 1e-5 / 1e-4.
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
from tests.test_gpu_ks import gpu_solver, modes

pytestmark = pytest.mark.gpu

# radix plans: 8 = 4*2 and 12 = 4*3 below the register transform's range;
# 75 = 3*5*5, 96 = 4*4*2*3, 100 = 4*5*5, 127 (prime above 31), 243 = 3^5,
# 250 = 2*5^3, 384 = 4^3*2*3 (four-wavefront team), 1000 = 4*2*5^3,
# 1022 = 2*7*73, 77 = 7*11 (primes without a butterfly)
SIZES = [8, 12, 75, 77, 96, 100, 127, 243, 250, 384, 1000, 1022]


def _case(n, scheme, Ls, dtype, seed, H=3):
    b, dt = len(Ls), 0.1
    gen = torch.Generator().manual_seed(seed)
    u = modes(b, n, Ls[0], gen, kmax=max(1, min(10, n // 4)))
    s = 0.1 * torch.randn(b, n, generator=gen, dtype=torch.float64)
    w = torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    so = KSOracle(ks_params(Ls, resolution=n, dt=dt, time_scheme=scheme))
    so.doublestep = False
    uo = u.clone().requires_grad_(True)
    sso = s.clone().requires_grad_(True)
    so.correction_s = sso
    cur, outs = uo, []
    for _ in range(H):
        cur = so.step(cur, so.domain_size)
        outs.append(cur)
    ref = torch.stack(outs, 1)
    (ref * w).sum().backward()
    sv = gpu_solver(Ls, n, dt, scheme, dtype=dtype)
    sv.doublestep = False
    ug = u.to("cuda", dtype).requires_grad_(True)
    sg = s.to("cuda", dtype).requires_grad_(True)
    sv.correction_s = sg
    traj = sv.rollout(ug, H, torch.tensor(Ls, device="cuda"))
    (traj * w.to("cuda", dtype)).sum().backward()
    return (traj, ref.detach()), (ug.grad, uo.grad), (sg.grad, sso.grad)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("scheme", ["ETD1", "ETRK2", "ETDRK4"])
def test_mixed_radix_rollout_and_gradient_fp64(cuda, n, scheme):
    # packed pairs with an unpaired last trajectory, and per-trajectory L
    # (unpacked); 3-step rollouts with a correction s
    for Ls in ([22.0 * max(n, 16) / 64] * 3, [22.0 * max(n, 16) / 64, 30.0 * max(n, 16) / 64]):
        st, gu, gs = _case(n, scheme, Ls, torch.float64, seed=n)
        assert rel_l2(*st) < 1e-11, Ls
        assert rel_l2(*gu) < 1e-10, Ls
        assert rel_l2(*gs) < 1e-10, Ls


@pytest.mark.parametrize("n", [12, 96, 127, 384, 1000])
def test_mixed_radix_fp32(cuda, n):
    st, gu, gs = _case(n, "ETDRK4", [22.0 * max(n, 16) / 64] * 5, torch.float32, seed=7 * n, H=1)
    assert rel_l2(*st) < 1e-5
    assert rel_l2(*gu) < 1e-4
    assert rel_l2(*gs) < 1e-4


# float32 sizes P * 2^k (P = 3, 5; 2^k in [32, 256]) run a radix-P step in
# registers before P power-of-two register transforms
REG_MIXED = [96, 192, 384, 768, 160, 320, 640]


@pytest.mark.parametrize("n", REG_MIXED)
@pytest.mark.parametrize("scheme", ["ETD1", "ETRK2", "ETDRK4"])
def test_register_mixed_radix_fp32(cuda, n, scheme):
    # packed pairs with an unpaired last trajectory, and per-trajectory L
    for Ls in ([22.0 * n / 64] * 5, [22.0 * n / 64, 30.0 * n / 64, 40.0 * n / 64]):
        st, gu, gs = _case(n, scheme, Ls, torch.float32, seed=3 * n, H=3)
        assert rel_l2(*st) < 1e-5, Ls
        assert rel_l2(*gu) < 1e-4, Ls
        assert rel_l2(*gs) < 1e-4, Ls


def test_register_mixed_radix_fused_mse_and_resolution_independence(cuda):
    # the loss-fused kernels on the register path, and a resolved state at
    # n = 96, 192, 384 against the power-of-two kernel at 512 (float32)
    b, n, L, H = 5, 192, 40.0, 6
    gen = torch.Generator().manual_seed(23)
    u = modes(b, n, L, gen)
    target = 0.3 * torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.1, time_scheme="ETDRK4"))
    so.doublestep = False
    uo = u.clone().requires_grad_(True)
    cur, outs = uo, []
    for _ in range(H):
        cur = so.step(cur, so.domain_size)
        outs.append(cur)
    ref = torch.stack(outs, 1)
    ref_loss = ((ref - target) ** 2).mean()
    ref_loss.backward()
    sv = gpu_solver([L], n, 0.1, "ETDRK4", dtype=torch.float32)
    sv.doublestep = False
    ug = u.float().to(cuda).requires_grad_(True)
    loss, traj = sv.rollout_mse(ug, target.float().to(cuda), torch.tensor([L], device=cuda))
    loss.backward()
    assert abs(float(loss) - float(ref_loss)) < 1e-5 * float(ref_loss)
    assert rel_l2(traj, ref.detach()) < 1e-5
    assert rel_l2(ug.grad, uo.grad) < 1e-4
    Lr, dt, steps = 22.0, 0.1, 10

    def run(m):
        x = torch.arange(m, dtype=torch.float64) * (Lr / m)
        u0 = (0.5 * torch.cos(2 * math.pi * x / Lr) + 0.25 * torch.sin(4 * math.pi * x / Lr + 0.3))[None]
        svm = gpu_solver([Lr], m, dt, "ETDRK4", dtype=torch.float32)
        svm.doublestep = False
        return svm.rollout(u0.float().to(cuda), steps, torch.tensor([Lr], device=cuda))[0, -1].double().cpu()

    def resample(v, m):
        k = v.shape[-1]
        V = torch.fft.fft(v)
        h = (min(k, m) - 1) // 2
        out = torch.zeros(m, dtype=torch.complex128)
        out[:h + 1] = V[:h + 1]
        out[m - h:] = V[k - h:]
        return torch.fft.ifft(out).real * (m / k)

    ref512 = run(512)
    for m in (96, 192, 384):
        assert rel_l2(resample(run(m), 512), ref512) < 1e-6, m


def test_mixed_radix_double_step_and_correction_u(cuda):
    # the reference's etd1/etrk2 wrappers: double step and correction_u
    n, L, dt = 96, 22.0 * 96 / 64, 0.1
    gen = torch.Generator().manual_seed(5)
    u = modes(4, n, L, gen)
    cu = 0.05 * torch.randn(4, n, generator=gen, dtype=torch.float64)
    for scheme in ("ETD1", "ETRK2"):
        so = KSOracle(ks_params([L] * 4, resolution=n, dt=dt, time_scheme=scheme))
        so.correction_u = cu
        sv = gpu_solver([L], n, dt, scheme)
        sv.correction_u = cu.to(cuda)
        out = sv.step(u.to(cuda), torch.tensor([L], device=cuda))
        assert rel_l2(out, so.step(u, so.domain_size)) < 1e-11, scheme


def test_mixed_radix_rollout_mse_matches_oracle(cuda):
    # the loss-fused kernels (forward partial sums, adjoint seeded in-kernel)
    b, n, L, H = 5, 150, 40.0, 6
    gen = torch.Generator().manual_seed(21)
    u = modes(b, n, L, gen)
    target = 0.3 * torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.1, time_scheme="ETDRK4"))
    so.doublestep = False
    uo = u.clone().requires_grad_(True)
    cur, outs = uo, []
    for _ in range(H):
        cur = so.step(cur, so.domain_size)
        outs.append(cur)
    ref = torch.stack(outs, 1)
    ref_loss = ((ref - target) ** 2).mean()
    ref_loss.backward()
    sv = gpu_solver([L], n, 0.1, "ETDRK4")
    sv.doublestep = False
    ug = u.to(cuda).requires_grad_(True)
    loss, traj = sv.rollout_mse(ug, target.to(cuda), torch.tensor([L], device=cuda))
    loss.backward()
    assert abs(float(loss) - float(ref_loss)) < 1e-12 * float(ref_loss)
    assert rel_l2(traj, ref.detach()) < 1e-11
    assert rel_l2(ug.grad, uo.grad) < 1e-10


@pytest.mark.parametrize("n,dtype", [(96, torch.float64), (192, torch.float32)])
def test_mixed_radix_inc_rollout_against_oracle(cuda, n, dtype):
    # the KS INC rollout in one library call with a CorrectionCNN, against
    # the oracle with a CPU float64 copy of the network: float64 at n = 96
    # (transform through LDS), float32 at n = 192 (radix-3 register step)
    b, L, H = 4, 40.0, 8
    gen = torch.Generator().manual_seed(31)
    u0 = modes(b, n, L, gen)
    target = 0.3 * torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    cnn = CorrectionCNN(3, 16, 5, generator=torch.Generator().manual_seed(3), backend="torch", scale=0.01).double()
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
    so.doublestep = False
    uo = u0.clone().requires_grad_(True)
    ref, _ = oracle_rollout(lambda v, L_: cnn(v), so, uo, H, "INC", L=so.domain_size)
    ((ref - target) ** 2).mean().backward()
    net = CorrectionCNN(3, 16, 5, scale=0.01).to(cuda, dtype)
    net.load_state_dict(cnn.state_dict())
    sv = KSSolverTorch(KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    u = u0.to(cuda, dtype).requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(net, sv, u, H, "INC", L=torch.tensor([L], dtype=torch.float64, device=cuda))
    ((traj - target.to(cuda, dtype)) ** 2).mean().backward()
    ftol, gtol = (1e-11, 1e-10) if dtype == torch.float64 else (1e-5, 1e-4)
    assert rel_l2(traj.detach(), ref.detach()) < ftol
    assert rel_l2(u.grad, uo.grad) < gtol
    for p, q in zip(net.parameters(), cnn.parameters()):
        assert rel_l2(p.grad, q.grad) < gtol


def test_resolution_independence_against_the_power_of_two_kernels(cuda):
    # a smooth, well-resolved state stepped at n = 96, 127 and 150 (mixed
    # radix) and at 256 (register transform), interpolated to 256 points,
    # agrees to rounding: the two transforms give the same physics
    L, dt, steps = 22.0, 0.1, 10

    def run(m):
        x = torch.arange(m, dtype=torch.float64) * (L / m)
        u = (0.5 * torch.cos(2 * math.pi * x / L) + 0.25 * torch.sin(4 * math.pi * x / L + 0.3))[None]
        sv = gpu_solver([L], m, dt, "ETDRK4")
        sv.doublestep = False
        return sv.rollout(u.to(cuda), steps, torch.tensor([L], device=cuda))[0, -1].cpu()

    def resample(u, m):
        n = u.shape[-1]
        U = torch.fft.fft(u)
        h = (min(n, m) - 1) // 2
        out = torch.zeros(m, dtype=torch.complex128)
        out[:h + 1] = U[:h + 1]
        out[m - h:] = U[n - h:]
        return torch.fft.ifft(out).real * (m / n)

    ref = run(256)
    for n in (96, 127, 150):
        assert rel_l2(resample(run(n), 256), ref) < 1e-13, n


def test_resolution_limits_raise(cuda):
    for n in (3, 1025, 2048):
        sv = gpu_solver([22.0], n, 0.1, "ETDRK4")
        with pytest.raises(ValueError, match="resolution"):
            sv.step(torch.zeros(2, n, dtype=torch.float64, device=cuda), torch.tensor([22.0], device=cuda))
