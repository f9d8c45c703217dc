"""GPU parity of the Burgers / WENO HIP kernels against the FP64 oracle and
the reference golden vectors.

Tolerances: float64 kernels agree with the oracle to rounding (1e-11
relative); float32 forward states within 1e-5 relative L2 per step and
float32 gradients within 1e-4 relative L2 of the FP64 oracle.
"""
import math

import numpy as np
import pytest
import torch

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch
from oracle import BurgersOracle, burgers_params
from oracle import weno_reconstruction as weno_ref
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
F64_TOL = 1e-11


def smooth(b, n, gen, kmax=8, L=2 * math.pi):
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        a = torch.randn(b, 1, generator=gen, dtype=torch.float64) / k
        ph = torch.rand(b, 1, generator=gen, dtype=torch.float64) * 2 * math.pi
        u += a * torch.sin(k * x + ph)
    return u / u.abs().amax(-1, keepdim=True)


def make_pair(n, b, dt=0.01, nu=0.01, scheme="Euler", conv="WENO", eq="homogeneous", forcing=None, beta=0.0):
    """(oracle solver, GPU solver factory(dtype)) with identical parameters."""
    kw = {}
    if forcing is not None:
        kw = dict(J=forcing[0].shape[-1], A=forcing[0], omega=forcing[1], phi=forcing[2], l=forcing[3])
    po = burgers_params(n, dt, nu=nu, batch_size=b, convection_term=conv, equation_type=eq, time_scheme=scheme,
                        beta=beta, **kw)
    oracle = BurgersOracle(po)

    def gpu(dtype=torch.float32, team=0):
        nu_g = nu.to("cuda", dtype) if isinstance(nu, torch.Tensor) else nu
        p = BGSimParams(dt=dt, resolution=n, nu=nu_g, batch_size=b, convection_term=conv, equation_type=eq,
                        time_scheme=scheme, beta=beta, device="cuda")
        if forcing is not None:
            p.A, p.omega, p.phi, p.l = (f.to("cuda", dtype) for f in forcing)
        p.team = team
        return BurgersSolverTorch(p)
    return oracle, gpu


@pytest.mark.parametrize("name", ["rand", "smooth", "step"])
@pytest.mark.parametrize("direction", ["left", "right"])
def test_weno_golden_fp64(cuda, golden, name, direction):
    g = golden("weno")
    f = torch.as_tensor(g[f"f_{name}"]).to(cuda)
    out = inc_amd.weno_reconstruction(f, direction).cpu().numpy()
    np.testing.assert_allclose(out, g[f"{direction}_{name}"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("direction", ["left", "right"])
def test_weno_fp32_and_adjoint(cuda, direction):
    gen = torch.Generator().manual_seed(11)
    f = smooth(5, 96, gen) + 0.1 * torch.randn(5, 96, generator=gen, dtype=torch.float64)
    w = torch.randn(5, 96, generator=gen, dtype=torch.float64)
    out32 = inc_amd.weno_reconstruction(f.float().to(cuda), direction)
    assert rel_l2(out32, weno_ref(f, direction)) < 1e-6
    for dtype, tol in ((torch.float64, 1e-11), (torch.float32, 1e-4)):
        fg = f.to(cuda, dtype).requires_grad_(True)
        (inc_amd.weno_reconstruction(fg, direction) * w.to(cuda, dtype)).sum().backward()
        fo = f.clone().requires_grad_(True)
        (weno_ref(fo, direction) * w).sum().backward()
        assert rel_l2(fg.grad, fo.grad) < tol, dtype


def test_burgers_pieces_golden_fp64(cuda, golden):
    g = golden("burgers")
    u = torch.as_tensor(g["u"])
    b, n = u.shape
    nu = torch.as_tensor(g["nu"])
    s = torch.as_tensor(g["s"])
    for conv, key in (("WENO", "nonlinear_weno"), ("FDM", "nonlinear_fdm")):
        p = BGSimParams(dt=float(g["dt"]), resolution=n, L=float(g["L"]), nu=nu.to(cuda), batch_size=b,
                        convection_term=conv, device="cuda")
        sv = BurgersSolverTorch(p)
        uc = u.to(cuda)
        np.testing.assert_allclose(sv.calc_nonlinear(uc).cpu(), g[key], rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(sv.calc_linear(uc).cpu(), g["linear"], rtol=1e-11, atol=1e-8)
        sv.correction_s = s.to(cuda)
        np.testing.assert_allclose(sv.step(uc).cpu(), g["euler_weno_s" if conv == "WENO" else "euler_fdm_s"],
                                   rtol=1e-12, atol=1e-12)
    # 'learning' forcing with per-trajectory time
    p = BGSimParams(dt=float(g["dt"]), resolution=n, L=float(g["L"]), nu=nu.to(cuda), batch_size=b,
                    equation_type="learning", device="cuda")
    p.A, p.omega, p.phi, p.l = (torch.as_tensor(g[k]).to(cuda) for k in ("A", "omega", "phi", "l"))
    sv = BurgersSolverTorch(p)
    sv.t = torch.as_tensor(g["t"]).to(cuda)
    np.testing.assert_allclose(sv._compute_source(u.to(cuda)).cpu(), g["forcing"], rtol=1e-11, atol=1e-12)
    sv.correction_s = s.to(cuda)
    np.testing.assert_allclose(sv.step(u.to(cuda)).cpu(), g["euler_learning_s"], rtol=1e-12, atol=1e-12)


def test_burgers_gradient_golden_fp64(cuda, golden):
    g = golden("burgers")
    u = torch.as_tensor(g["u"])
    b, n = u.shape
    p = BGSimParams(dt=float(g["dt"]), resolution=n, L=float(g["L"]), nu=torch.as_tensor(g["nu"]).to(cuda),
                    batch_size=b, device="cuda")
    sv = BurgersSolverTorch(p)
    ss = torch.as_tensor(g["grad_ss"]).to(cuda).requires_grad_(True)
    u0 = u.to(cuda).requires_grad_(True)
    cur, traj = u0, []
    for k in range(ss.shape[1]):
        sv.correction_s = ss[:, k]
        cur = sv.step(cur)
        traj.append(cur)
    loss = ((torch.stack(traj, 1) - torch.as_tensor(g["grad_target"]).to(cuda)) ** 2).mean()
    loss.backward()
    np.testing.assert_allclose(u0.grad.cpu(), g["grad_u0"], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(ss.grad.cpu(), g["grad_s"], rtol=1e-10, atol=1e-13)


CASES = [
    # n, b, scheme, conv, eq, team
    (256, 3, "SSPRK3", "WENO", "homogeneous", 0),
    (128, 8, "SSPRK3", "WENO", "homogeneous", 0),
    (128, 5, "Euler", "WENO", "learning", 0),
    (64, 4, "SSPRK3", "FDM", "linear", 0),
    (1024, 2, "SSPRK3", "WENO", "homogeneous", 0),
    (100, 3, "SSPRK3", "WENO", "homogeneous", 0),  # ragged: n not a power of two
    (512, 2, "Euler", "WENO", "homogeneous", 256),  # several wavefronts per team
    (2048, 1, "SSPRK3", "WENO", "homogeneous", 0),  # 8 points per thread, one team per block
]


def _forcing(b, gen, J=2):
    A = torch.randn(b, 1, J, generator=gen, dtype=torch.float64) * 0.5
    om = torch.randn(b, 1, J, generator=gen, dtype=torch.float64)
    ph = torch.rand(b, 1, J, generator=gen, dtype=torch.float64) * 2 * math.pi
    ll = torch.randint(1, 4, (b, 1, J), generator=gen).to(torch.float64)
    return A, om, ph, ll


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_burgers_step_forward_and_gradient(cuda, case):
    n, b, scheme, conv, eq, team = case
    gen = torch.Generator().manual_seed(n + b)
    nu = torch.exp(torch.empty(b, 1, dtype=torch.float64).uniform_(math.log(0.003), math.log(0.03), generator=gen))
    dx = 2 * math.pi / n
    dt = min(0.4 * dx, 0.6 * dx * dx / float(nu.max())) * (0.5 if scheme == "Euler" else 1.0)
    forcing = _forcing(b, gen) if eq == "learning" else None
    oracle, gpu = make_pair(n, b, dt=dt, nu=nu, scheme=scheme, conv=conv, eq=eq, forcing=forcing, beta=-0.3)
    u = smooth(b, n, gen) + 0.02 * torch.randn(b, n, generator=gen, dtype=torch.float64)
    s = 0.2 * torch.randn(b, n, generator=gen, dtype=torch.float64)
    w = torch.randn(b, n, generator=gen, dtype=torch.float64)
    uo = u.clone().requires_grad_(True)
    so = s.clone().requires_grad_(True)
    oracle.t = 0.25
    oracle.correction_s = so
    ref = oracle.step(uo)
    (ref * w).sum().backward()
    for dtype in (torch.float64, torch.float32):
        sv = gpu(dtype, team)
        sv.t = 0.25
        ug = u.to(cuda, dtype).requires_grad_(True)
        sg = s.to(cuda, dtype).requires_grad_(True)
        sv.correction_s = sg
        out = sv.step(ug)
        (out * w.to(cuda, dtype)).sum().backward()
        ftol, gtol = (F64_TOL, 1e-10) if dtype == torch.float64 else (1e-5, 1e-4)
        assert rel_l2(out, ref.detach()) < ftol, ("forward", dtype)
        assert rel_l2(ug.grad, uo.grad) < gtol, ("grad u", dtype)
        assert rel_l2(sg.grad, so.grad) < gtol, ("grad s", dtype)
        assert abs(sv.t - 0.25 - dt) < 1e-15


def test_burgers_fused_rollout_matches_step_loop(cuda):
    n, b, H = 128, 6, 9
    gen = torch.Generator().manual_seed(5)
    u = (smooth(b, n, gen)).to(cuda, torch.float32)
    p = BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b, time_scheme="SSPRK3", device="cuda")
    w = torch.randn(b, H, n, generator=gen).to(cuda)
    s1 = BurgersSolverTorch(p)
    u1 = u.clone().requires_grad_(True)
    traj1, corr = inc_amd.simulate_rollout(None, s1, u1, H, None)
    assert corr is None and traj1.shape == (b, H, n)
    (traj1 * w).sum().backward()
    s2 = BurgersSolverTorch(p)
    u2 = u.clone().requires_grad_(True)
    cur, outs = u2, []
    for _ in range(H):
        cur = s2.step(cur)
        outs.append(cur)
    traj2 = torch.stack(outs, 1)
    (traj2 * w).sum().backward()
    assert rel_l2(traj1, traj2) < 1e-7
    assert rel_l2(u1.grad, u2.grad) < 1e-6
    assert abs(s1.t - s2.t) < 1e-12


def test_burgers_multistep_fp64_matches_oracle(cuda):
    n, b, H = 256, 2, 50
    gen = torch.Generator().manual_seed(8)
    u = smooth(b, n, gen)
    oracle, gpu = make_pair(n, b, dt=0.005, nu=0.01, scheme="SSPRK3")
    sv = gpu(torch.float64)
    traj = sv.rollout(u.to(cuda), H)
    cur = u
    for k in range(H):
        cur = oracle.step(cur)
        assert rel_l2(traj[:, k], cur) < 1e-11, k


def test_burgers_adaptive_cfl_matches_oracle(cuda):
    n, b = 64, 2
    gen = torch.Generator().manual_seed(2)
    u = 2.0 * smooth(b, n, gen)
    po = burgers_params(n, 0.05, nu=0.01, batch_size=b, adaptive_CFL=0.3)
    oracle = BurgersOracle(po)
    p = BGSimParams(dt=0.05, resolution=n, nu=0.01, batch_size=b, adaptive_CFL=0.3, device="cuda")
    sv = BurgersSolverTorch(p)
    out = sv.step(u.to(cuda))
    ref = oracle.step(u)
    assert rel_l2(out, ref) < 1e-11
    assert abs(sv.t - oracle.t) < 1e-14


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK3"])
def test_burgers_adaptive_per_trajectory_matches_oracle(cuda, scheme):
    # adaptive_mode='trajectory': every trajectory sub-steps by its own max|u|
    # on the GPU.  The oracle run on one trajectory at a time is the
    # reference's batch-wide sub-stepping restricted to that trajectory, so
    # it pins both the states and the gradients (sub-step sizes constant, as
    # with the reference's .item()).  Amplitudes 0.5..3 give different
    # sub-step counts to trajectories that share a block.
    n, b, dt, cfl, H = 64, 6, 0.05, 0.3, 3
    gen = torch.Generator().manual_seed(21)
    amps = torch.tensor([0.5, 1.0, 1.5, 2.0, 2.5, 3.0], dtype=torch.float64)[:, None]
    u0 = amps * smooth(b, n, gen)
    w = torch.randn(b, H, n, generator=gen, dtype=torch.float64)
    ref_traj, ref_grad = [], []
    for i in range(b):
        po = burgers_params(n, dt, nu=0.01, batch_size=1, adaptive_CFL=cfl, time_scheme=scheme)
        oracle = BurgersOracle(po)
        ui = u0[i:i + 1].clone().requires_grad_(True)
        cur, outs = ui, []
        for _ in range(H):
            cur = oracle.step(cur)
            outs.append(cur)
        tr = torch.stack(outs, 1)
        (tr * w[i:i + 1]).sum().backward()
        ref_traj.append(tr.detach())
        ref_grad.append(ui.grad)
    ref_traj, ref_grad = torch.cat(ref_traj), torch.cat(ref_grad)
    p = BGSimParams(dt=dt, resolution=n, nu=0.01, batch_size=b, adaptive_CFL=cfl, adaptive_mode="trajectory",
                    max_substeps=64, time_scheme=scheme, device="cuda")
    sv = BurgersSolverTorch(p)
    u = u0.to(cuda).requires_grad_(True)
    traj, _ = inc_amd.simulate_rollout(None, sv, u, H, None)
    (traj * w.to(cuda)).sum().backward()
    assert rel_l2(traj, ref_traj) < F64_TOL
    assert rel_l2(u.grad, ref_grad) < 1e-10
    nsub = sv.last_substeps.cpu()
    assert nsub.shape == (b, H) and int(nsub.min()) >= 1 and len(set(nsub[:, 0].tolist())) >= 3
    assert abs(sv.t - H * dt) < 1e-14
    # the step-by-step path agrees with the fused rollout
    sv2 = BurgersSolverTorch(p)
    cur = u0.to(cuda)
    for k in range(H):
        cur = sv2.step(cur)
        assert rel_l2(cur, traj[:, k].detach()) < 1e-14


def test_burgers_adaptive_substep_cap(cuda):
    # max_substeps=2 with a CFL that would need more: each step takes the
    # CFL-limited sub-step h1 = cfl*dx/max|u| and then one covering dt - h1
    n, b, dt, cfl = 64, 3, 0.05, 0.05
    gen = torch.Generator().manual_seed(22)
    u0 = 2.0 * smooth(b, n, gen)
    p = BGSimParams(dt=dt, resolution=n, nu=0.01, batch_size=b, adaptive_CFL=cfl, adaptive_mode="trajectory",
                    max_substeps=2, time_scheme="SSPRK3", device="cuda")
    sv = BurgersSolverTorch(p)
    out = sv.step(u0.to(cuda))
    assert sv.last_substeps.cpu().flatten().tolist() == [2] * b
    oracle = BurgersOracle(burgers_params(n, dt, nu=0.01, batch_size=1, time_scheme="SSPRK3"))
    dx = 2 * math.pi / n
    for i in range(b):
        ui = u0[i:i + 1]
        h1 = cfl * dx / float(ui.abs().max())
        ref = oracle._advance(oracle._advance(ui, h1), dt - h1)
        assert rel_l2(out[i:i + 1], ref) < F64_TOL


def test_burgers_argmax_tie_takes_first_index(cuda):
    # sin on a grid samples |u| = 1 at two points: the gradient through
    # alpha = max|u| goes to the first one, as torch.max(dim) does.
    n = 64
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.sin(x)[None].repeat(2, 1)
    assert (u.abs() == u.abs().max()).sum() == 4
    oracle, gpu = make_pair(n, 2, dt=0.01, nu=0.0)
    uo = u.clone().requires_grad_(True)
    w = torch.randn(2, n, generator=torch.Generator().manual_seed(1), dtype=torch.float64)
    (oracle.step(uo) * w).sum().backward()
    ug = u.to(cuda).requires_grad_(True)
    (gpu(torch.float64).step(ug) * w.to(cuda)).sum().backward()
    assert rel_l2(ug.grad, uo.grad) < 1e-11


def test_burgers_nan_propagates(cuda):
    n = 64
    u = torch.zeros(2, n, device=cuda)
    u[0, 5] = float("nan")
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=n, batch_size=2, device="cuda"))
    out = sv.step(u)
    assert torch.isnan(out[0]).any() and not torch.isnan(out[1]).any()
