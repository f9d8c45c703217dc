"""Pin the FP64 Burgers oracle: reference golden vectors + closed forms."""
import math

import numpy as np
import pytest
import torch

from oracle import BurgersOracle, burgers_params, weno_reconstruction

T = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64)  # noqa: E731


@pytest.mark.parametrize("name", ["rand", "smooth", "step"])
@pytest.mark.parametrize("direction", ["left", "right"])
def test_weno_matches_reference_golden(golden, name, direction):
    g = golden("weno")
    out = weno_reconstruction(T(g[f"f_{name}"]), direction).numpy()
    np.testing.assert_allclose(out, g[f"{direction}_{name}"], rtol=1e-13, atol=1e-13)


def test_weno_bad_direction():
    with pytest.raises(ValueError):
        weno_reconstruction(torch.zeros(1, 8, dtype=torch.float64), "up")


def test_weno_linear_data_closed_form():
    # For linear data every candidate is exact but periodic wrap breaks the
    # line, so use a quadratic-free slope on the interior of a long grid:
    # with f_i = c*i + d, betas are equal (b terms all equal to 2c, a's 0),
    # weights are the linear ones and the value is f at x_{i+1/2}.
    n = 64
    i = torch.arange(n, dtype=torch.float64)
    f = (0.3 * i + 1.0)[None]
    left = weno_reconstruction(f, "left")[0]
    interior = slice(3, n - 4)
    np.testing.assert_allclose(left[interior], (0.3 * (i + 0.5) + 1.0)[interior], rtol=0, atol=1e-12)
    right = weno_reconstruction(f, "right")[0]
    np.testing.assert_allclose(right[interior], (0.3 * (i + 1.5) + 1.0)[interior], rtol=0, atol=1e-12)


def test_weno_equal_beta_gives_linear_weights():
    # Quadratic data f_i = i^2 gives a_k = 2 and |b_k| = 4|i| for all three
    # left candidates, so the betas coincide, the weights are gamma and the
    # reconstruction is the fixed 5-point combination
    # (2 f_{i-2} - 13 f_{i-1} + 47 f_i + 27 f_{i+1} - 3 f_{i+2}) / 60.
    n = 64
    i = torch.arange(n, dtype=torch.float64)
    f = ((i - 20.0) ** 2)[None]
    left = weno_reconstruction(f, "left")[0]
    r = lambda k: torch.roll(f[0], -k)  # noqa: E731
    lin = (2 * r(-2) - 13 * r(-1) + 47 * r(0) + 27 * r(1) - 3 * r(2)) / 60
    interior = slice(3, n - 4)
    np.testing.assert_allclose(left[interior], lin[interior], rtol=1e-12, atol=1e-10)


def test_weno_alternating_data_right_equals_shifted_left():
    # On period-2 data the mirrored beta_2 term of 'right' has the same
    # magnitude as the shifted left one, so right == roll(left, -1).
    n = 16
    f = torch.cos(math.pi * torch.arange(n, dtype=torch.float64))[None]
    left = weno_reconstruction(f, "left")
    right = weno_reconstruction(f, "right")
    np.testing.assert_allclose(right, torch.roll(left, -1, -1), rtol=0, atol=1e-14)


def test_weno_right_is_shifted_left_except_beta2():
    # On random data the mirrored beta_2 term makes 'right' differ from the
    # shifted 'left'; on data symmetric about i+2 they agree.
    g = torch.Generator().manual_seed(0)
    f = torch.randn(2, 32, generator=g, dtype=torch.float64)
    left = weno_reconstruction(f, "left")
    right = weno_reconstruction(f, "right")
    assert (right - torch.roll(left, -1, -1)).abs().max() > 1e-6


def _burgers_from_golden(g, conv, eq="homogeneous", forcing=False):
    u = T(g["u"])
    b, n = u.shape
    kw = {}
    if forcing:
        kw = dict(J=2, A=T(g["A"]), omega=T(g["omega"]), phi=T(g["phi"]), l=T(g["l"]))
    p = burgers_params(n, float(g["dt"]), L=float(g["L"]), nu=T(g["nu"]), batch_size=b,
                       convection_term=conv, equation_type=eq, **kw)
    return BurgersOracle(p), u, T(g["s"])


def test_burgers_pieces_match_reference_golden(golden):
    g = golden("burgers")
    sw, u, s = _burgers_from_golden(g, "WENO")
    sf, _, _ = _burgers_from_golden(g, "FDM")
    np.testing.assert_allclose(sw.calc_nonlinear_weno(u), g["nonlinear_weno"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sf.calc_nonlinear_FDM(u), g["nonlinear_fdm"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sw.calc_linear(u), g["linear"], rtol=1e-12, atol=1e-9)
    sw.correction_s = s
    sf.correction_s = s
    np.testing.assert_allclose(sw.step(u), g["euler_weno_s"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sf.step(u), g["euler_fdm_s"], rtol=1e-12, atol=1e-12)


def test_burgers_learning_forcing_matches_reference_golden(golden):
    g = golden("burgers")
    sl, u, s = _burgers_from_golden(g, "WENO", "learning", forcing=True)
    sl.t = T(g["t"])
    np.testing.assert_allclose(sl._compute_source(u, sl.t), g["forcing"], rtol=1e-12, atol=1e-12)
    sl.correction_s = s
    np.testing.assert_allclose(sl.step(u), g["euler_learning_s"], rtol=1e-12, atol=1e-12)


def test_burgers_gradient_matches_reference_autograd(golden):
    g = golden("burgers")
    sw, u, _ = _burgers_from_golden(g, "WENO")
    ss = T(g["grad_ss"])
    tgt = T(g["grad_target"])
    u0 = u.clone().requires_grad_(True)
    s_leaf = ss.clone().requires_grad_(True)
    cur, traj = u0, []
    for k in range(ss.shape[1]):
        sw.correction_s = s_leaf[:, k]
        cur = sw.step(cur)
        traj.append(cur)
    loss = ((torch.stack(traj, 1) - tgt) ** 2).mean()
    loss.backward()
    np.testing.assert_allclose(loss.item(), float(g["grad_loss"]), rtol=1e-13)
    np.testing.assert_allclose(u0.grad, g["grad_u0"], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(s_leaf.grad, g["grad_s"], rtol=1e-11, atol=1e-13)


def test_burgers_mean_conservation_and_constant_fixed_point():
    n = 64
    p = burgers_params(n, 0.01, nu=0.02, batch_size=2, time_scheme="SSPRK3")
    so = BurgersOracle(p)
    g = torch.Generator().manual_seed(1)
    u = torch.randn(2, n, generator=g, dtype=torch.float64)
    u1 = so.step(u)
    np.testing.assert_allclose(u1.sum(-1), u.sum(-1), rtol=0, atol=1e-11)
    c = torch.full((2, n), 0.7, dtype=torch.float64)
    np.testing.assert_allclose(so.step(c), c, rtol=0, atol=1e-13)


def test_laplacian_eigenvalue():
    n, L = 32, 2 * math.pi
    p = burgers_params(n, 0.01, L=L, nu=1.0)
    so = BurgersOracle(p)
    x = so.x
    for k in (1, 3, 7):
        v = torch.sin(k * x)[None]
        lam = -(4 / so.dx ** 2) * math.sin(k * so.dx / 2) ** 2
        np.testing.assert_allclose(so.calc_linear(v), lam * v, rtol=0, atol=1e-10)


def test_sspk3_amplification_factor_on_diffusive_mode():
    # u = eps*sin(kx): at tiny amplitude convection is O(eps^2) and the step
    # is linear diffusion: u' = (1+z+z^2/2+z^3/6) u with z = dt*lambda.
    n, L, nu, dt = 32, 2 * math.pi, 0.05, 0.05
    p = burgers_params(n, dt, L=L, nu=nu, time_scheme="SSPRK3", convection_term="FDM")
    so = BurgersOracle(p)
    k = 3
    eps = 1e-7
    v = eps * torch.sin(k * so.x)[None]
    z = dt * nu * (-(4 / so.dx ** 2) * math.sin(k * so.dx / 2) ** 2)
    amp = 1 + z + z * z / 2 + z ** 3 / 6
    np.testing.assert_allclose(so.step(v) / eps, amp * v / eps, rtol=0, atol=1e-8)
    # Euler: 1 + z
    so_e = BurgersOracle(burgers_params(n, dt, L=L, nu=nu, convection_term="FDM"))
    np.testing.assert_allclose(so_e.step(v) / eps, (1 + z) * v / eps, rtol=0, atol=1e-8)


def test_burgers_gradient_central_difference():
    n = 32
    p = burgers_params(n, 0.01, nu=0.02, batch_size=2, time_scheme="SSPRK3")
    so = BurgersOracle(p)
    g = torch.Generator().manual_seed(3)
    u = torch.randn(2, n, generator=g, dtype=torch.float64)
    s = 0.1 * torch.randn(2, n, generator=g, dtype=torch.float64)
    w = torch.randn(2, n, generator=g, dtype=torch.float64)

    def f(uu, sv):
        so.correction_s = sv
        return (so.step(uu) * w).sum()

    u0 = u.clone().requires_grad_(True)
    s0 = s.clone().requires_grad_(True)
    f(u0, s0).backward()
    for _ in range(3):
        du = torch.randn(2, n, generator=g, dtype=torch.float64)
        ds = torch.randn(2, n, generator=g, dtype=torch.float64)
        h = 1e-6
        fd = (f(u + h * du, s + h * ds) - f(u - h * du, s - h * ds)) / (2 * h)
        an = (u0.grad * du).sum() + (s0.grad * ds).sum()
        np.testing.assert_allclose(fd.item(), an.item(), rtol=1e-6)


def _learning_solver(dt, time_scheme, b=2, n=64, l=1.0, nu=0.1):
    """'learning' forcing sum_j A_j sin(omega_j t + 2 pi l x / L + phi_j) with
    two modes per trajectory (solver_1d.py:74-85), seeded."""
    g = torch.Generator().manual_seed(11)
    A = 0.5 + torch.rand(b, 1, 2, generator=g, dtype=torch.float64)
    om = 1 + 3 * torch.rand(b, 1, 2, generator=g, dtype=torch.float64)
    ph = 2 * math.pi * torch.rand(b, 1, 2, generator=g, dtype=torch.float64)
    ll = torch.full((b, 1, 2), l, dtype=torch.float64)
    p = burgers_params(n, dt, nu=nu, batch_size=b, time_scheme=time_scheme, equation_type="learning",
                       J=2, A=A, omega=om, phi=ph, l=ll)
    return BurgersOracle(p), (A, om, ph)


@pytest.mark.parametrize("scheme", ["SSPRK3", "Euler"])
def test_stage_times_on_a_spatially_constant_state(scheme):
    # A constant state has zero convection (every cell reconstructs the same
    # flux) and zero diffusion; with l = 0 the forcing f(t) is constant in
    # space, so u_t = f(t) and one SSP-RK3 step is exactly Simpson's rule
    # u + dt/6 (f(t) + 4 f(t + dt/2) + f(t + dt)) -- the stage times t,
    # t + dt, t + dt/2 -- and one Euler step is u + dt f(t).
    dt, t0 = 0.07, 0.3
    so, (A, om, ph) = _learning_solver(dt, scheme, l=0.0)
    so.t = t0
    u = torch.tensor([[0.4], [-0.9]], dtype=torch.float64).expand(2, 64).contiguous()

    def f(t):
        return (A[:, 0] * torch.sin(om[:, 0] * t + ph[:, 0])).sum(-1, keepdim=True)

    if scheme == "SSPRK3":
        want = u + dt / 6 * (f(t0) + 4 * f(t0 + dt / 2) + f(t0 + dt))
    else:
        want = u + dt * f(t0)
    np.testing.assert_allclose(so.step(u), want, rtol=0, atol=1e-14)
    assert so.t == pytest.approx(t0 + dt)


def test_sspk3_third_order_with_time_dependent_forcing():
    # Smooth viscous solution with 'learning' forcing (l = 1) over t in
    # [0.2, 0.7]: the error against a dt = 1/1280 run falls ~8x per halving
    # of dt (order 3).  A forcing evaluated at a wrong stage time drops the
    # order to 1 (ratio ~2), so this pins the stage times under transport.
    def run(dt, tend=0.5):
        so, _ = _learning_solver(dt, "SSPRK3")
        so.t = 0.2
        x = so.x
        u = torch.stack([0.3 * torch.sin(x), 0.2 * torch.cos(2 * x) + 0.1])
        for _ in range(int(round(tend / dt))):
            u = so.step(u)
        return u

    ref = run(1 / 1280)
    errs = [float((run(dt) - ref).norm()) for dt in (1 / 20, 1 / 40, 1 / 80)]
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(6.5 < r < 9.5 for r in ratios), (errs, ratios)
