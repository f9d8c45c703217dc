"""Pin the FP64 KS oracle: reference golden vectors, closed forms, ETDRK4
coefficient identities and convergence order."""
import math

import numpy as np
import pytest
import torch

from oracle import KSOracle, etdrk4_coefficients, ks_params

T = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64)  # noqa: E731
CASES = [f"n{n}_{s}_{d}" for n in (32, 64) for s in ("ETD1", "ETRK2") for d in ("mixed", "twopi")]


def _solver(key, g):
    n = int(key.split("_")[0][1:])
    scheme = key.split("_")[1]
    return KSOracle(ks_params(T(g[f"{key}_L"]), resolution=n, dt=0.1, time_scheme=scheme)), T(g[f"{key}_L"])


@pytest.mark.parametrize("key", CASES)
def test_ks_steps_match_reference_golden(golden, key):
    g = golden("ks")
    so, L = _solver(key, g)
    u, s, cu = T(g[f"{key}_u"]), T(g[f"{key}_s"]), T(g[f"{key}_cu"])
    tol = dict(rtol=1e-11, atol=1e-12)
    so.doublestep = False
    np.testing.assert_allclose(so.step(u, L), g[f"{key}_single"], **tol)
    so.doublestep = True
    np.testing.assert_allclose(so.step(u, L), g[f"{key}_double"], **tol)
    so.correction_s = s
    np.testing.assert_allclose(so.step(u, L), g[f"{key}_double_s"], **tol)
    so.correction_u = cu
    np.testing.assert_allclose(so.step(u, L), g[f"{key}_double_s_cu"], **tol)


@pytest.mark.parametrize("key", CASES)
def test_ks_gradient_matches_reference_autograd(golden, key):
    g = golden("ks")
    so, L = _solver(key, g)
    u = T(g[f"{key}_u"])
    ss = T(g[f"{key}_gss"]).clone().requires_grad_(True)
    tgt = T(g[f"{key}_gtarget"])
    u0 = u.clone().requires_grad_(True)
    cur, traj = u0, []
    for k in range(ss.shape[1]):
        so.correction_s = ss[:, k]
        cur = so.step(cur, L)
        traj.append(cur)
    np.testing.assert_allclose(torch.stack(traj, 1).detach(), g[f"{key}_gtraj"], rtol=1e-10, atol=1e-11)
    ((torch.stack(traj, 1) - tgt) ** 2).mean().backward()
    np.testing.assert_allclose(u0.grad, g[f"{key}_gu0"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ss.grad, g[f"{key}_gs"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("n", [64, 75, 96])
@pytest.mark.parametrize("scheme", ["ETD1", "ETRK2", "ETDRK4"])
def test_single_mode_linear_decay(scheme, n):
    # u = eps*cos(k x): the nonlinear term is O(eps^2); the mode evolves by
    # e^{(k^2-k^4) dt} per sub-step (odd and non-power-of-two n included).
    L, dt = 22.0, 0.25
    so = KSOracle(ks_params([L], resolution=n, dt=dt, time_scheme=scheme))
    so.doublestep = False
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    for j in (1, 3, 5):
        k = 2 * math.pi * j / L
        eps = 1e-9
        u = eps * torch.cos(k * x)[None]
        out = so.step(u, so.domain_size) / eps
        np.testing.assert_allclose(out, math.exp((k * k - k ** 4) * dt) * u / eps, rtol=0, atol=1e-7)


@pytest.mark.parametrize("scheme", ["ETD1", "ETRK2", "ETDRK4"])
def test_mean_conserved_and_constant_forcing(scheme):
    n, L, dt = 64, 32.0, 0.2
    so = KSOracle(ks_params([L, L], resolution=n, dt=dt, time_scheme=scheme))
    so.doublestep = False
    g = torch.Generator().manual_seed(5)
    u = torch.randn(2, n, generator=g, dtype=torch.float64) * 0.3
    out = so.step(u, so.domain_size)
    np.testing.assert_allclose(out.mean(-1), u.mean(-1), rtol=0, atol=1e-13)
    # constant forcing s = c adds c*dt to the mean (k = 0: E = 1, phi1 = dt,
    # and f1 + 4 f2 + f3 = dt for ETDRK4)
    so.correction_s = torch.full((2, n), 0.7, dtype=torch.float64)
    out = so.step(u, so.domain_size)
    np.testing.assert_allclose(out.mean(-1), u.mean(-1) + 0.7 * dt, rtol=0, atol=1e-12)


def test_etdrk4_coefficients_against_direct_and_taylor():
    h = 0.25
    big = torch.tensor([-4.0, -10.0, -100.0, -16000.0, 3.0], dtype=torch.float64) / h
    e, e2, q, f1, f2, f3 = etdrk4_coefficients(big, h)
    z = h * big
    ez = torch.exp(z)
    np.testing.assert_allclose(q, h * (torch.exp(z / 2) - 1) / z, rtol=1e-12)
    np.testing.assert_allclose(f1, h * (-4 - z + ez * (4 - 3 * z + z * z)) / z ** 3, rtol=1e-10)
    np.testing.assert_allclose(f2, h * (2 + z + ez * (z - 2)) / z ** 3, rtol=1e-10)
    np.testing.assert_allclose(f3, h * (-4 - 3 * z - z * z + ez * (4 - z)) / z ** 3, rtol=1e-10)
    # small |z|: Taylor series of phi functions
    small = torch.tensor([0.0, 1e-8, -1e-4, 2e-3, -2.4e-3], dtype=torch.float64) / h
    _, _, q, f1, f2, f3 = etdrk4_coefficients(small, h)
    z = h * small
    # series through z^3 (next terms are below 1e-11 relative here)
    np.testing.assert_allclose(q, h * (1 / 2 + z / 8 + z ** 2 / 48 + z ** 3 / 384), rtol=1e-11)
    np.testing.assert_allclose(f1, h * (1 / 6 + z / 6 + 3 * z ** 2 / 40 + z ** 3 / 45), rtol=1e-10)
    np.testing.assert_allclose(f2, h * (1 / 6 + z / 12 + z ** 2 / 40 + z ** 3 / 180), rtol=1e-10)
    np.testing.assert_allclose(f3, h * (1 / 6 - z ** 2 / 120 - z ** 3 / 360), rtol=1e-10)


def test_etdrk4_phi_identity():
    h = 0.25
    lop = torch.linspace(-500, 0.3, 997, dtype=torch.float64)
    e, _, _, f1, f2, f3 = etdrk4_coefficients(lop, h)
    safe = torch.where(lop == 0, torch.ones_like(lop), lop)
    want = torch.where(lop == 0, torch.full_like(lop, h), (e - 1) / safe)
    np.testing.assert_allclose(f1 + 4 * f2 + f3, want, rtol=1e-11, atol=1e-14)


def test_etdrk4_fourth_order_convergence():
    # smooth KS solution over T = 1 on L = 22: error of ETDRK4 against a
    # fine-step reference drops by ~16x per halving of dt.
    n, L, tend = 64, 22.0, 1.0
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u0 = (torch.cos(2 * math.pi * x / L) * (1 + torch.sin(2 * math.pi * x / L)))[None] * 0.8

    def run(dt):
        so = KSOracle(ks_params([L], resolution=n, dt=dt, time_scheme="ETDRK4", dealias=False))
        so.doublestep = False
        u = u0
        for _ in range(int(round(tend / dt))):
            u = so.step(u, so.domain_size)
        return u

    ref = run(1 / 1024)
    errs = [float((run(dt) - ref).norm()) for dt in (1 / 16, 1 / 32, 1 / 64)]
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    # order 4 means a ratio of 16 asymptotically; order 3 would give 8
    assert ratios[0] > 11 and ratios[1] > 13, (errs, ratios)


def test_ks_fft_is_textbook_dft():
    # the oracle's torch.fft agrees with a brute-force O(N^2) DFT
    n = 32
    g = torch.Generator().manual_seed(9)
    u = torch.randn(n, generator=g, dtype=torch.float64)
    j = torch.arange(n, dtype=torch.float64)
    w = torch.exp(-2j * math.pi * torch.outer(j, j) / n)
    np.testing.assert_allclose(torch.fft.fft(u), w @ u.to(torch.complex128), rtol=0, atol=1e-12)


def test_ks_gradient_central_difference_etdrk4():
    n, L = 32, 22.0
    so = KSOracle(ks_params([L, 30.0], resolution=n, dt=0.25, time_scheme="ETDRK4"))
    g = torch.Generator().manual_seed(4)
    u = torch.randn(2, n, generator=g, dtype=torch.float64) * 0.5
    s = torch.randn(2, n, generator=g, dtype=torch.float64) * 0.1
    w = torch.randn(2, n, generator=g, dtype=torch.float64)

    def f(uu, sv):
        so.correction_s = sv
        return (so.step(uu, so.domain_size) * w).sum()

    u0 = u.clone().requires_grad_(True)
    s0 = s.clone().requires_grad_(True)
    f(u0, s0).backward()
    for _ in range(4):
        du = torch.randn(2, n, generator=g, dtype=torch.float64)
        ds = torch.randn(2, n, generator=g, dtype=torch.float64)
        h = 1e-6
        fd = (f(u + h * du, s + h * ds) - f(u - h * du, s - h * ds)) / (2 * h)
        an = (u0.grad * du).sum() + (s0.grad * ds).sum()
        np.testing.assert_allclose(fd.item(), an.item(), rtol=1e-7)


def _two_mode_state(n, k1, k2, L):
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    return (0.3 * torch.cos(2 * math.pi * k1 * x / L) + 0.2 * torch.cos(2 * math.pi * k2 * x / L + 0.4))[None]


@pytest.mark.parametrize("n,k1", [(256, 43), (256, 60), (256, 85), (512, 86), (512, 120), (512, 170)])
def test_dealias_kills_the_product_of_a_mode_above_n_over_6(n, k1):
    # u = a cos(2 pi k1 x / L) with n/6 < k1 < n/3: u^2 holds modes 0 and
    # 2 k1 only.  Mode 0 has g = 0 and |2 k1| (aliased or not) is above the
    # 2/3 cut n/3, so with the mask the nonlinear term vanishes at every
    # stage and one ETDRK4 step is the exact linear map e^{(k^2 - k^4) dt} of
    # the single mode.  Without the mask the product mode survives (2 k1 is
    # never the Nyquist index n/2 here, where g = 0 anyway).
    assert n / 6 < k1 < n / 3 and 2 * k1 != n // 2
    L = 2 * math.pi * k1 / 0.9  # wavenumber 0.9: the mode neither grows nor decays much
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u = 0.3 * torch.cos(2 * math.pi * k1 * x / L)[None]
    k = 2 * math.pi * k1 / L
    exact = math.exp((k * k - k ** 4) * 0.25) * u
    for dealias in (True, False):
        so = KSOracle(ks_params([L], resolution=n, dt=0.25, time_scheme="ETDRK4", dealias=dealias))
        so.doublestep = False
        err = float((so.step(u, so.domain_size) - exact).norm() / exact.norm())
        if dealias:
            assert err < 1e-13, err
        else:
            assert err > 1e-3, err


@pytest.mark.parametrize("n,k1,k2", [(256, 43, 42), (512, 86, 85)])
def test_dealias_cut_is_between_floor_and_ceil_of_n_over_3(n, k1, k2):
    # Two modes whose products (k1 + k2, 2 k1, 2 k2) reach exactly the last
    # kept wavenumber (85 at n = 256, 170 at n = 512) and the first cut one
    # (86, 171).  After one ETDRK4 step with
    # the mask every |k| > n/3 is zero to rounding and |k| = floor(n/3) is
    # not; without the mask the first cut wavenumber is populated.
    L = 2 * math.pi * k1 / 0.8
    u = _two_mode_state(n, k1, k2, L)
    kk = (torch.fft.fftfreq(n, d=1.0 / n)).abs()
    last, first_cut = n // 3, n // 3 + 1
    assert {last, first_cut} <= {k1 + k2, 2 * k1, 2 * k2}
    for dealias in (True, False):
        so = KSOracle(ks_params([L], resolution=n, dt=0.25, time_scheme="ETDRK4", dealias=dealias))
        so.doublestep = False
        sp = torch.fft.fft(so.step(u, so.domain_size)[0]).abs()
        sp = sp / sp.max()
        if dealias:
            assert float(sp[kk >= first_cut].max()) < 1e-14
            assert float(sp[kk == last].max()) > 1e-3
        else:
            assert float(sp[kk == first_cut].max()) > 1e-3


def test_initial_conditions_noise_has_unit_frobenius_norm_over_the_batch():
    # solver_1d.py:214-223: cos(2 pi x) + alpha sin(2 pi x) on the
    # endpoint-inclusive linspace(0, 1, N), plus eps / ||eps|| with the norm
    # taken over the whole [B, N] batch (Frobenius), not per trajectory.
    n = 48
    alpha = torch.tensor([0.0, 0.5, -1.3], dtype=torch.float64)
    so = KSOracle(ks_params([22.0, 30.0, 64.0], resolution=n, alpha=alpha))
    u = so.initial_conditions(generator=torch.Generator().manual_seed(3))
    x = torch.linspace(0, 1, n, dtype=torch.float64)
    assert float(x[-1]) == 1.0
    det = torch.cos(2 * math.pi * x) + alpha[:, None] * torch.sin(2 * math.pi * x)
    noise = u - det
    np.testing.assert_allclose(float(torch.linalg.norm(noise)), 1.0, rtol=1e-12)
    # per-trajectory norms are not each 1 (the batch shares one normaliser)
    assert float((noise.norm(dim=1) - 1).abs().max()) > 1e-2
    with pytest.raises(ValueError):
        KSOracle(ks_params([22.0, 30.0], resolution=n, alpha=alpha)).initial_conditions()


def _resample(u, m):
    """Spectral interpolation of a periodic row onto m points (the modes
    |k| < min(n, m) / 2 kept)."""
    n = u.shape[-1]
    U = torch.fft.fft(u)
    h = (min(n, m) - 1) // 2
    out = torch.zeros(m, dtype=torch.complex128)
    out[:h + 1] = U[:h + 1]
    out[m - h:] = U[n - h:]
    return torch.fft.ifft(out).real * (m / n)


@pytest.mark.parametrize("n", [75, 96, 127, 150])
@pytest.mark.parametrize("scheme", ["ETD1", "ETRK2", "ETDRK4"])
def test_resolution_independence_of_a_resolved_state(scheme, n):
    # A smooth state (modes 1 and 2, amplitude 0.5, L = 22) stays resolved
    # to rounding at every n here over 10 steps: its spectrum falls below
    # 1e-16 well inside n/3, so neither the grid nor the 2/3 cut matters,
    # and the solution at odd, prime and mixed-radix n, interpolated to 256
    # points, is the n = 256 solution.  A wrong wavenumber of any mode
    # (the negative half of an odd-n spectrum, the Nyquist handling of
    # an even one) breaks the agreement.
    L, dt = 22.0, 0.1

    def run(m):
        x = torch.arange(m, dtype=torch.float64) * (L / m)
        u = (0.5 * torch.cos(2 * math.pi * x / L) + 0.25 * torch.sin(4 * math.pi * x / L + 0.3))[None]
        so = KSOracle(ks_params([L], resolution=m, dt=dt, time_scheme=scheme))
        so.doublestep = False
        for _ in range(10):
            u = so.step(u, so.domain_size)
        return u[0]

    ref = run(256)
    got = _resample(run(n), 256)
    assert float((got - ref).norm() / ref.norm()) < 1e-13
