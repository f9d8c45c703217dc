"""The closed-form pins of the added schemes, run through the HIP solvers.

The same cases pin the FP64 oracle (tests/test_oracle_ks.py,
tests/test_oracle_burgers.py); here the kernels are checked against the
mathematics directly, not against the oracle:

* 2/3-rule mask: a single KS mode above n/6 has its whole product cut, so one
  ETDRK4 step is the exact linear map of that mode with the mask and is not
  without it; two modes whose products reach floor(n/3) and floor(n/3) + 1
  show the cut between them.
* SSP-RK3 stage times: on a spatially constant state with spatially constant
  'learning' forcing one step is Simpson's rule exactly; with transport and
  time-dependent forcing the error falls ~8x per halving of dt.

Tolerances: float64 to rounding (1e-12 relative); float32 to its FFT /
stencil rounding (2e-6 relative, 1e-6 absolute on O(1) values, 1e-5 for
the spectrum of the cut modes).
"""
import math

import pytest
import torch

from inc_amd import BGSimParams, BurgersSolverTorch, KSSimParams, KSSolverTorch

pytestmark = pytest.mark.gpu


def _ks(L, n, dtype_dealias, radix=0):
    p = KSSimParams(torch.tensor([L], dtype=torch.float64, device="cuda"), resolution=n, dt=0.25,
                    time_scheme="ETDRK4", device="cuda", dealias=dtype_dealias)
    p.radix = radix
    sv = KSSolverTorch(p)
    sv.doublestep = False
    return sv


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n,k1", [(256, 43), (256, 85), (512, 86), (512, 170)])
def test_ks_dealias_single_mode_is_the_exact_linear_map(cuda, dtype, n, k1):
    L = 2 * math.pi * k1 / 0.9
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u = 0.3 * torch.cos(2 * math.pi * k1 * x / L)[None].repeat(2, 1)  # an FFT-packed pair
    u[1] *= -0.5
    k = 2 * math.pi * k1 / L
    exact = math.exp((k * k - k ** 4) * 0.25) * u
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    Ld = torch.tensor([L], dtype=torch.float64, device=cuda)
    for dealias in (True, False):
        out = _ks(L, n, dealias).step(u.to(cuda, dtype), Ld).double().cpu()
        err = float((out - exact).norm() / exact.norm())
        if dealias:
            assert err < tol, err
        else:
            assert err > 1e-3, err


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n,k1,k2", [(256, 43, 42), (512, 86, 85)])
def test_ks_dealias_cut_between_floor_and_ceil_of_n_over_3(cuda, dtype, n, k1, k2):
    L = 2 * math.pi * k1 / 0.8
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u = (0.3 * torch.cos(2 * math.pi * k1 * x / L) + 0.2 * torch.cos(2 * math.pi * k2 * x / L + 0.4))[None]
    kk = torch.fft.fftfreq(n, d=1.0 / n).abs()
    last, first_cut = n // 3, n // 3 + 1
    assert {last, first_cut} <= {k1 + k2, 2 * k1, 2 * k2}
    zero_tol = 1e-13 if dtype == torch.float64 else 1e-5
    Ld = torch.tensor([L], dtype=torch.float64, device=cuda)
    for dealias in (True, False):
        out = _ks(L, n, dealias).step(u.to(cuda, dtype), Ld).double().cpu()
        sp = torch.fft.fft(out[0]).abs()
        sp = sp / sp.max()
        if dealias:
            assert float(sp[kk >= first_cut].max()) < zero_tol
            assert float(sp[kk == last].max()) > 1e-3
        else:
            assert float(sp[kk == first_cut].max()) > 1e-3


def _learning(dt, scheme, dtype, b=2, n=64, l=1.0, nu=0.1):
    g = torch.Generator().manual_seed(11)
    A = 0.5 + torch.rand(b, 1, 2, generator=g, dtype=torch.float64)
    om = 1 + 3 * torch.rand(b, 1, 2, generator=g, dtype=torch.float64)
    ph = 2 * math.pi * torch.rand(b, 1, 2, generator=g, dtype=torch.float64)
    ll = torch.full((b, 1, 2), l, dtype=torch.float64)
    p = BGSimParams(dt=dt, resolution=n, nu=nu, batch_size=b, equation_type="learning", time_scheme=scheme,
                    A=A.cuda(), omega=om.cuda(), phi=ph.cuda(), l=ll.cuda(), device="cuda")
    return BurgersSolverTorch(p), (A, om, ph)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("scheme", ["SSPRK3", "Euler"])
def test_burgers_stage_times_simpson_on_constant_state(cuda, dtype, scheme):
    dt, t0 = 0.07, 0.3
    sv, (A, om, ph) = _learning(dt, scheme, dtype, l=0.0)
    sv.t = t0
    u = torch.tensor([[0.4], [-0.9]], dtype=torch.float64).expand(2, 64).contiguous()

    def f(t):
        return (A[:, 0] * torch.sin(om[:, 0] * t + ph[:, 0])).sum(-1, keepdim=True)

    want = (u + dt / 6 * (f(t0) + 4 * f(t0 + dt / 2) + f(t0 + dt))) if scheme == "SSPRK3" else u + dt * f(t0)
    out = sv.step(u.to(cuda, dtype)).double().cpu()
    tol = 1e-14 if dtype == torch.float64 else 1e-6
    assert float((out - want).abs().max()) < tol


def test_burgers_sspk3_third_order_with_time_dependent_forcing(cuda):
    def run(dt, tend=0.5):
        sv, _ = _learning(dt, "SSPRK3", torch.float64)
        sv.t = 0.2
        x = torch.arange(64, dtype=torch.float64) * (2 * math.pi / 64)
        u = torch.stack([0.3 * torch.sin(x), 0.2 * torch.cos(2 * x) + 0.1]).to(cuda)
        for _ in range(int(round(tend / dt))):
            u = sv.step(u)
        return u.cpu()

    ref = run(1 / 1280)
    errs = [float((run(dt) - ref).norm()) for dt in (1 / 20, 1 / 40, 1 / 80)]
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(6.5 < r < 9.5 for r in ratios), (errs, ratios)


def test_ks_initial_conditions_on_the_gpu(cuda):
    torch.manual_seed(1)
    p = KSSimParams([22.0, 30.0, 64.0], resolution=64, device="cuda")
    u = KSSolverTorch(p).initial_conditions()
    assert u.is_cuda and u.shape == (3, 64)
    x = torch.linspace(0, 1, 64, device=cuda)
    det = torch.cos(2 * torch.pi * x) + p.alpha[:, None] * torch.sin(2 * torch.pi * x)
    assert abs(float(torch.linalg.norm((u - det).double())) - 1.0) < 1e-5
