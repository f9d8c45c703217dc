"""Edge cases on the GPU: empty batches, smallest and largest supported
sizes, odd batches, NaN propagation, argument errors mapped to ValueError."""
import math

import pytest
import torch

import inc_amd
from inc_amd import BGSimParams, BurgersSolverTorch, KSSimParams, KSSolverTorch, RaggedBurgers
from oracle import BurgersOracle, KSOracle, burgers_params, ks_params
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_empty_batches(cuda):
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=64, batch_size=0, device="cuda"))
    assert sv.step(torch.zeros(0, 64, device=cuda)).shape == (0, 64)
    ks = KSSolverTorch(KSSimParams([22.0], resolution=64, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    u = torch.zeros(0, 64, device=cuda, requires_grad=True)
    out = ks.step(u, torch.tensor([22.0], device=cuda))
    out.sum().backward()
    assert out.shape == (0, 64) and u.grad.shape == (0, 64)
    assert inc_amd.weno_reconstruction(torch.zeros(0, 16, device=cuda)).shape == (0, 16)


def test_smallest_and_largest_sizes(cuda):
    gen = torch.Generator().manual_seed(1)
    # Burgers: n = 8 (smallest) and n = 4096 (16 points per thread, fp32)
    for n in (8, 4096):
        x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
        u = torch.sin(x)[None] + 0.1 * torch.randn(2, n, generator=gen, dtype=torch.float64)
        dt = 0.2 * 2 * math.pi / n
        ref = BurgersOracle(burgers_params(n, dt, nu=0.01, batch_size=2, time_scheme="SSPRK3")).step(u)
        sv = BurgersSolverTorch(BGSimParams(dt=dt, resolution=n, nu=0.01, batch_size=2, time_scheme="SSPRK3",
                                            device="cuda"))
        assert rel_l2(sv.step(u.float().to(cuda)), ref) < 1e-5, n
    # KS: n = 16 and n = 1024, fp64 kernels
    for n, Lk in ((16, 8.0), (1024, 200.0)):
        u = 0.3 * torch.randn(3, n, generator=gen, dtype=torch.float64)
        so = KSOracle(ks_params([Lk] * 3, resolution=n, dt=0.1, time_scheme="ETDRK4"))
        so.doublestep = False
        sv = KSSolverTorch(KSSimParams([Lk], resolution=n, dt=0.1, time_scheme="ETDRK4", device="cuda"))
        sv.doublestep = False
        assert rel_l2(sv.step(u.to(cuda), torch.tensor([Lk], device=cuda)), so.step(u, so.domain_size)) < 1e-11


def test_unsupported_sizes_raise_value_error(cuda):
    ks = KSSolverTorch(KSSimParams([22.0], resolution=48, dt=0.25, time_scheme="ETD1", device="cuda"))
    with pytest.raises(ValueError, match="power of two"):
        ks.step(torch.zeros(2, 48, device=cuda), torch.tensor([22.0], device=cuda))
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=4, batch_size=1, device="cuda"))
    with pytest.raises(ValueError, match=">= 8"):
        sv.step(torch.zeros(1, 4, device=cuda))
    with pytest.raises(KeyError):
        KSSolverTorch(KSSimParams([22.0], resolution=64, time_scheme="RK4", device="cuda"))
    with pytest.raises(ValueError):
        BurgersSolverTorch(BGSimParams(convection_term="spectral", device="cuda"))
    bad_t = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=64, batch_size=2, equation_type="learning",
                                           device="cuda"))
    bad_t.param.A = bad_t.param.omega = bad_t.param.phi = bad_t.param.l = torch.zeros(2, 1, 1, device=cuda)
    bad_t.t = torch.zeros(3, 1, device=cuda)
    with pytest.raises(ValueError, match="t tensor"):
        bad_t.step(torch.zeros(2, 64, device=cuda))


def test_ks_nan_propagates_to_its_fft_pair_only(cuda):
    ks = KSSolverTorch(KSSimParams([22.0], resolution=64, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    u = 0.1 * torch.randn(4, 64, generator=torch.Generator().manual_seed(0)).to(cuda)
    u[1, 3] = float("nan")
    out = ks.step(u, torch.tensor([22.0], device=cuda))
    # trajectories 0 and 1 share one packed complex FFT, which mixes a NaN into
    # both halves: the pair partner becomes NaN too (the reference would keep
    # trajectory 0 finite); other pairs are untouched and nothing is clamped
    assert torch.isnan(out[1]).all() and torch.isnan(out[0]).all()
    assert torch.isfinite(out[2:]).all()


def test_ragged_single_trajectory_and_duplicates(cuda):
    n = 128
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.sin(x)
    for ns in ([n], [n, n, n]):
        rb = RaggedBurgers(ns, [0.01] * len(ns), [0.005] * len(ns), dtype=torch.float64)
        out = rb.unpack(rb.advance(rb.pack([u] * len(ns)), None, steps=4))
        sv = BurgersSolverTorch(BGSimParams(dt=0.005, resolution=n, nu=0.01, batch_size=1, time_scheme="SSPRK3",
                                            device="cuda"))
        ref = sv.rollout(u[None].to(cuda), 4)[:, -1]
        for o in out:
            assert rel_l2(o, ref[0]) < 1e-13
