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
    ks = KSSolverTorch(KSSimParams([22.0], resolution=2048, dt=0.25, time_scheme="ETD1", device="cuda"))
    with pytest.raises(ValueError, match="resolution"):
        ks.step(torch.zeros(2, 2048, device=cuda), torch.tensor([22.0], device=cuda))
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=4, batch_size=1, device="cuda"))
    with pytest.raises(ValueError, match=">= 8"):
        sv.step(torch.zeros(1, 4, device=cuda))
    # a state whose length is not the solver's resolution (dx and the
    # stencil belong to the solver's grid)
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=64, batch_size=2, device="cuda"))
    with pytest.raises(ValueError, match="resolution"):
        sv.step(torch.zeros(2, 32, device=cuda))
    with pytest.raises(ValueError, match="resolution"):
        sv.stability_numbers(torch.zeros(2, 128, device=cuda))
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


def test_ks_nan_stays_in_its_trajectory_when_unpacked(cuda):
    # the reference transforms each row on its own (torch.fft along the last
    # axis, solver_1d.py:249-255): a NaN in trajectory 1 makes only that
    # trajectory NaN.  pack_pairs=False runs one trajectory per FFT and keeps
    # that behaviour; the default packs two trajectories into one complex FFT,
    # which mixes the NaN into the pair partner (trajectory 0) as well.
    u = 0.1 * torch.randn(4, 64, generator=torch.Generator().manual_seed(0)).to(cuda)
    u[1, 3] = float("nan")
    Ld = torch.tensor([22.0], device=cuda)
    for pack in (False, True):
        ks = KSSolverTorch(KSSimParams([22.0], resolution=64, dt=0.25, time_scheme="ETDRK4", device="cuda",
                                       pack_pairs=pack))
        out = ks.step(u, Ld)
        assert torch.isnan(out[1]).all()
        assert torch.isfinite(out[2:]).all()
        if pack:
            assert torch.isnan(out[0]).all()
        else:
            alone = ks.step(u[[0, 2]], Ld)
            assert torch.isfinite(out[0]).all()
            assert rel_l2(out[0], alone[0]) < 1e-6


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ks_unpacked_matches_oracle_and_packed(cuda, dtype):
    gen = torch.Generator().manual_seed(4)
    n, Lk = 128, 40.0
    u = 0.4 * torch.randn(5, n, generator=gen, dtype=torch.float64)
    s = 0.05 * torch.randn(5, n, generator=gen, dtype=torch.float64)
    so = KSOracle(ks_params([Lk] * 5, resolution=n, dt=0.25, time_scheme="ETDRK4"))
    so.doublestep = False
    so.correction_s = s
    ref = so.step(u, so.domain_size)
    outs = []
    for pack in (False, True):
        sv = KSSolverTorch(KSSimParams([Lk], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda",
                                       pack_pairs=pack))
        sv.doublestep = False
        sv.correction_s = s.to(cuda, dtype)
        ug = u.to(cuda, dtype).requires_grad_(True)
        out = sv.step(ug, torch.tensor([Lk], device=cuda))
        out.square().sum().backward()
        outs.append((out.detach(), ug.grad))
        assert rel_l2(out, ref) < (1e-12 if dtype == torch.float64 else 1e-5)
    assert rel_l2(outs[0][1], outs[1][1]) < (1e-12 if dtype == torch.float64 else 1e-5)


def test_model_parameter_dtype_and_device_are_checked_before_launch(cuda):
    from inc_amd.correction import CorrectionCNN, circular_conv1d
    cnn = CorrectionCNN(3, 32, 5, generator=torch.Generator().manual_seed(0))
    u64 = torch.randn(4, 64, dtype=torch.float64, device=cuda)
    # parameters on the host: refused before any pointer reaches a kernel
    with pytest.raises(ValueError, match="GPU"):
        cnn(u64.float())
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=64, batch_size=4, time_scheme="SSPRK3", device="cuda"))
    with torch.no_grad(), pytest.raises(ValueError, match="GPU"):
        inc_amd.simulate_rollout(cnn, sv, u64.float(), 3, "INC")
    cnn = cnn.cuda()
    # float32 parameters under a float64 state: TypeError, in the network,
    # the fused rollout and the per-step rollout alike
    with pytest.raises(TypeError):
        cnn(u64)
    with torch.no_grad(), pytest.raises(TypeError):
        inc_amd.simulate_rollout(cnn, sv, u64, 3, "INC")
    with pytest.raises(TypeError):
        sv.rollout_inc(cnn, u64, 3)
    with pytest.raises(TypeError):
        circular_conv1d(u64[:, None], cnn.convs[0].weight, cnn.convs[0].bias)
    # matching dtypes run
    with torch.no_grad():
        res, corr = inc_amd.simulate_rollout(cnn.double(), sv, u64, 3, "INC")
    assert torch.isfinite(res).all() and corr.shape == (4, 3, 64)


@pytest.mark.parametrize("shape", [(64,), (3, 1), (), (3, 64)])
def test_burgers_correction_broadcasts_like_the_reference(cuda, shape):
    # source + correction_s broadcasts in the reference (solver_1d.py:116-123):
    # a [N], [B, 1] or scalar correction acts on every trajectory / point,
    # and its gradient is summed back to its own shape
    gen = torch.Generator().manual_seed(7)
    n, b = 64, 3
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.sin(x)[None] * torch.tensor([[1.0], [0.6], [-0.8]], dtype=torch.float64)
    s = 0.1 * torch.randn(shape, generator=gen, dtype=torch.float64)
    w = torch.randn(b, n, generator=gen, dtype=torch.float64)
    so = BurgersOracle(burgers_params(n, 0.01, nu=0.01, batch_size=b, time_scheme="SSPRK3"))
    so_s = s.clone().requires_grad_(True)
    so.correction_s = so_s
    uo = u.clone().requires_grad_(True)
    ref = so.step(uo)
    (ref * w).sum().backward()
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b, time_scheme="SSPRK3",
                                        device="cuda"))
    sg = s.to(cuda).requires_grad_(True)
    sv.correction_s = sg
    ug = u.to(cuda).requires_grad_(True)
    out = sv.step(ug)
    (out * w.to(cuda)).sum().backward()
    assert sg.grad.shape == s.shape
    assert rel_l2(out, ref.detach()) < 1e-12
    assert rel_l2(ug.grad, uo.grad) < 1e-11
    assert rel_l2(sg.grad, so_s.grad) < 1e-11


def test_ks_correction_broadcasts_over_the_batch(cuda):
    gen = torch.Generator().manual_seed(8)
    n, b, Lk = 64, 4, 22.0
    u = 0.3 * torch.randn(b, n, generator=gen, dtype=torch.float64)
    s = 0.1 * torch.randn(n, generator=gen, dtype=torch.float64)
    so = KSOracle(ks_params([Lk] * b, resolution=n, dt=0.1, time_scheme="ETRK2"))
    so.correction_s = s
    ref = so.step(u, so.domain_size)
    sv = KSSolverTorch(KSSimParams([Lk], resolution=n, dt=0.1, time_scheme="ETRK2", device="cuda"))
    sg = s.to(cuda).requires_grad_(True)
    sv.correction_s = sg
    out = sv.step(u.to(cuda), torch.tensor([Lk], device=cuda))
    assert rel_l2(out, ref) < 1e-11
    out.sum().backward()
    assert sg.grad.shape == (n,)
    sv.correction_s = torch.zeros(b, 1, device=cuda, dtype=torch.float64)
    with pytest.raises(ValueError, match="correction_s"):
        sv.step(u.to(cuda), torch.tensor([Lk], device=cuda))


def test_ks_domain_size_cache_follows_values(cuda):
    # a solver reused across batches with a fresh device tensor of domain
    # sizes each time (domain_size = batch[2].to(device)) must use the new
    # values, whatever address the allocator hands the new tensor, and an
    # in-place edit of the tensor must be seen too
    gen = torch.Generator().manual_seed(9)
    n = 64
    u = 0.3 * torch.randn(2, n, generator=gen, dtype=torch.float64)
    sv = KSSolverTorch(KSSimParams([22.0], resolution=n, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False

    def want(Lk):
        so = KSOracle(ks_params([Lk] * 2, resolution=n, dt=0.25, time_scheme="ETDRK4"))
        so.doublestep = False
        return so.step(u, so.domain_size)

    for Lk in (22.0, 30.0, 35.0, 22.0):
        d = torch.tensor([Lk], dtype=torch.float64).to(cuda)
        assert rel_l2(sv.step(u.to(cuda), d), want(Lk)) < 1e-11
        del d
    d = torch.tensor([Lk], dtype=torch.float64).to(cuda)
    sv.step(u.to(cuda), d)
    d.fill_(41.0)
    assert rel_l2(sv.step(u.to(cuda), d), want(41.0)) < 1e-11
    # host values (a list) are keyed by value
    assert rel_l2(sv.step(u.to(cuda), [27.0]), want(27.0)) < 1e-11


def test_burgers_forcing_cache_sees_in_place_edits(cuda):
    n, b = 64, 2
    A = torch.full((b, 1, 1), 0.5, dtype=torch.float64, device=cuda)
    om = torch.ones(b, 1, 1, dtype=torch.float64, device=cuda)
    ph = torch.zeros(b, 1, 1, dtype=torch.float64, device=cuda)
    ll = torch.ones(b, 1, 1, dtype=torch.float64, device=cuda)
    sv = BurgersSolverTorch(BGSimParams(dt=0.01, resolution=n, batch_size=b, equation_type="learning",
                                        A=A, omega=om, phi=ph, l=ll, device="cuda"))
    u = torch.zeros(b, n, dtype=torch.float64, device=cuda)
    sv.t = 0.5
    first = sv._advance(u, 0.01)
    A.mul_(3.0)
    second = sv._advance(u, 0.01)
    assert rel_l2(second, 3.0 * first) < 1e-13
    nu = torch.full((b, 1), 0.01, dtype=torch.float64, device=cuda)
    sv.nu = nu
    x = torch.arange(n, dtype=torch.float64, device=cuda) * (2 * math.pi / n)
    v = torch.sin(x).expand(b, n).contiguous()
    a = sv._advance(v, 0.01)
    nu.fill_(0.05)
    assert rel_l2(sv._advance(v, 0.01), a) > 1e-6


def test_ragged_advance_checks_correction_size_and_dtype(cuda):
    rb = RaggedBurgers([64, 128], [0.01, 0.01], [0.001, 0.001], dtype=torch.float32)
    u = torch.zeros(rb.total, device=cuda)
    with pytest.raises(ValueError, match="correction"):
        rb.advance(u, torch.zeros(64, device=cuda))
    with pytest.raises(TypeError):
        rb.advance(u.double())
    with pytest.raises(TypeError):
        rb.advance(u, torch.zeros(rb.total, device=cuda, dtype=torch.float64))


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
