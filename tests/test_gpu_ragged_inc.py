"""GPU parity of the ragged correction-network path: the ragged-rows
convolutions against torch's FP64 convolution row by row, and the INC rollout
of a ragged batch in one library call (inc_ragged_inc_fwd / _adj) against the
per-resolution-group path (the network per group and one ragged launch per
step, whose pieces are pinned elsewhere) and against FP64 central differences.

Tolerances: float64 1e-12 relative for the convolutions and 1e-10 for the
rollout and its gradients (same arithmetic, different launch shapes and
summation order of the weight gradient); float32 1e-5 for the convolutions,
and for the rollout against the float64 rollout of the same inputs 1e-5
(trajectory) and 1e-4 (parameter gradients); the float32 gradient of the
initial state is held to the per-group path's own float32 error (below 2e-3,
see the test).
"""
import ctypes
import math

import pytest
import torch
from torch.nn import functional as F

import inc_amd._lib as L
from inc_amd import RaggedBurgers
from inc_amd.batched import stable_dt
from inc_amd.correction import CorrectionCNN
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def ref_conv(x, w, b):
    p = w.shape[-1] // 2
    return F.conv1d(F.pad(x, (p, p), mode="circular"), w, b)


def u0_field(n, seed, kmax=8):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        u += torch.randn(1, generator=gen, dtype=torch.float64) / k * torch.sin(
            k * x + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
    return u / u.abs().max()


def _pack_channels(rb, rows):
    """Per-row [C, n] tensors (internal row order) -> packed [C * total]."""
    return torch.cat([r.reshape(-1) for r in rows])


def _unpack_channels(rb, packed, c):
    out, o = [], 0
    for n in rb.n:
        out.append(packed[o:o + c * n].view(c, n))
        o += c * n
    return out


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(1, 32, 5), (32, 32, 5), (32, 1, 5), (48, 48, 7), (16, 8, 3), (32, 16, 1)],
                         ids=lambda s: "x".join(map(str, s)))
def test_rows_convolutions_match_each_row(cuda, dtype, shape):
    ci, co, k = shape
    ns = [64, 1024, 16, 128, 100, 256, 300, 8]
    rb = RaggedBurgers(ns, [0.01] * len(ns), [1e-3] * len(ns), dtype=dtype)
    gen = torch.Generator().manual_seed(sum(shape))
    xs = [torch.randn(ci, n, generator=gen, dtype=torch.float64) for n in rb.n]
    gs = [torch.randn(co, n, generator=gen, dtype=torch.float64) for n in rb.n]
    ms = [torch.rand(ci, n, generator=gen, dtype=torch.float64) for n in rb.n]
    w = torch.randn(co, ci, k, generator=gen, dtype=torch.float64) / (ci * k) ** 0.5
    b = torch.randn(co, generator=gen, dtype=torch.float64)
    c = lambda t: t.to(cuda, dtype).contiguous()  # noqa: E731
    # every device tensor a launch reads is held by a name until the results
    # are read (a temporary's block could be handed to the next copy first)
    x, g, m = c(_pack_channels(rb, xs)), c(_pack_channels(rb, gs)), c(_pack_channels(rb, ms))
    wd, bd = c(w), c(b)
    rows, _ = rb._rows(1)
    lib = L.load()
    code = L.dtype_code(dtype)
    y = torch.empty(co * rb.total, device=cuda, dtype=dtype)
    dy = torch.empty_like(y)
    L.check(lib.inc_conv1d_rows_fwd_act(code, L.INC_ACT_GELU, ctypes.byref(rows), ci, co, k, L.ptr(x), L.ptr(wd),
                                        L.ptr(bd), L.ptr(y), L.ptr(dy), L.stream()), "rows fwd")
    gx = torch.empty(ci * rb.total, device=cuda, dtype=dtype)
    L.check(lib.inc_conv1d_rows_bwd_data_mul(code, ctypes.byref(rows), ci, co, k, L.ptr(g), L.ptr(wd), L.ptr(m),
                                             L.ptr(gx), L.stream()), "rows bwd data")
    gw, gb = torch.empty(co, ci, k, device=cuda, dtype=dtype), torch.empty(co, device=cuda, dtype=dtype)
    work = torch.empty(512 * (co * ci * k + co), device=cuda, dtype=dtype)
    L.check(lib.inc_conv1d_rows_bwd_weight(code, ctypes.byref(rows), ci, co, k, L.ptr(g), L.ptr(x), L.ptr(gw),
                                           L.ptr(gb), L.ptr(work), work.numel(), L.stream()), "rows bwd weight")
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    ys, dys, gxs = _unpack_channels(rb, y, co), _unpack_channels(rb, dy, co), _unpack_channels(rb, gx, ci)
    gw_ref, gb_ref = torch.zeros_like(w), torch.zeros_like(b)
    for r in range(rb.count):
        v = ref_conv(xs[r][None], w, b)[0].requires_grad_(True)
        (dv,) = torch.autograd.grad(F.gelu(v).sum(), v)
        assert rel_l2(ys[r], F.gelu(v).detach()) < tol, r
        assert rel_l2(dys[r], dv) < tol, r
        xr = xs[r][None].clone().requires_grad_(True)
        wr, br = w.clone().requires_grad_(True), b.clone().requires_grad_(True)
        (ref_conv(xr, wr, br)[0] * gs[r]).sum().backward()
        assert rel_l2(gxs[r], xr.grad[0] * ms[r]) < tol, r
        gw_ref += wr.grad
        gb_ref += br.grad
    assert rel_l2(gw, gw_ref) < tol and rel_l2(gb, gb_ref) < tol


def _inc_case(count, steps, dtype, seed=9, width=32, k=5, depth=4):
    gen = torch.Generator().manual_seed(seed)
    choices = [64, 128, 256, 512, 1024]
    ns = [choices[int(i)] for i in torch.randint(0, 5, (count,), generator=gen)]
    nus = torch.exp(torch.empty(count, dtype=torch.float64).uniform_(math.log(0.003), math.log(0.03),
                                                                      generator=gen)).tolist()
    dts = [stable_dt(n, nu) for n, nu in zip(ns, nus)]
    hz = torch.randint(max(1, steps // 4), steps + 1, (count,), generator=gen).tolist()
    rb = RaggedBurgers(ns, nus, dts, horizon=hz, dtype=dtype)
    u0 = rb.pack([u0_field(n, 1000 + b) for b, n in enumerate(ns)])
    cnn = CorrectionCNN(depth, width, k, generator=torch.Generator().manual_seed(seed + 1), scale=0.01)
    cnn = cnn.to(cuda_dev(), dtype)
    with torch.no_grad():
        target = RaggedBurgers(ns, [1.2 * v for v in nus], dts, horizon=hz, dtype=dtype).rollout(u0, steps)
    mask = rb.valid_mask(steps)
    return rb, u0, cnn, target, mask


def cuda_dev():
    return torch.device("cuda")


def _loss_and_grads(rb, u0, cnn, target, mask, steps, fused, weights=None):
    """The trainer's masked MSE to the target, or with ``weights`` the linear
    functional sum(weights * traj), whose cotangent is the same in every
    precision (the MSE's 2 (traj - target) / M suffers cancellation, which
    would measure the target's rounding rather than the adjoint's)."""
    u = u0.clone().requires_grad_(True)
    cnn.zero_grad(set_to_none=True)
    traj = rb.rollout(u, steps, model=cnn, fused=fused)
    if weights is None:
        loss = (((traj - target) ** 2) * mask).sum() / mask.sum()
    else:
        loss = (traj * weights).sum()
    loss.backward()
    return traj.detach(), loss.detach(), u.grad, [p.grad.clone() for p in cnn.parameters()]


def test_ragged_inc_rollout_one_call_matches_per_group_path(cuda):
    # float64: the one-call path against the per-group path (1e-10); float32:
    # the one-call path against the float64 result of the same inputs, at
    # the float32 accuracy bar (trajectory 1e-5, gradients 1e-4)
    steps = 10
    res = {}
    for dtype in (torch.float64, torch.float32):
        rb, u0, cnn, target, mask = _inc_case(24, steps, dtype)
        wts = torch.randn(steps, rb.total, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
        wts = wts.to(cuda, dtype) * mask / rb.total
        res[dtype] = _loss_and_grads(rb, u0, cnn, target, mask, steps, True, weights=wts)
        if dtype == torch.float64:
            b = _loss_and_grads(rb, u0, cnn, target, mask, steps, False, weights=wts)
            a = res[dtype]
            assert rel_l2(a[0], b[0]) < 1e-10
            assert rel_l2(a[1].reshape(1), b[1].reshape(1)) < 1e-10
            assert rel_l2(a[2], b[2]) < 1e-10
            for ga, gb in zip(a[3], b[3]):
                assert rel_l2(ga, gb) < 1e-10
            m1 = _loss_and_grads(rb, u0, cnn, target, mask, steps, True)
            m2 = _loss_and_grads(rb, u0, cnn, target, mask, steps, False)
            for x1, x2 in zip(m1[1:3], m2[1:3]):
                assert rel_l2(x1.reshape(-1), x2.reshape(-1)) < 1e-10
            for ga, gb in zip(m1[3], m2[3]):
                assert rel_l2(ga, gb) < 1e-10
            # trajectories past their horizon carry their state
            per = rb.unpack(a[0])
            for i, bidx in enumerate(rb.perm):
                h = rb.horizon[i]
                if h < steps:
                    assert torch.equal(per[bidx][h:], per[bidx][h - 1:h].expand(steps - h, -1))
            # the evaluation form (no gradient) gives the same trajectory
            with torch.no_grad():
                ev = rb.rollout(u0, steps, model=cnn)
            assert torch.equal(ev, a[0])
    a, b = res[torch.float32], res[torch.float64]
    assert rel_l2(a[0], b[0]) < 1e-5
    for ga, gb in zip(a[3], b[3]):
        assert rel_l2(ga, gb) < 1e-4
    # The float32 gradient of u0 is as far from float64 through the
    # per-group path (5.5e-4 measured; 1.1e-4 without the network): the
    # WENO weights' derivatives make the step's Jacobian sensitive to the
    # state's rounding on the viscous-limited n = 512 / 1024 trajectories.
    # The one-call path is held to the per-group path's own error.
    rb, u0, cnn, target, mask = _inc_case(24, steps, torch.float32)
    wts = torch.randn(steps, rb.total, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
    wts = wts.to(cuda, torch.float32) * mask / rb.total
    c = _loss_and_grads(rb, u0, cnn, target, mask, steps, False, weights=wts)
    e_one, e_group = rel_l2(a[2], b[2]), rel_l2(c[2], b[2])
    assert e_one < 2e-3 and e_one < 1.2 * e_group, (e_one, e_group)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ragged_inc_stored_activations_match_recompute(cuda, dtype):
    # uneven batch with horizons: the forward's stored activations (rows past
    # a horizon never written, never read) give the adjoint of the network's
    # repeat bit for bit
    steps = 9
    rb, u0, cnn, target, mask = _inc_case(30, steps, dtype, seed=17, width=48, k=7, depth=5)
    out = {}
    for keep in (False, True):
        rb.keep_activations = keep
        out[keep] = _loss_and_grads(rb, u0, cnn, target, mask, steps, True)
    a, b = out[True], out[False]
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) and torch.equal(a[2], b[2])
    for i, (ga, gb) in enumerate(zip(a[3], b[3])):
        assert torch.equal(ga, gb), i


def test_ragged_inc_rollout_gradient_central_differences(cuda):
    steps = 6
    rb, u0, cnn, target, mask = _inc_case(12, steps, torch.float64, seed=4, width=16, k=3, depth=3)

    def loss_fn(u):
        traj = rb.rollout(u, steps, model=cnn)
        return (((traj - target) ** 2) * mask).sum() / mask.sum()

    u = u0.clone().requires_grad_(True)
    cnn.zero_grad(set_to_none=True)
    loss_fn(u).backward()
    params = list(cnn.parameters())
    gen = torch.Generator().manual_seed(5)
    # unit directions, h = 1e-5: the loss's own rounding (~1e-16 of it) over
    # 2h stays far below the directional derivatives
    h = 1e-5
    with torch.no_grad():
        for _ in range(3):
            d = torch.randn(rb.total, generator=gen, dtype=torch.float64).to(cuda)
            d /= d.norm()
            fd = (loss_fn(u0 + h * d) - loss_fn(u0 - h * d)) / (2 * h)
            an = (u.grad * d).sum()
            assert abs(float(fd - an)) <= 1e-6 * abs(float(an)) + 1e-13, (float(fd), float(an))
        # unit directions in each parameter tensor, h = 1e-5: the loss's own
        # rounding (~1e-16 of it) over 2h stays far below the derivatives
        hp = 1e-5
        for p in params:
            d = torch.randn(p.shape, generator=gen, dtype=torch.float64).to(cuda)
            d /= d.norm()
            p0 = p.detach().clone()
            p.copy_(p0 + hp * d)
            fp = loss_fn(u0)
            p.copy_(p0 - hp * d)
            fm = loss_fn(u0)
            p.copy_(p0)
            fd = (fp - fm) / (2 * hp)
            an = (p.grad * d).sum()
            assert abs(float(fd - an)) <= 1e-6 * abs(float(an)) + 1e-13, (float(fd), float(an))
