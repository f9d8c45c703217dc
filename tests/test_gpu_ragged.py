"""GPU parity of the ragged (uneven-batch) Burgers launch against the uniform
per-trajectory solver, and of a HIP-graph-captured INC rollout against eager.

Tolerances: float64 ragged results agree with the uniform float64 kernels
(themselves pinned to the FP64 oracle) to 1e-12 relative; graph replays agree
with eager execution to 1e-6 (same kernels; the CNN's MIOpen algorithms may
differ between eager and captured runs).
"""
import math

import pytest
import torch

from inc_amd import BGSimParams, BurgersSolverTorch, RaggedBurgers
from inc_amd.batched import stable_dt
from inc_amd.correction import CorrectionCNN
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def u0_field(n, seed, kmax=8):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        u += torch.randn(1, generator=gen, dtype=torch.float64) / k * torch.sin(
            k * x + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
    return u / u.abs().max()


def mixed_batch(ns, seed=0):
    gen = torch.Generator().manual_seed(seed)
    nus = torch.exp(torch.empty(len(ns), dtype=torch.float64).uniform_(math.log(0.003), math.log(0.03),
                                                                        generator=gen)).tolist()
    dts = [stable_dt(n, nu) for n, nu in zip(ns, nus)]
    return nus, dts


@pytest.mark.parametrize("scheme", ["SSPRK3", "Euler"])
def test_ragged_matches_uniform_solver_fp64(cuda, scheme):
    ns = [64, 1024, 16, 128, 100, 256, 64, 512]
    hz = [5, 3, 5, 1, 4, 5, 2, 5]
    nus, dts = mixed_batch(ns)
    if scheme == "Euler":
        dts = [0.5 * d for d in dts]
    steps = 5
    rb = RaggedBurgers(ns, nus, dts, horizon=hz, scheme=scheme, dtype=torch.float64)
    fields = [u0_field(n, 100 + b) for b, n in enumerate(ns)]
    ss = [0.1 * u0_field(n, 200 + b) for b, n in enumerate(ns)]
    w = torch.randn(steps, rb.total, generator=torch.Generator().manual_seed(3), dtype=torch.float64).to(cuda)
    u = rb.pack(fields).requires_grad_(True)
    s = rb.pack(ss).requires_grad_(True)
    traj = rb.advance(u, s, steps=steps, want_traj=True)
    (traj * w).sum().backward()
    per_traj = rb.unpack(traj.detach())
    w_per = rb.unpack(w)
    g_u, g_s = rb.unpack(u.grad), rb.unpack(s.grad)
    for b, n in enumerate(ns):
        p = BGSimParams(dt=dts[b], resolution=n, nu=nus[b], batch_size=1, time_scheme=scheme, device="cuda")
        sv = BurgersSolverTorch(p)
        ub = fields[b].to(cuda)[None].requires_grad_(True)
        sb = ss[b].to(cuda)[None].requires_grad_(True)
        sv.correction_s = sb
        cur, outs = ub, []
        for k in range(steps):
            cur = sv.step(cur) if k < hz[b] else cur
            outs.append(cur)
        ref = torch.stack(outs, 1)[0]
        (ref * w_per[b]).sum().backward()
        assert rel_l2(per_traj[b], ref.detach()) < 1e-12, b
        assert rel_l2(g_u[b], ub.grad[0]) < 1e-11, b
        assert rel_l2(g_s[b], sb.grad[0]) < 1e-11, b


def test_ragged_team_sizes_agree(cuda):
    ns = [1024, 64, 256, 8, 512]
    nus, dts = mixed_batch(ns, seed=4)
    fields = [u0_field(n, 10 + b) for b, n in enumerate(ns)]
    outs = []
    for team in (64, 128, 256):
        rb = RaggedBurgers(ns, nus, dts, team=team, dtype=torch.float32)
        outs.append(rb.advance(rb.pack(fields), None, steps=3))
    assert rel_l2(outs[1], outs[0]) < 1e-6 and rel_l2(outs[2], outs[0]) < 1e-6


def _inc_setup(count=48, steps=12, seed=9, dtype=torch.float32):
    gen = torch.Generator().manual_seed(seed)
    choices = [64, 128, 256, 512, 1024]
    ns = [choices[int(i)] for i in torch.randint(0, 5, (count,), generator=gen)]
    nus, dts = mixed_batch(ns, seed)
    hz = torch.randint(steps // 4, steps + 1, (count,), generator=gen).tolist()
    rb = RaggedBurgers(ns, nus, dts, horizon=hz, dtype=dtype)
    u0 = rb.pack([u0_field(n, 1000 + b) for b, n in enumerate(ns)])
    with torch.no_grad():
        rb_t = RaggedBurgers(ns, [1.2 * v for v in nus], dts, horizon=hz, dtype=dtype)
        target = rb_t.rollout(u0, steps)
    cnn = CorrectionCNN(generator=torch.Generator().manual_seed(0)).cuda().to(dtype)
    mask = rb.valid_mask(steps)

    def loss_fn(u_init):
        traj = rb.rollout(u_init, steps, model=lambda v: 0.01 * cnn(v))
        return (((traj - target) ** 2) * mask).sum() / mask.sum()
    return rb, u0, cnn, loss_fn


def test_ragged_inc_rollout_gradient_fd(cuda):
    # float64 throughout (ragged kernels and CNN): central differences
    rb, u0, cnn, loss_fn = _inc_setup(count=12, steps=6, dtype=torch.float64)
    u = u0.clone().requires_grad_(True)
    loss_fn(u).backward()
    gen = torch.Generator().manual_seed(5)
    with torch.no_grad():
        for _ in range(4):
            d = torch.randn(rb.total, generator=gen, dtype=torch.float64).to(cuda)
            h = 1e-6
            fd = (loss_fn(u0 + h * d) - loss_fn(u0 - h * d)) / (2 * h)
            an = (u.grad * d).sum()
            assert abs(float(fd - an)) <= 1e-6 * abs(float(an)) + 1e-12


@pytest.mark.skip(reason="capturing the per-group INC rollout (then with MIOpen convolutions, autograd backward "
                         "and solver launches) segfaulted inside hipStreamEndCapture; no cause shown, so the case "
                         "stays off the GPU and RaggedBurgers.rollout refuses it under capture; the one-call path "
                         "is captured in tests/test_gpu_graph.py")
def test_ragged_inc_rollout_hip_graph_replays_match_eager(cuda):
    rb, u0, cnn, loss_fn = _inc_setup()
    params = list(cnn.parameters())
    # eager reference
    u_e = u0.clone().requires_grad_(True)
    loss_e = loss_fn(u_e)
    loss_e.backward()
    g_e = [p.grad.clone() for p in params]
    gu_e = u_e.grad.clone()
    # capture forward + backward
    static_u = u0.clone().requires_grad_(True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            for p in params:
                p.grad = None
            static_u.grad = None
            loss_fn(static_u).backward()
    torch.cuda.current_stream().wait_stream(side)
    for p in params:
        p.grad = None
    static_u.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        static_loss = loss_fn(static_u)
        static_loss.backward()
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert rel_l2(static_loss.detach().reshape(1), loss_e.detach().reshape(1)) < 1e-6, rep
        assert rel_l2(static_u.grad, gu_e) < 1e-5, rep
        for a, b in zip(params, g_e):
            assert rel_l2(a.grad, b) < 1e-5, rep
        # eager work between replays
        _ = loss_fn(u0.clone()).item()
