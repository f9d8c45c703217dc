"""Benchmark: KS ETDRK4 forward+adjoint rollouts on one MI355X.

Workload (plain KS, fp32 state, ETDRK4 with 2/3 dealiasing): B = 256
independent trajectories, N = 256, L = 64, dt = 0.25, 1000 steps, seeded
u0 = sum_{k=1..10} b_k cos(2 pi k x / L + psi_k), b_k ~ N(0, 0.5^2).
One bench step = one forward rollout of 1000 steps (every step's state
written, as simulate_rollout returns them) plus the adjoint rollout of a
trajectory cotangent back to u0, through the drop-in API
(``simulate_rollout`` + ``torch.autograd.backward``).

Prints one JSON line.  value = trajectory-steps per second, forward+adjoint
(B * 1000 per bench step).  roofline: the dominant kernel's algorithmic HBM
bytes (adjoint: read u^n, read lambda^{n+1}, write lambda^n = 3*4*N bytes
per trajectory-step; forward: 2*4*N) over its duration measured with HIP
events on the launch stream, against 8 TB/s, next to a measured 1 GiB
device copy (measured_copy_GBps).  compute: the same kernel's FP32 flops per
launch (rocprofv3 SQ_INSTS_VALU_FLOPS_FP32, recorded under profiles/) over
its duration, against the 157.3 TFLOP/s vector peak.  cpu_baseline: the FP64
oracle (PyTorch CPU restatement of the same scheme) on a bounded sample,
fwd+adj.  other_workloads (not the metric): the same kernels at saturation,
the rollout with its MSE loss fused in against torch's MSELoss, Burgers at
saturation, the INC training steps (Burgers with the correction network inside
the rollout and adjoint kernels, against the per-step path; KS with
checkpointing; ragged Burgers), and the Burgers INC evaluation rollout with the
network inside the kernel.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B, N, L, DT, H = 256, 256, 64.0, 0.25, 1000
HBM_PEAK_GBS = 8000.0


def initial_state(b, n, L, seed, kmax=10):
    import torch
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        amp = torch.randn(b, 1, generator=gen, dtype=torch.float64) * 0.5
        ph = torch.rand(b, 1, generator=gen, dtype=torch.float64) * 2 * math.pi
        u += amp * torch.cos(2 * math.pi * k * x / L + ph)
    return u


VALU_PEAK_TFS = 157.3  # fp32 vector, packed-FP32 rate (the published figure)


def read_profile_value(name, key):
    """A figure recorded in a committed rocprofv3 summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def read_traffic():
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    return read_profile_value("ks_adj_traffic.md", "traffic_bytes_per_launch")


def copy_bandwidth(dev, nbytes=1 << 30, reps=5):
    """Device-to-device copy of nbytes, best of reps: read + write bytes per
    second (a STREAM-copy measurement of the box's HBM)."""
    import torch
    src = torch.empty(nbytes // 4, device=dev)
    src.fill_(1.0)
    dst = torch.empty_like(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        a.record()
        dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b) * 1e-3)
    del src, dst
    return 2 * nbytes / best / 1e9


def cpu_baseline(budget_s=12.0):
    import torch
    from oracle import KSOracle, ks_params
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    so = KSOracle(ks_params([L] * B, resolution=N, dt=DT, time_scheme="ETDRK4"))
    so.doublestep = False
    u0 = initial_state(B, N, L, 1001)

    def run(h):
        u = u0.clone().requires_grad_(True)
        cur, outs = u, []
        for _ in range(h):
            cur = so.step(cur, so.domain_size)
            outs.append(cur)
        torch.autograd.backward(torch.stack(outs, 1), torch.ones(B, h, N, dtype=torch.float64))
        return u.grad

    t0 = time.perf_counter()
    run(2)
    t1 = time.perf_counter()
    h = max(2, min(200, int(budget_s / max(t1 - t0, 1e-3) * 2)))
    t0 = time.perf_counter()
    run(h)
    dt = time.perf_counter() - t0
    return {"value": B * h / dt, "unit": "trajectory-steps/s (fwd+adj)", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (torch CPU) ETDRK4, B={B}, N={N}, {h} steps fwd+adj, {dt:.1f} s"}


def _timed(fn, reps=3):
    """Median wall time (s) of fn() over reps runs after one warm-up."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def sine_state(n, seed, kmax=8):
    """u0 = sum_{k<=8} a_k sin(k x + phi_k), a_k ~ N(0,1)/k, scaled to max|u0| = 1."""
    import torch
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        u += torch.randn(1, generator=gen, dtype=torch.float64) / k * torch.sin(
            k * x + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
    return u / u.abs().max()


def extra_burgers_saturation(b=8192, n=256, h=200):
    """The Burgers solver kernels alone (SSP-RK3 + WENO5, fp32, nu=0.01, dt=0.005, no correction) at a
    batch that fills the SIMDs: forward rollout with every state written, adjoint from a trajectory
    cotangent; algorithmic bytes 2*4*N (forward) and 3*4*N (adjoint) per trajectory-step."""
    import torch
    import inc_amd
    dev = torch.device("cuda")
    u0 = torch.stack([sine_state(n, 3000 + i) for i in range(b)]).float().to(dev)
    cot = torch.randn(b, h, n, device=dev) * 1e-3
    sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.005, resolution=n, nu=0.01, batch_size=b,
                                                        time_scheme="SSPRK3", device="cuda"))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]

    def run():
        u = u0.clone().requires_grad_(True)
        ev[0].record()
        traj = sv.rollout(u, h)
        ev[1].record()
        torch.autograd.backward(traj, cot)
        ev[2].record()
    run()
    torch.cuda.synchronize()
    tf = ta = float("inf")
    for _ in range(3):
        run()
        torch.cuda.synchronize()
        tf = min(tf, ev[0].elapsed_time(ev[1]) * 1e-3)
        ta = min(ta, ev[1].elapsed_time(ev[2]) * 1e-3)
    ts = b * h
    return {"workload": f"Burgers SSP-RK3/WENO5 fwd+adj B={b} N={n} {h} steps fp32",
            "trajectory_steps_per_s": ts / (tf + ta), "fwd_GBps_algorithmic": 2 * 4 * n * ts / tf / 1e9,
            "adj_GBps_algorithmic": 3 * 4 * n * ts / ta / 1e9, "fwd_ms": tf * 1e3, "adj_ms": ta * 1e3}


def extra_burgers_inc():
    """Burgers INC fwd+adj: B=512, N=128, nu=0.01, dt=0.01, 16 steps, SSP-RK3 + WENO5,
    correction from a 4-layer 32-channel k=5 circular CNN, MSE to a nu=0.012 target.
    The network runs inside the rollout and adjoint kernels (the default path);
    the per-step path (convolution kernels and one solver launch per step) is
    timed beside it."""
    import torch
    import inc_amd
    from inc_amd.correction import CorrectionCNN
    b, n, h = 512, 128, 16
    dev = torch.device("cuda")
    u0 = torch.stack([sine_state(n, 2000 + i) for i in range(b)]).float().to(dev)
    cnn = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2)).to(dev)
    with torch.no_grad():
        tgt_solver = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.01, resolution=n, nu=0.012, batch_size=b,
                                                                    time_scheme="SSPRK3", device="cuda"))
        target = tgt_solver.rollout(u0, h)

    def train_step(fused):
        sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b,
                                                            time_scheme="SSPRK3", device="cuda"))
        res, _ = inc_amd.simulate_rollout(cnn, sv, u0, h, "INC", fused=fused)
        loss = ((res - target) ** 2).mean()
        loss.backward()
        cnn.zero_grad(set_to_none=True)
    t = _timed(lambda: train_step(True))
    tp = _timed(lambda: train_step(False))

    def train_step_mse():
        sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b,
                                                            time_scheme="SSPRK3", device="cuda"))
        loss, _, _ = inc_amd.simulate_rollout_mse(sv, u0, target, model=cnn)
        loss.backward()
        cnn.zero_grad(set_to_none=True)
    tm = _timed(train_step_mse)
    # the fused kernels, timed with HIP events on the launch stream: the
    # forward (burgers_inc_fwd_kernel) and the adjoint (burgers_inc_adj_kernel)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    tf = ta = float("inf")
    for _ in range(4):
        sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b,
                                                            time_scheme="SSPRK3", device="cuda"))
        u = u0.clone().requires_grad_(True)
        ev[0].record()
        res, _ = inc_amd.simulate_rollout(cnn, sv, u, h, "INC")
        ev[1].record()
        g = (2.0 / res.numel()) * (res.detach() - target)
        ev[2].record()
        torch.autograd.backward(res, g)
        ev[3].record()
        torch.cuda.synchronize()
        tf = min(tf, ev[0].elapsed_time(ev[1]) * 1e-3)
        ta = min(ta, ev[2].elapsed_time(ev[3]) * 1e-3)
        cnn.zero_grad(set_to_none=True)
    pts = b * n * h
    # matrix-core work of the adjoint: the hidden 32 -> 32 layers evaluated again, their input
    # gradients and their weight gradients (3 x 2 x 32 x 32 x 5 x 2 flops per point-step)
    mflops = 3 * 2 * (2 * 32 * 32 * 5) * pts
    return {"workload": "Burgers INC fwd+adj B=512 N=128 16 steps SSP-RK3/WENO5 + CNN(4x32,k5)",
            "trajectory_steps_per_s": b * h / t, "ms_per_train_step": t * 1e3,
            "per_step_path_ms_per_train_step": tp * 1e3, "fused_over_per_step": tp / t,
            "loss_fused_ms_per_train_step": tm * 1e3,
            "fwd_kernel_ms": tf * 1e3, "adj_kernel_ms": ta * 1e3,
            "roofline": {"bound": "mfma", "kernel": "burgers_inc_adj_kernel<32,5,2,2>",
                         "achieved": mflops / ta / 1e12, "peak": MFMA_F32_PEAK_TFS, "unit": "TFLOP/s",
                         "frac": mflops / ta / 1e12 / MFMA_F32_PEAK_TFS, "flops_per_launch": mflops,
                         "hbm_algorithmic_bytes": 4 * 4 * n * b * h,
                         "hbm_frac": 4 * 4 * n * b * h / ta / 1e9 / HBM_PEAK_GBS}}


def extra_burgers_inc_eval(b=512, n=128, h=1000):
    """Burgers INC evaluation rollout (no gradient, as the trainer's validation
    and test passes run it): B=512, N=128, 1000 SSP-RK3/WENO5 steps with the
    4-layer 32-channel k=5 correction network, evaluated inside the rollout
    kernel against the per-step path (network + solver launch per step).  The
    network's last layer is scaled by 0.01 (the correction stays small
    against the dynamics, as a trained corrector's does): at its He-normal
    initialisation a few trajectories of the 1000-step rollout diverge.  The
    final states are checked to be finite."""
    import torch
    import inc_amd
    from inc_amd.correction import CorrectionCNN
    dev = torch.device("cuda")
    u0 = torch.stack([sine_state(n, 5000 + i) for i in range(b)]).float().to(dev)
    cnn = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2)).to(dev)
    with torch.no_grad():
        cnn.convs[-1].weight.mul_(0.01)
        cnn.convs[-1].bias.mul_(0.01)

    def solver():
        return inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b,
                                                              time_scheme="SSPRK3", device="cuda"))

    def fused():
        with torch.no_grad():
            return inc_amd.simulate_rollout(cnn, solver(), u0, h, "INC")

    def per_step():
        with torch.no_grad():
            sv, u = solver(), u0
            for _ in range(h):
                sv.correction_s = cnn(u)
                u = sv.step(u)
    tf, tp = _timed(fused, reps=3), _timed(per_step, reps=3)
    finite = bool(torch.isfinite(fused()[0]).all())
    return {"workload": f"Burgers INC evaluation rollout B={b} N={n} {h} steps SSP-RK3/WENO5 + CNN(4x32,k5), "
                        "no gradient", "in_kernel_ms": tf * 1e3, "per_step_ms": tp * 1e3, "finite": finite,
            "in_kernel_trajectory_steps_per_s": b * h / tf, "per_step_trajectory_steps_per_s": b * h / tp}


def extra_ks_inc(step_path=True):
    """KS INC fwd+adj: L=100, N=512, dt=0.25, B=1024, 64 steps, correction
    from a 6-layer 48-channel k=7 circular CNN (output x 0.01).  The
    CorrectionCNN goes to simulate_rollout, which runs the rollout in one
    library call per direction (inc_ks_inc_fwd / _adj, every step's hidden
    activations kept for the adjoint: 32 GB); with step_path the
    step-by-step path (the network called from Python every step,
    checkpointed every 8 steps) is timed beside it."""
    import torch
    import inc_amd
    from inc_amd.correction import CorrectionCNN
    b, n, h, Lk = 1024, 512, 64, 100.0
    dev = torch.device("cuda")
    u0 = initial_state(b, n, Lk, 1003, kmax=16).float().to(dev)
    cnn = CorrectionCNN(6, 48, 7, generator=torch.Generator().manual_seed(3), scale=0.01).to(dev)
    Ld = torch.tensor([Lk], device=dev)
    sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([Lk], resolution=n, dt=0.25, time_scheme="ETDRK4",
                                                   device="cuda"))
    sv.doublestep = False
    target = torch.zeros(b, h, n, device=dev)

    def train_step():
        res, _ = inc_amd.simulate_rollout(cnn, sv, u0, h, "INC", L=Ld)
        ((res - target) ** 2).mean().backward()
        cnn.zero_grad(set_to_none=True)

    def train_step_loop():
        res, _ = inc_amd.simulate_rollout(cnn, sv, u0, h, "INC", L=Ld, checkpoint_every=8, fused=False)
        ((res - target) ** 2).mean().backward()
        cnn.zero_grad(set_to_none=True)
    t = _timed(train_step, reps=2)
    t_loop = _timed(train_step_loop, reps=2) if step_path else None
    # dominant kernel, measured live: the hidden 48 -> 48 k7 layer with GELU over the batch
    import ctypes
    import inc_amd._lib as L
    x = torch.randn(b, 48, n, device=dev)
    y, dy = torch.empty_like(x), torch.empty_like(x)
    w = cnn.convs[1].weight.detach().contiguous()
    bias = cnn.convs[1].bias.detach().contiguous()
    lib = L.load()

    def hidden():
        L.check(lib.inc_conv1d_fwd_act(L.INC_F32, L.INC_ACT_GELU, b, 48, 48, n, 7, L.ptr(x), L.ptr(w), L.ptr(bias),
                                       L.ptr(y), L.ptr(dy), L.stream()), "hidden layer")
    hidden()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    kt = []
    for _ in range(10):
        ev[0].record()
        hidden()
        ev[1].record()
        torch.cuda.synchronize()
        kt.append(ev[0].elapsed_time(ev[1]) * 1e-3)
    tk = sorted(kt)[len(kt) // 2]
    flops = 2.0 * 48 * 48 * 7 * b * n
    return {"workload": "KS INC fwd+adj B=1024 N=512 L=100 64 steps ETDRK4, CNN(6x48,k7), one call per direction",
            "trajectory_steps_per_s": b * h / t, "ms_per_train_step": t * 1e3,
            "step_path_ms_per_train_step": t_loop * 1e3 if t_loop else None,
            "roofline": {"bound": "mfma", "kernel": "conv_fwd_mfma_kernel (hidden 48->48 k7 + GELU, B=1024, N=512)",
                         "achieved": flops / tk / 1e12, "peak": MFMA_F32_PEAK_TFS, "unit": "TFLOP/s",
                         "frac": flops / tk / 1e12 / MFMA_F32_PEAK_TFS, "kernel_us": tk * 1e6,
                         "flops_per_launch": flops}}


def extra_reference_net(kind, shapes=((512, 64), (512, 128), (2048, 256)), torch_too=True):
    """One of the reference's correction networks (ResNet1D: 6 residual
    blocks, 32 features; UNet1D: 32 initial features) forward + backward, on
    this library's kernels against the same module run by torch
    (torch.nn.Conv1d, MaxPool1d, ConvTranspose1d, cat)."""
    import torch
    from inc_amd import ResNet1D, UNet1D
    cls = {"resnet1d": ResNet1D, "unet1d": UNet1D}[kind]
    name = "ResNet1D(32 features, 6 blocks)" if kind == "resnet1d" else "UNet1D(32 initial features)"
    out = {"workload": f"{name} forward+backward, fp32"}
    for b, n in shapes:
        torch.manual_seed(0)
        hip = cls().cuda()
        ref = cls(backend="torch").cuda()
        ref.load_state_dict(hip.state_dict())
        u = torch.randn(b, n, device="cuda")

        def step(m):
            m(u, 2 * math.pi).square().sum().backward()
        t_hip = _timed(lambda: step(hip), reps=10)
        out[f"B{b}_N{n}"] = {"ms": t_hip * 1e3}
        if torch_too:
            t_ref = _timed(lambda: step(ref), reps=10)
            out[f"B{b}_N{n}"].update({"torch_ms": t_ref * 1e3, "speedup": t_ref / t_hip})
    return out


MFMA_F32_PEAK_TFS = 157.3  # dense fp32 matrix peak


def ragged_workload(count=2048, steps=64, seed=1004):
    """The uneven INC batch: `count` Burgers trajectories, N uniform in
    {64, 128, 256, 512, 1024}, nu ~ LogU[0.003, 0.03], dt = min(CFL 0.4,
    0.3 dx^2 / nu) per trajectory (inc_amd.batched.stable_dt), horizons ~ U{16..64}, SSP-RK3 + WENO5, fp32;
    the correction network is CorrectionCNN(4 x 32, k = 5) with its output
    scaled by 0.01; the loss is the MSE to the same batch at 1.2 nu over the
    steps inside each trajectory's horizon."""
    import torch
    gen = torch.Generator().manual_seed(seed)
    choices = [64, 128, 256, 512, 1024]
    ns = [choices[int(i)] for i in torch.randint(0, 5, (count,), generator=gen)]
    nus = torch.exp(torch.empty(count, dtype=torch.float64).uniform_(math.log(0.003), math.log(0.03),
                                                                      generator=gen)).tolist()
    from inc_amd.batched import stable_dt
    dts = [stable_dt(n, nu) for n, nu in zip(ns, nus)]
    hz = torch.randint(16, steps + 1, (count,), generator=gen).tolist()
    return ns, nus, dts, hz


def batched_ragged_inc(plain_sample=16, replays=20):
    """The batched workload of the metric: one training step of the uneven INC
    batch (ragged_workload) -- rollout with the network, masked MSE cotangent,
    adjoint through solver and network -- captured once in a hipGraph and
    replayed (3 warm-up replays, then the median of `replays`); the launch
    shape is autotuned.  Beside it: the same step run eagerly, the
    per-resolution-group path, and the plain per-trajectory path (one
    BurgersSolverTorch launch per trajectory per step, network per step) on a
    sample of trajectories."""
    import torch
    import inc_amd
    from inc_amd import autotune
    from inc_amd.correction import CorrectionCNN
    steps = 64
    ns, nus, dts, hz = ragged_workload(steps=steps)
    count = len(ns)
    rb = inc_amd.RaggedBurgers(ns, nus, dts, horizon=hz)
    tuning = {}
    rb.team = autotune.tune_ragged(rb, record=tuning)
    dev = torch.device("cuda")
    u0 = rb.pack([sine_state(n, 3000 + i) for i, n in enumerate(ns)])
    cnn = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2), scale=0.01).to(dev)
    params = list(cnn.parameters())
    mask = rb.valid_mask(steps)
    msum = float(mask.sum())
    with torch.no_grad():
        target = inc_amd.RaggedBurgers(ns, [1.2 * v for v in nus], dts, horizon=hz, team=rb.team).rollout(u0, steps)

    def train_step():
        traj = rb.rollout(u0, steps, model=cnn)
        torch.autograd.backward(traj, (2.0 / msum) * (traj - target) * mask)
        return traj

    def eager_step(fused=True):
        traj = rb.rollout(u0, steps, model=cnn, fused=fused)
        ((((traj - target) ** 2) * mask).sum() / msum).backward()
        cnn.zero_grad(set_to_none=True)

    # capture: warm-up on a side stream, then one captured training step
    for p in params:
        p.grad = None
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            train_step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        traj_g = train_step()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for r in range(3 + replays):
        a.record()
        graph.replay()
        b.record()
        torch.cuda.synchronize()
        if r >= 3:
            ts.append(a.elapsed_time(b) * 1e-3)
    t_graph = sorted(ts)[len(ts) // 2]
    finite = bool(torch.isfinite(traj_g).all()) and all(bool(torch.isfinite(p.grad).all()) for p in params)
    del graph, traj_g
    for p in params:
        p.grad = None
    t_eager = _timed(eager_step, reps=3)
    t_group = _timed(lambda: eager_step(False), reps=2)
    traj_steps = sum(hz)
    dom = ragged_dominant_kernel(rb, params[2], params[3])
    # plain path on a sample: one solver launch per trajectory per step, the
    # network per step, fwd + adj
    idx = list(range(0, count, count // plain_sample))[:plain_sample]

    def plain_step():
        for i in idx:
            sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=dts[i], resolution=ns[i], nu=nus[i],
                                                                batch_size=1, time_scheme="SSPRK3", device="cuda"))
            res, _ = inc_amd.simulate_rollout(cnn, sv, sine_state(ns[i], 3000 + i).float().to(dev)[None], hz[i],
                                              "INC", fused=False)
            (res ** 2).mean().backward()
        cnn.zero_grad(set_to_none=True)
    tp = _timed(plain_step, reps=1)
    plain_rate = sum(hz[i] for i in idx) / tp
    value = traj_steps / t_graph
    solver_bytes = sum(7 * 4 * n * h for n, h in zip(ns, hz))
    return {
        "metric": "fwd+adjoint trajectory-steps/s, batched (hipGraph-captured, autotuned, uneven batch)",
        "value": value, "unit": "trajectory-steps/s", "ms_per_step": t_graph * 1e3, "dtype": "f32",
        "config": {"workload": f"uneven Burgers INC training step: {count} trajectories, N in {{64..1024}}, "
                               "nu LogU[0.003,0.03], dt min(CFL 0.4, viscous), horizons U{16..64}, SSP-RK3/WENO5, "
                               "CorrectionCNN 4x32 k5 (x0.01), masked MSE", "trajectory_steps": traj_steps,
                   "team": rb.team, "autotune_ms": tuning},
        "eager_ms_per_step": t_eager * 1e3, "per_group_path_ms_per_step": t_group * 1e3,
        "plain_trajectory_steps_per_s": plain_rate, "batched_over_plain": value / plain_rate,
        "finite": finite,
        "roofline": dom,
        "hbm": {"solver_algorithmic_bytes": solver_bytes,
                "GBps_over_step": solver_bytes / t_graph / 1e9,
                "frac": solver_bytes / t_graph / 1e9 / HBM_PEAK_GBS},
    }


def ragged_dominant_kernel(rb=None, w=None, bias=None, reps=20, calibrate_copy=False):
    """The dominant kernel of the uneven INC training step, measured live:
    the hidden 32 -> 32 layer's forward with GELU (conv_fwd_mfma_kernel through
    inc_conv1d_rows_fwd_act) over every trajectory of the batch, as at step 0,
    timed with HIP events on the launch stream (median of `reps`).  The
    algorithmic flops are 2 * 32 * 32 * 5 per point of the batch; the traffic
    is the copy-calibrated FETCH_SIZE + WRITE_SIZE of this launch recorded
    under profiles/."""
    import ctypes
    import torch
    import inc_amd
    import inc_amd._lib as L
    dev = torch.device("cuda")
    if rb is None:
        ns, nus, dts, hz = ragged_workload()
        rb = inc_amd.RaggedBurgers(ns, nus, dts, horizon=hz, team=128)
        gen = torch.Generator().manual_seed(5)
        w = (torch.randn(32, 32, 5, generator=gen) * (2.0 / 160) ** 0.5).to(dev)
        bias = torch.zeros(32, device=dev)
    if calibrate_copy:
        src = torch.randn(64 * 2 ** 20, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rows, _ = rb._rows(1)
    x = torch.randn(32 * rb.total, device=dev)
    y = torch.empty_like(x)
    w = w.detach().contiguous()
    bias = bias.detach().contiguous()
    lib = L.load()

    def hidden():
        L.check(lib.inc_conv1d_rows_fwd_act(L.INC_F32, L.INC_ACT_GELU, ctypes.byref(rows), 32, 32, 5, L.ptr(x),
                                            L.ptr(w), L.ptr(bias), L.ptr(y), None, L.stream()), "hidden layer")
    hidden()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kt = []
    for _ in range(reps):
        a.record()
        hidden()
        b.record()
        torch.cuda.synchronize()
        kt.append(a.elapsed_time(b) * 1e-3)
    t_kernel = sorted(kt)[len(kt) // 2]
    flops = 2.0 * 32 * 32 * 5 * rb.total
    return {"bound": "mfma", "kernel": "conv_fwd_mfma_kernel<2,2> (hidden 32->32 layer + GELU, every trajectory)",
            "achieved": flops / t_kernel / 1e12, "peak": MFMA_F32_PEAK_TFS, "unit": "TFLOP/s",
            "frac": flops / t_kernel / 1e12 / MFMA_F32_PEAK_TFS, "kernel_us": t_kernel * 1e6,
            "flops_per_launch": flops, "traffic": read_profile_value("ragged_inc_hidden_conv_traffic.md",
                                                                    "traffic_bytes_per_launch"),
            "algorithmic_bytes_per_launch": 2 * 32 * 4 * rb.total}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_time(fn, threads, reps=1, warmup=False):
    """Median wall time of fn() on `threads` host threads."""
    import torch
    torch.set_num_threads(threads)
    if warmup:
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_baselines():
    """The FP64 oracle (a PyTorch-CPU restatement of the reference's op
    sequence) on a bounded sample of every workload, forward + adjoint
    (autograd through the oracle), on 1 host thread and on all of them.
    burgers_reference_B1_N256 is the reference's own algorithm exactly
    (Euler, WENO5 with LF splitting, dense-matrix diffusion) at its full
    size."""
    import torch
    from oracle import BurgersOracle, KSOracle, burgers_params, ks_params
    from oracle import simulate_rollout as o_rollout
    from inc_amd.correction import CorrectionCNN
    cores = len(os.sched_getaffinity(0))
    out = {"cpu_model": cpu_model(), "cores": cores}

    def burgers_reference():
        so = BurgersOracle(burgers_params(256, 0.005, nu=0.01, batch_size=1))
        u = sine_state(256, 1000)[None].clone().requires_grad_(True)
        cur, outs = u, []
        for _ in range(200):
            cur = so.step(cur)
            outs.append(cur)
        torch.stack(outs, 1).square().mean().backward()

    def burgers_reference_fwd():
        so = BurgersOracle(burgers_params(256, 0.005, nu=0.01, batch_size=1))
        with torch.no_grad():
            cur = sine_state(256, 1000)[None]
            for _ in range(200):
                cur = so.step(cur)

    def ks_case(b, n, Lk, h, cnn=None, ckpt=None):
        so = KSOracle(ks_params([Lk] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
        so.doublestep = False
        u = initial_state(b, n, Lk, 1001, kmax=16 if cnn else 10).requires_grad_(True)
        if cnn is None:
            res, _ = o_rollout(None, so, u, h, None, L=so.domain_size)
        else:
            res, _ = o_rollout(lambda v, L_: cnn(v), so, u, h, "INC", L=so.domain_size)
        res.square().mean().backward()

    def burgers_inc(b, n, h, cnn):
        so = BurgersOracle(burgers_params(n, 0.01, nu=0.01, batch_size=b, time_scheme="SSPRK3"))
        u = torch.stack([sine_state(n, 2000 + i) for i in range(b)]).requires_grad_(True)
        res, _ = o_rollout(cnn, so, u, h, "INC")
        res.square().mean().backward()

    cnn2 = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2), backend="torch").double()
    cnn3 = CorrectionCNN(6, 48, 7, generator=torch.Generator().manual_seed(3), backend="torch", scale=0.01).double()
    cnn4 = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2), backend="torch", scale=0.01).double()
    ns, nus, dts, hz = ragged_workload()
    sample = list(range(0, len(ns), len(ns) // 4))[:4]

    def uneven_inc():
        for i in sample:
            so = BurgersOracle(burgers_params(ns[i], dts[i], nu=nus[i], batch_size=1, time_scheme="SSPRK3"))
            u = sine_state(ns[i], 3000 + i)[None].clone().requires_grad_(True)
            res, _ = o_rollout(cnn4, so, u, hz[i], "INC")
            res.square().mean().backward()

    cases = {
        "burgers_reference_B1_N256": (burgers_reference, 200,
                                      "reference algorithm: Burgers Euler/WENO5/dense D2, B=1, N=256, 200 steps, fp64"),
        "burgers_reference_B1_N256_forward_only": (burgers_reference_fwd, 200, "the same, forward only"),
        "ks_etdrk4_B256_N256": (lambda: ks_case(256, 256, 64.0, 2), 256 * 2,
                                "KS ETDRK4, B=256, N=256, 2 of 1000 steps"),
        "burgers_inc_B512_N128": (lambda: burgers_inc(16, 128, 16, cnn2), 16 * 16,
                                  "Burgers INC + CNN(4x32,k5), 16 of 512 trajectories, N=128, 16 steps"),
        "ks_inc_B1024_N512": (lambda: ks_case(2, 512, 100.0, 8, cnn=cnn3), 2 * 8,
                              "KS INC + CNN(6x48,k7), 2 of 1024 trajectories, N=512, 8 of 64 steps"),
        "uneven_inc_2048": (uneven_inc, sum(hz[i] for i in sample),
                 "uneven Burgers INC + CNN(4x32,k5), 4 of 2048 trajectories with their own N, nu, dt, horizon"),
    }
    first = True
    for key, (fn, tsteps, what) in cases.items():
        row = {"sample": what, "trajectory_steps": tsteps,
               "unit": "trajectory-steps/s" + (" (fwd)" if key.endswith("only") else " (fwd+adj)")}
        for threads in (1, cores):
            row[f"value_{threads}_threads"] = tsteps / _cpu_time(fn, threads, warmup=first)
            first = False
        out[key] = row
    torch.set_num_threads(cores)
    return out


def extra_ks_saturation(b=8192, h=200):
    """The headline kernels at a batch large enough to fill the SIMDs (4096
    trajectory pairs): shows their throughput-bound rate next to the
    latency-bound headline workload."""
    import torch
    import inc_amd
    dev = torch.device("cuda")
    sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=N, dt=DT, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    Ld = torch.tensor([L], device=dev)
    u0 = initial_state(b, N, L, 1005).float().to(dev)
    cot = torch.randn(b, h, N, device=dev) * 1e-3
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]

    def run():
        u = u0.clone().requires_grad_(True)
        ev[0].record()
        traj = sv.rollout(u, h, Ld)
        ev[1].record()
        torch.autograd.backward(traj, cot)
        ev[2].record()
    run()
    torch.cuda.synchronize()
    tf = ta = float("inf")
    for _ in range(3):
        run()
        torch.cuda.synchronize()
        tf = min(tf, ev[0].elapsed_time(ev[1]) * 1e-3)
        ta = min(ta, ev[1].elapsed_time(ev[2]) * 1e-3)
    ts = b * h
    return {"workload": f"KS ETDRK4 fwd+adj B={b} N={N} {h} steps fp32", "trajectory_steps_per_s": ts / (tf + ta),
            "fwd_GBps_algorithmic": 2 * 4 * N * ts / tf / 1e9, "adj_GBps_algorithmic": 3 * 4 * N * ts / ta / 1e9,
            "fwd_ms": tf * 1e3, "adj_ms": ta * 1e3}


def extra_ks_mixed_radix(b=256, h=200, sizes=(250, 256, 384, 512)):
    """KS ETDRK4 fwd+adj at resolutions that are not powers of two (the
    mixed-radix transform through LDS) beside the power-of-two ones next to
    them (the register-transpose transform), L scaled with N."""
    import torch
    import inc_amd
    dev = torch.device("cuda")
    out = {"workload": f"KS ETDRK4 fwd+adj B={b}, {h} steps, fp32, L = 64 N / 256"}
    for n in sizes:
        Ln = 64.0 * n / 256
        sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([Ln], resolution=n, dt=DT, time_scheme="ETDRK4", device="cuda"))
        sv.doublestep = False
        Ld = torch.tensor([Ln], device=dev)
        u0 = initial_state(b, n, Ln, 1007).float().to(dev)
        cot = torch.randn(b, h, n, device=dev) * 1e-3

        def run():
            u = u0.clone().requires_grad_(True)
            torch.autograd.backward(sv.rollout(u, h, Ld), cot)
        run()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(3):
            a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            run()
            e.record()
            torch.cuda.synchronize()
            best = min(best, a.elapsed_time(e) * 1e-3)
        pm = [q for q in (3, 5) if n % q == 0 and 32 <= n // q <= 256 and (n // q) & (n // q - 1) == 0]
        kind = ("register" if n & (n - 1) == 0 else
                (f"radix-{pm[0]} step + register transforms" if pm else "mixed-radix through LDS"))
        out[f"N{n}"] = {"transform": kind, "ms": best * 1e3, "trajectory_steps_per_s": b * h / best}
    return out


def extra_ks_checkpoint(b=B, h=H, ks=(10, 50, 100)):
    """The headline workload (KS ETDRK4 fwd+adj, B=256, N=256, 1000 steps)
    with checkpoint-and-recompute every k steps in the fused rollout: time
    per fwd+adj and the peak memory the rollout and its adjoint allocate,
    against the adjoint over the stages of all steps."""
    import torch
    import inc_amd
    dev = torch.device("cuda")
    sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=N, dt=DT, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    Ld = torch.tensor([L], device=dev)
    u0 = initial_state(b, N, L, 1006).float().to(dev)
    cot = torch.randn(b, h, N, device=dev) * 1e-3
    out = {"workload": f"KS ETDRK4 fwd+adj B={b} N={N} {h} steps fp32, checkpoint every k steps"}
    for k in (None,) + tuple(ks):
        def run():
            u = u0.clone().requires_grad_(True)
            traj = sv.rollout(u, h, Ld, checkpoint_every=k)
            torch.autograd.backward(traj, cot)
        run()
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        t = _timed(run)
        peak = torch.cuda.max_memory_allocated() - base
        out["all_stages" if k is None else f"k{k}"] = {"ms": t * 1e3, "trajectory_steps_per_s": b * h / t,
                                                       "peak_MB": peak / 1e6}
    return out


def extra_ks_mse(b=B, h=H):
    """The headline rollout trained against a target trajectory: MSE loss and
    its gradient w.r.t. the initial state, with the loss fused into the
    kernels (simulate_rollout_mse) and with torch's MSELoss on the returned
    trajectory (simulate_rollout + F.mse_loss), same inputs."""
    import torch
    import inc_amd
    from torch.nn import functional as F
    dev = torch.device("cuda")
    sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=N, dt=DT, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    Ld = torch.tensor([L], device=dev)
    u0 = initial_state(b, N, L, 1006).float().to(dev)
    with torch.no_grad():
        target = sv.rollout(initial_state(b, N, L, 1007).float().to(dev), h, Ld)

    def fused():
        u = u0.clone().requires_grad_(True)
        loss, _ = inc_amd.simulate_rollout_mse(sv, u, target, L=Ld)
        loss.backward()

    def unfused():
        u = u0.clone().requires_grad_(True)
        traj, _ = inc_amd.simulate_rollout(None, sv, u, h, None, L=Ld)
        F.mse_loss(traj, target).backward()
    tf, tu = _timed(fused, reps=5), _timed(unfused, reps=5)
    return {"workload": f"KS ETDRK4 B={b} N={N} {h} steps fp32, MSE to a target, loss + gradient",
            "fused_ms": tf * 1e3, "unfused_ms": tu * 1e3, "fused_trajectory_steps_per_s": b * h / tf,
            "unfused_trajectory_steps_per_s": b * h / tu}


def latency_bound(solver, Ld, u0, t_headline, b=2):
    """The chain of dependent steps alone: the same rollout and adjoint for a
    single FFT pair (B = 2), whose time is set by step latency, not by how
    many trajectories share the chip.  ceiling_ms = that time, the bound the
    headline workload cannot beat with this algorithm however little HBM it
    moves; frac = ceiling / the headline kernels' time."""
    import torch
    u = u0[:b].clone()
    cot = torch.randn(b, H, N, device=u0.device) * 1e-3
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    best = float("inf"), float("inf")
    for _ in range(4):
        uu = u.clone().requires_grad_(True)
        ev[0].record()
        traj = solver.rollout(uu, H, Ld)
        ev[1].record()
        torch.autograd.backward(traj, cot)
        ev[2].record()
        torch.cuda.synchronize()
        best = (min(best[0], ev[0].elapsed_time(ev[1]) * 1e-3), min(best[1], ev[1].elapsed_time(ev[2]) * 1e-3))
    ceiling = best[0] + best[1]
    return {"what": "fwd+adj rollout of one FFT pair (B=2), same N, L, dt, steps", "ceiling_ms": ceiling * 1e3,
            "fwd_us_per_step": best[0] / H * 1e6, "adj_us_per_step": best[1] / H * 1e6,
            "us_per_fft": (best[0] + best[1]) / H / 16 * 1e6, "frac": ceiling / t_headline}


def plain_path(solver, u0, Ld, ntraj=8, h=50):
    """One launch per trajectory per step (no batching), fwd + adjoint."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in range(ntraj):
        u = u0[b:b + 1].clone().requires_grad_(True)
        cur, outs = u, []
        for _ in range(h):
            cur = solver.step(cur, Ld)
            outs.append(cur)
        torch.autograd.backward(torch.stack(outs, 1), torch.ones(1, h, N, device=u0.device))
    torch.cuda.synchronize()
    return ntraj * h / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-trajectory plain path")
    ap.add_argument("--no-extras", action="store_true", help="skip the INC and ragged-batch workloads")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device-to-device copy first (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: trajectories are independent, a multi-GPU run would shard the "
              "batch with no exchange on this path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import inc_amd
    dev = torch.device("cuda:0")
    solver = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=N, dt=DT, time_scheme="ETDRK4",
                                                       device="cuda"))
    solver.doublestep = False
    Ld = torch.tensor([L], device=dev)
    u0 = initial_state(B, N, L, 1001).float().to(dev)
    gen = torch.Generator().manual_seed(7)
    cot = (torch.randn(B, H, N, generator=gen) * 1e-3).to(dev)

    def one_step(ev=None):
        u = u0.clone().requires_grad_(True)
        if ev is not None:
            ev[0].record()
        traj, _ = inc_amd.simulate_rollout(None, solver, u, H, None, L=Ld)
        if ev is not None:
            ev[1].record()
        torch.autograd.backward(traj, cot)
        if ev is not None:
            ev[2].record()
        return u

    if args.calibrate_copy:
        src = torch.randn(64 * 2 ** 20, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        u = one_step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_fwd = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps * 1e-3
    t_adj = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps * 1e-3
    assert torch.isfinite(u.grad).all()

    ts = B * H
    value = ts * args.steps / wall
    fwd_bytes, adj_bytes = 2 * 4 * N * ts, 3 * 4 * N * ts
    dom_is_adj = t_adj >= t_fwd
    dom_t, dom_bytes = (t_adj, adj_bytes) if dom_is_adj else (t_fwd, fwd_bytes)
    achieved = dom_bytes / dom_t / 1e9
    out = {
        "metric": "fwd+adjoint trajectory-steps/s, batched vs plain, 1 MI355X; % HBM roofline",
        "value": value, "unit": "trajectory-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded low-mode cosine initial states)",
        "config": {"workload": "KS ETDRK4 plain rollout fwd+adj: B=256, N=256, L=64, dt=0.25, 1000 steps, "
                               "2/3 dealias, fp32", "batch": B, "n": N, "horizon": H},
        "roofline": {"bound": "hbm", "kernel": "ks_adj_kernel" if dom_is_adj else "ks_fwd_kernel",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(), "measured_copy_GBps": copy_bandwidth(dev)},
        "forward_trajectory_steps_per_s": ts / t_fwd,
        "kernel_ms": {"fwd": t_fwd * 1e3, "adj": t_adj * 1e3},
    }
    adj_flops = read_profile_value("ks_valu_counters.md", "adj_flops_per_launch")
    if adj_flops is not None and dom_is_adj:
        # FP32 flops of the dominant kernel (rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 x 64 lanes, per
        # launch of this workload) over its live-measured duration, against the vector peak
        out["compute"] = {"bound": "valu", "kernel": "ks_adj_kernel", "achieved": adj_flops / t_adj / 1e12,
                          "peak": VALU_PEAK_TFS, "unit": "TFLOP/s", "frac": adj_flops / t_adj / 1e12 / VALU_PEAK_TFS,
                          "flops_per_launch": adj_flops}
    out["latency_bound"] = latency_bound(solver, Ld, u0, t_fwd + t_adj)
    if not args.no_plain:
        plain = plain_path(solver, u0, Ld)
        out["plain_trajectory_steps_per_s"] = plain
        out["batched_over_plain"] = value / plain
    if not args.no_extras:
        out["batched"] = batched_ragged_inc()
        out["other_workloads"] = {"ks_saturated": extra_ks_saturation(), "ks_mse_fused": extra_ks_mse(),
                                  "ks_checkpoint": extra_ks_checkpoint(), "ks_mixed_radix": extra_ks_mixed_radix(),
                                  "burgers_saturated": extra_burgers_saturation(), "burgers_inc": extra_burgers_inc(),
                                  "burgers_inc_eval": extra_burgers_inc_eval(),
                                  "ks_inc": extra_ks_inc(), "resnet1d": extra_reference_net("resnet1d"),
                                  "unet1d": extra_reference_net("unet1d")}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baselines"] = cpu_baselines()
        if "batched" in out:
            c4 = out["cpu_baselines"]["uneven_inc_2048"]
            out["batched"]["cpu_baseline"] = {
                "value": c4[f"value_{out['cpu_baselines']['cores']}_threads"], "unit": c4["unit"],
                "cores": out["cpu_baselines"]["cores"], "kind": "port", "sample": c4["sample"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
