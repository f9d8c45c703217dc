"""Accuracy of the HIP solver step against the FP64 oracle, per workload and
rollout horizon (test infrastructure; needs a GPU).

    python tests/accuracy_report.py            # prints a markdown table

For every workload it reports, as relative L2 errors against the FP64 oracle
(or, where the oracle would take minutes on the CPU, against the float64 HIP
kernels, which tests/test_gpu_*.py pin to the oracle at 1e-11):
  * teacher-forced one-step error of the float32 kernels (the 1e-5 bar),
  * free-running error of the float32 rollout at several horizons (reported,
    not gated: KS is chaotic and decorrelates),
  * the float64 kernels' free-running error (rounding level),
  * for the INC workloads the float32 gradient error (the 1e-4 bar).
"""
import math
import os
import sys

import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import inc_amd  # noqa: E402
from inc_amd.correction import CorrectionCNN  # noqa: E402
from oracle import BurgersOracle, KSOracle, burgers_params, ks_params  # noqa: E402

DEV = torch.device("cuda")


def rel(a, b):
    a, b = a.double().cpu(), b.double().cpu()
    return float((a - b).norm() / b.norm())


def sine_states(b, n, seed):
    out = []
    for i in range(b):
        gen = torch.Generator().manual_seed(seed + i)
        x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
        u = torch.zeros(n, dtype=torch.float64)
        for k in range(1, 9):
            u += torch.randn(1, generator=gen, dtype=torch.float64) / k * torch.sin(
                k * x + 2 * math.pi * torch.rand(1, generator=gen, dtype=torch.float64))
        out.append(u / u.abs().max())
    return torch.stack(out)


def cos_modes(b, n, L, seed, kmax, sd=0.5):
    gen = torch.Generator().manual_seed(seed)
    x = torch.arange(n, dtype=torch.float64) * (L / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        u += torch.randn(b, 1, generator=gen, dtype=torch.float64) * sd * torch.cos(
            2 * math.pi * k * x / L + 2 * math.pi * torch.rand(b, 1, generator=gen, dtype=torch.float64))
    return u


def row(name, tf, horizons, f64, grad=None, wgrad=None, g64=None):
    hz = ", ".join(f"{h}: {e:.1e}" for h, e in horizons)
    g = "" if grad is None else f"{grad:.1e}"
    if wgrad is not None:
        g += f" (weights: {wgrad:.1e})"
    if g64 is not None:
        g += f"; fp64 {g64:.1e}"
    return f"| {name} | {tf:.1e} | {hz} | {f64:.1e} | {g} |"


def linear_weights(shape, seed):
    """Fixed weights w of the functional sum(w * trajectory): its cotangent is
    w in every precision, so the gradient error measures the adjoint itself
    (an MSE to a nearby target would measure the target's rounding)."""
    return torch.randn(shape, generator=torch.Generator().manual_seed(seed), dtype=torch.float64)


def burgers_plain():
    """Burgers, no correction, N=256, nu=0.01, dt=0.005, 200 steps (SSP-RK3); gradient of a fixed
    linear functional of the trajectory with respect to u0."""
    n, H = 256, 200
    u0 = sine_states(1, n, 1000)
    wts = linear_weights((1, H, n), 7)
    so = BurgersOracle(burgers_params(n, 0.005, nu=0.01, time_scheme="SSPRK3"))
    ug = u0.clone().requires_grad_(True)
    ref = [ug]
    for _ in range(H):
        ref.append(so.step(ref[-1]))
    (torch.stack(ref[1:], 1) * wts).sum().backward()
    gref = ug.grad
    ref = [r.detach() for r in ref]
    out, grads = {}, {}
    for dt_ in (torch.float32, torch.float64):
        sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.005, resolution=n, nu=0.01, time_scheme="SSPRK3",
                                                            device="cuda"))
        u = u0.to(DEV, dt_).requires_grad_(True)
        traj = sv.rollout(u, H)
        (traj * wts.to(DEV, dt_)).sum().backward()
        out[dt_] = traj.detach().cpu()
        grads[dt_] = u.grad.cpu()
    sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.005, resolution=n, nu=0.01, time_scheme="SSPRK3",
                                                        device="cuda"))
    tf = max(rel(sv.step(ref[k].float().to(DEV)), ref[k + 1]) for k in range(0, H, 10))
    hz = [(h, rel(out[torch.float32][:, h - 1], ref[h])) for h in (1, 20, 200)]
    return row("Burgers plain N=256, 200 steps (gradient: linear functional)", tf, hz,
               rel(out[torch.float64][:, -1], ref[H]), rel(grads[torch.float32], gref),
               g64=rel(grads[torch.float64], gref))


def ks_plain():
    """KS ETDRK4, B=256 (oracle on 16 of them), N=256, L=64, dt=0.25, 1000 steps."""
    n, b, L, H = 256, 16, 64.0, 1000
    u0 = cos_modes(b, n, L, 1001, 10)
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
    so.doublestep = False
    ref = [u0]
    for _ in range(H):
        ref.append(so.step(ref[-1], so.domain_size))
    out = {}
    for dt_ in (torch.float32, torch.float64):
        sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4",
                                                       device="cuda"))
        sv.doublestep = False
        out[dt_] = sv.rollout(u0.to(DEV, dt_), H, torch.tensor([L], device=DEV)).cpu()
    tf = max(rel(sv.step(ref[k].float().to(DEV), torch.tensor([L], device=DEV)), ref[k + 1])
             for k in range(0, H, 50))
    hz = [(h, rel(out[torch.float32][:, h - 1], ref[h])) for h in (1, 10, 100, 400, 1000)]
    return row("KS ETDRK4 N=256 L=64, 1000 steps", tf, hz, rel(out[torch.float64][:, -1], ref[H]))


def burgers_inc():
    """Burgers INC: N=128, nu=0.01, dt=0.01, 16 steps, CNN(4x32,k5) with output scale 0.1; oracle on
    8 trajectories.  float32 runs the network inside the rollout and adjoint kernels (the training
    path); its gradient column gives the initial state's and, in brackets, the weights' (worst
    parameter tensor)."""
    n, b, H = 128, 8, 16
    u0 = sine_states(b, n, 2000)
    cnn = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2), backend="torch").double()
    target = torch.randn(b, H, n, generator=torch.Generator().manual_seed(9), dtype=torch.float64) * 0.3

    def oracle_run():
        so = BurgersOracle(burgers_params(n, 0.01, nu=0.01, batch_size=b, time_scheme="SSPRK3"))
        u = u0.clone().requires_grad_(True)
        cur, outs = u, []
        for _ in range(H):
            so.correction_s = 0.1 * cnn(cur)
            cur = so.step(cur)
            outs.append(cur)
        traj = torch.stack(outs, 1)
        ((traj - target) ** 2).mean().backward()
        return traj.detach(), u.grad
    ref, gref = oracle_run()
    wref = [p.grad.clone() for p in cnn.parameters()]
    res = {}
    for dt_ in (torch.float32, torch.float64):
        net = CorrectionCNN(4, 32, 5, scale=0.1).to(DEV, dt_)
        net.load_state_dict(cnn.state_dict())
        sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b,
                                                            time_scheme="SSPRK3", device="cuda"))
        u = u0.to(DEV, dt_).requires_grad_(True)
        traj, _ = inc_amd.simulate_rollout(net, sv, u, H, "INC")
        ((traj - target.to(DEV, dt_)) ** 2).mean().backward()
        res[dt_] = (traj.detach().cpu(), u.grad.cpu(),
                    max(rel(p.grad.cpu(), w) for p, w in zip(net.parameters(), wref)))
    sv = inc_amd.BurgersSolverTorch(inc_amd.BGSimParams(dt=0.01, resolution=n, nu=0.01, batch_size=b,
                                                        time_scheme="SSPRK3", device="cuda"))
    net32 = CorrectionCNN(4, 32, 5).to(DEV)
    net32.load_state_dict(cnn.state_dict())
    tf = 0.0
    so = BurgersOracle(burgers_params(n, 0.01, nu=0.01, batch_size=b, time_scheme="SSPRK3"))
    prev = u0
    for k in range(H):
        with torch.no_grad():
            so.correction_s = 0.1 * cnn(prev)
            nxt = so.step(prev)
            sv.correction_s = so.correction_s.float().to(DEV)
            tf = max(tf, rel(sv.step(prev.float().to(DEV)), nxt))
        prev = ref[:, k]
    hz = [(h, rel(res[torch.float32][0][:, h - 1], ref[:, h - 1])) for h in (1, 8, 16)]
    return row("Burgers INC N=128, 16 steps (fwd+adj)", tf, hz, rel(res[torch.float64][0], ref),
               rel(res[torch.float32][1], gref), res[torch.float32][2])


def ks_inc():
    """KS INC: L=100, N=512, dt=0.25, 64 steps, checkpoint every 8, CNN(6x48,k7); oracle on 4 trajectories."""
    n, b, L, H = 512, 4, 100.0, 64
    u0 = cos_modes(b, n, L, 1003, 16)
    cnn = CorrectionCNN(6, 48, 7, generator=torch.Generator().manual_seed(3), backend="torch").double()

    def oracle_run():
        so = KSOracle(ks_params([L] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
        so.doublestep = False
        u = u0.clone().requires_grad_(True)
        cur, outs = u, []
        for _ in range(H):
            so.correction_s = 0.01 * cnn(cur)
            cur = so.step(cur, so.domain_size)
            outs.append(cur)
        traj = torch.stack(outs, 1)
        (traj ** 2).mean().backward()
        return traj.detach(), u.grad
    ref, gref = oracle_run()
    res = {}
    for dt_ in (torch.float32, torch.float64):
        net = CorrectionCNN(6, 48, 7).to(DEV, dt_)
        net.load_state_dict(cnn.state_dict())
        sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4",
                                                       device="cuda"))
        sv.doublestep = False
        u = u0.to(DEV, dt_).requires_grad_(True)
        traj, _ = inc_amd.simulate_rollout(lambda v, L_: 0.01 * net(v), sv, u, H, "INC",
                                           L=torch.tensor([L], device=DEV), checkpoint_every=8)
        (traj ** 2).mean().backward()
        res[dt_] = (traj.detach().cpu(), u.grad.cpu())
    hz = [(h, rel(res[torch.float32][0][:, h - 1], ref[:, h - 1])) for h in (1, 16, 64)]
    # teacher forcing: one fp32 step from the oracle state with the oracle's correction
    sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4",
                                                   device="cuda"))
    sv.doublestep = False
    so = KSOracle(ks_params([L] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
    so.doublestep = False
    tf, prev = 0.0, u0
    for k in range(H):
        with torch.no_grad():
            s = 0.01 * cnn(prev)
            so.correction_s = s
            nxt = so.step(prev, so.domain_size)
            sv.correction_s = s.float().to(DEV)
            tf = max(tf, rel(sv.step(prev.float().to(DEV), torch.tensor([L], device=DEV)), nxt))
        prev = ref[:, k]
    return row("KS INC N=512 L=100, 64 steps, ckpt 8 (fwd+adj)", tf, hz, rel(res[torch.float64][0], ref),
               rel(res[torch.float32][1], gref))


def ks_plain_gradient():
    """KS ETDRK4 gradient of a fixed linear functional of the trajectory with respect to u0, at
    horizons 10 and 100 (the chaotic dynamics amplify any perturbation over longer horizons)."""
    n, b, L = 256, 8, 64.0
    u0 = cos_modes(b, n, L, 1001, 10)
    errs = []
    for H in (10, 100):
        wts = linear_weights((b, H, n), 8)
        so = KSOracle(ks_params([L] * b, resolution=n, dt=0.25, time_scheme="ETDRK4"))
        so.doublestep = False
        ug = u0.clone().requires_grad_(True)
        cur, outs = ug, []
        for _ in range(H):
            cur = so.step(cur, so.domain_size)
            outs.append(cur)
        (torch.stack(outs, 1) * wts).sum().backward()
        row_e = []
        for dt_ in (torch.float32, torch.float64):
            sv = inc_amd.KSSolverTorch(inc_amd.KSSimParams([L], resolution=n, dt=0.25, time_scheme="ETDRK4",
                                                           device="cuda"))
            sv.doublestep = False
            u = u0.to(DEV, dt_).requires_grad_(True)
            traj = sv.rollout(u, H, torch.tensor([L], dtype=torch.float64, device=DEV))
            (traj * wts.to(DEV, dt_)).sum().backward()
            row_e.append(rel(u.grad, ug.grad))
        errs.append((H, row_e[0], row_e[1]))
    txt = "; ".join(f"horizon {H}: fp32 {e32:.1e}, fp64 {e64:.1e}" for H, e32, e64 in errs)
    return f"| KS ETDRK4 N=256 L=64 gradient of u0 (linear functional) | | | | {txt} |"


def ragged_inc():
    """The uneven INC batch: 12 trajectories drawn from it (N from 64 to 1024, their own nu, dt and
    horizon), 64 steps, CNN(4x32,k5) x 0.01, the one-call path (inc_ragged_inc_fwd / _adj); oracle per
    trajectory.  Gradient of a fixed linear functional of the trajectory: u0 and, in brackets, the
    weights (worst parameter tensor)."""
    sys.path.insert(0, ROOT)
    import bench
    ns, nus, dts, hz = bench.ragged_workload()
    pick = list(range(0, len(ns), len(ns) // 12))[:12]
    ns, nus, dts, hz = ([v[i] for i in pick] for v in (ns, nus, dts, hz))
    steps = 64
    cnn = CorrectionCNN(4, 32, 5, generator=torch.Generator().manual_seed(2), backend="torch", scale=0.01).double()
    u0s = [sine_states(1, n, 3000 + i)[0] for i, n in enumerate(ns)]
    wts = [linear_weights((steps, n), 100 + i) for i, n in enumerate(ns)]
    refs, grefs, tf = [], [], 0.0
    for i, n in enumerate(ns):
        so = BurgersOracle(burgers_params(n, dts[i], nu=nus[i], batch_size=1, time_scheme="SSPRK3"))
        ug = u0s[i][None].clone().requires_grad_(True)
        cur, outs = ug, []
        for k in range(steps):
            if k < hz[i]:
                so.correction_s = cnn(cur)
                cur = so.step(cur)
            outs.append(cur)
        traj = torch.cat(outs, 0)
        (traj * wts[i]).sum().backward()
        refs.append(traj.detach())
        grefs.append(ug.grad[0])
    wref = [p.grad.clone() for p in cnn.parameters()]
    res = {}
    for dt_ in (torch.float32, torch.float64):
        rb = inc_amd.RaggedBurgers(ns, nus, dts, horizon=hz, dtype=dt_)
        net = CorrectionCNN(4, 32, 5, scale=0.01).to(DEV, dt_)
        net.load_state_dict(cnn.state_dict())
        u = rb.pack(u0s).requires_grad_(True)
        traj = rb.rollout(u, steps, model=net)
        w_packed = torch.stack([rb.pack([w[k] for w in wts]) for k in range(steps)])
        (traj * w_packed.to(dt_)).sum().backward()
        per = rb.unpack(traj.detach().cpu())
        gu = rb.unpack(u.grad.cpu())
        res[dt_] = (per, gu, max(rel(p.grad.cpu(), w) for p, w in zip(net.parameters(), wref)))
        if dt_ == torch.float32:
            # teacher forcing: one float32 ragged step from the oracle states with the oracle's correction
            for k in (0, 15, 31):
                prev = [u0s[i] if k == 0 else refs[i][k - 1] for i in range(len(ns))]
                with torch.no_grad():
                    s = [cnn(p[None])[0] for p in prev]
                    nxt = rb.advance(rb.pack(prev), rb.pack(s), steps=1, step0=k)
                for i, v in enumerate(rb.unpack(nxt.cpu())):
                    if k < hz[i]:
                        tf = max(tf, rel(v, refs[i][k]))
    def stacked(per):
        return torch.cat([p.reshape(-1) for p in per])
    hzs = []
    for h in (1, 16, 64):
        a = torch.cat([res[torch.float32][0][i][h - 1] for i in range(len(ns))])
        b = torch.cat([refs[i][h - 1] for i in range(len(ns))])
        hzs.append((h, rel(a, b)))
    f64 = rel(stacked(res[torch.float64][0]), stacked(refs))
    g32 = rel(stacked(res[torch.float32][1]), stacked(grefs))
    g64 = rel(stacked(res[torch.float64][1]), stacked(grefs))
    return row("uneven INC batch: 12 of its trajectories, 64 steps (fwd+adj, one-call path)", tf, hzs, f64, g32,
               res[torch.float32][2], g64=g64)


def main():
    print("| workload | fp32 one-step (teacher-forced) | fp32 free-running, by horizon | fp64 kernels "
          "at the last horizon | fp32 gradient |")
    print("|---|---|---|---|---|")
    for fn in (burgers_plain, ks_plain, ks_plain_gradient, burgers_inc, ks_inc, ragged_inc):
        print(fn(), flush=True)


if __name__ == "__main__":
    main()
