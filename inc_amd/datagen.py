"""Dataset generation on the GPU with the fused rollout kernels.

The reference trains from ``.pth`` dictionaries it downloads
(README.md:150) and ships no generator; its loaders define the schema
(data/data_loader_1d.py:8-118):

* KS:      {"trajectories": [B, T+1, D], "domain_size": [B, 1],
            "metadata": {"gen_dt", "resolution", ...}}
* Burgers: {"trajectories": [B, T+1, D], "nu": [B, 1],
            "A", "omega", "phi", "l": [B, 1, J],
            "metadata": {"gen_dt", "resolution", "L", ...}}

(``init_BGSim`` reads metadata['resolution'] and metadata['L'],
trainer/trainer_1d.py:346-349.)  Each generator advances the whole batch in
one launch per chunk of steps and returns CPU tensors in that schema, ready
for ``torch.save``.
"""
import math

import torch

from .config_1d import BGSimParams, KSSimParams
from .solver_1d import BurgersSolverTorch, KSSolverTorch


def _chunks(total, chunk):
    done = 0
    while done < total:
        k = min(chunk, total - done)
        yield k
        done += k


@torch.no_grad()
def generate_ks(count, resolution=64, domain_sizes=(22.0,), dt=0.25, steps=1000, warmup=0, scheme="ETDRK4",
                dealias=True, doublestep=False, modes=10, amplitude=0.5, seed=0, device="cuda",
                dtype=torch.float32, chunk=256):
    """KS trajectories from u0 = sum_{k<=modes} b_k cos(2 pi k x / L + psi_k),
    b_k ~ N(0, amplitude^2), with L drawn uniformly from ``domain_sizes`` per
    trajectory; ``warmup`` steps are run and discarded first."""
    gen = torch.Generator().manual_seed(seed)
    ds = torch.as_tensor(domain_sizes, dtype=torch.float64)
    Ls = ds[torch.randint(0, ds.numel(), (count,), generator=gen)]
    x = torch.arange(resolution, dtype=torch.float64)[None] / resolution  # x / L
    u = torch.zeros(count, resolution, dtype=torch.float64)
    for k in range(1, modes + 1):
        b = torch.randn(count, 1, generator=gen, dtype=torch.float64) * amplitude
        psi = torch.rand(count, 1, generator=gen, dtype=torch.float64) * 2 * math.pi
        u += b * torch.cos(2 * math.pi * k * x + psi)
    p = KSSimParams(Ls.to(device), resolution=resolution, dt=dt, time_scheme=scheme, device=device, dealias=dealias)
    solver = KSSolverTorch(p)
    solver.doublestep = doublestep
    Ld = Ls.to(device)
    cur = u.to(device, dtype)
    for k in _chunks(warmup, chunk):
        cur = solver.rollout(cur, k, Ld)[:, -1].contiguous()
    out = [cur.cpu()]
    for k in _chunks(steps, chunk):
        traj = solver.rollout(cur, k, Ld)
        out.append(traj.cpu())
        cur = traj[:, -1].contiguous()
    traj = torch.cat([out[0][:, None]] + out[1:], 1)
    return {"trajectories": traj, "domain_size": Ls[:, None].to(dtype),
            "metadata": {"gen_dt": float(dt * (2 if doublestep else 1)), "resolution": int(resolution),
                         "time_scheme": scheme, "dealias": bool(dealias), "warmup_steps": int(warmup),
                         "seed": int(seed)}}


@torch.no_grad()
def generate_burgers(count, resolution=128, L=2 * math.pi, dt=0.01, steps=200, nu_range=(0.003, 0.03), J=5,
                     amp=0.5, omega=0.4, max_l=3, scheme="SSPRK3", seed=0, device="cuda", dtype=torch.float32,
                     chunk=256):
    """Forced Burgers trajectories ('learning' equation type): per trajectory
    nu ~ LogUniform(nu_range), J forcing modes A ~ U(-amp, amp),
    omega ~ U(-omega, omega), phi ~ U(0, 2 pi), l ~ U{1..max_l}; the initial
    state is the forcing field at t = 0 (BurgersSolverTorch.initial_conditions,
    solver_1d.py:68-72)."""
    gen = torch.Generator().manual_seed(seed)
    lo, hi = math.log(nu_range[0]), math.log(nu_range[1])
    nu = torch.exp(torch.empty(count, 1, dtype=torch.float64).uniform_(lo, hi, generator=gen))
    A = (torch.rand(count, 1, J, generator=gen, dtype=torch.float64) * 2 - 1) * amp
    om = (torch.rand(count, 1, J, generator=gen, dtype=torch.float64) * 2 - 1) * omega
    ph = torch.rand(count, 1, J, generator=gen, dtype=torch.float64) * 2 * math.pi
    ll = torch.randint(1, max_l + 1, (count, 1, J), generator=gen).to(torch.float64)
    p = BGSimParams(dt=dt, resolution=resolution, L=L, nu=nu.to(device, dtype), equation_type="learning",
                    batch_size=count, time_scheme=scheme, device=device)
    p.A, p.omega, p.phi, p.l = (t.to(device, dtype) for t in (A, om, ph, ll))
    solver = BurgersSolverTorch(p)
    x = torch.arange(resolution, device=device, dtype=dtype) * (L / resolution)
    cur = solver.initial_conditions(x, 0.0).contiguous()
    out = [cur.cpu()]
    for k in _chunks(steps, chunk):
        traj = solver.rollout(cur, k)
        out.append(traj.cpu())
        cur = traj[:, -1].contiguous()
    traj = torch.cat([out[0][:, None]] + out[1:], 1)
    return {"trajectories": traj, "nu": nu.to(dtype), "A": A.to(dtype), "omega": om.to(dtype), "phi": ph.to(dtype),
            "l": ll.to(dtype),
            "metadata": {"gen_dt": float(dt), "resolution": int(resolution), "L": float(L), "time_scheme": scheme,
                         "nu_range": [float(nu_range[0]), float(nu_range[1])], "J": int(J), "seed": int(seed)}}
