"""Write golden input/output vectors from the reference 1D solvers.

Runs ONLY in the development container, where the reference checkout lives
at /root/reference; the vectors it writes (``*.npz`` next to this file) are
data and travel with the tests, the reference does not.  It imports
``solvers/solver_1d.py`` (which needs only torch and math) with float64 as
the default dtype, because the reference creates its matrices and wavenumber
grids in the default dtype (solver_1d.py:76,89,97,208).

What the vectors pin that closed forms cannot:
  * weno_reconstruction 'right' (solver_1d.py:24-33): the shifted stencil and
    the mirrored beta_2 term.
  * calc_nonlinear_weno (solver_1d.py:129-142) and its autograd gradient,
    including the torch.max(dim) subgradient of the Lax-Friedrichs alpha.
  * The Euler step composed from the reference's pieces (its ``step`` cannot
    be called as shipped: solver_1d.py:116 vs :188), FDM convection, the
    'learning' forcing (solver_1d.py:74-85).
  * KS ETD1 / ETRK2 with the double step, correction_s / correction_u,
    per-trajectory domain sizes, an L = 2*pi case that hits the Lop == 0
    branch at |k| = 1, and autograd gradients through 3-step rollouts.

Usage:  python tests/golden/make_golden.py   (rewrites the .npz files)
"""
import importlib.util
import math
import os
import sys
from types import SimpleNamespace

import numpy as np
import torch

REF = "/root/reference/solvers/solver_1d.py"
HERE = os.path.dirname(os.path.abspath(__file__))


def load_reference():
    spec = importlib.util.spec_from_file_location("_ref_solver_1d", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def smooth_field(gen, b, n, kmax=8):
    x = torch.arange(n, dtype=torch.float64) * (2 * math.pi / n)
    u = torch.zeros(b, n, dtype=torch.float64)
    for k in range(1, kmax + 1):
        a = torch.randn(b, 1, generator=gen, dtype=torch.float64) / k
        ph = torch.rand(b, 1, generator=gen, dtype=torch.float64) * 2 * math.pi
        u = u + a * torch.sin(k * x + ph)
    return u / u.abs().amax(-1, keepdim=True)


def weno_vectors(ref, gen):
    out = {}
    b, n = 3, 32
    f_rand = torch.randn(b, n, generator=gen, dtype=torch.float64)
    f_smooth = smooth_field(gen, b, n)
    f_step = torch.where(torch.arange(n) < n // 2, 1.0, -0.5).to(torch.float64).repeat(b, 1)
    f_step = f_step + 0.01 * torch.randn(b, n, generator=gen, dtype=torch.float64)
    for name, f in (("rand", f_rand), ("smooth", f_smooth), ("step", f_step)):
        out[f"f_{name}"] = f.numpy()
        for d in ("left", "right"):
            out[f"{d}_{name}"] = ref.weno_reconstruction(f, d).numpy()
    return out


def burgers_ns(n, b, L, nu, conv, eq="homogeneous", J=0, forcing=None):
    ns = SimpleNamespace(resolution=n, dt=0.01, x=torch.arange(n, dtype=torch.float64) * (L / n),
                         adaptive_CFL=None, L=L, nu=nu, device="cpu", convection_term=conv,
                         equation_type=eq, beta=0.0, batch_size=b, J=J)
    if forcing is not None:
        ns.A, ns.omega, ns.phi, ns.l = forcing
    return ns


def burgers_vectors(ref, gen):
    out = {}
    b, n, L, dt = 3, 32, 2 * math.pi, 0.01
    nu = torch.tensor([[0.01], [0.02], [0.05]], dtype=torch.float64)
    u = smooth_field(gen, b, n) * torch.tensor([[1.0], [0.7], [1.3]], dtype=torch.float64)
    u = u + 0.05 * torch.randn(b, n, generator=gen, dtype=torch.float64)
    s = 0.3 * torch.randn(b, n, generator=gen, dtype=torch.float64)
    out.update(u=u.numpy(), s=s.numpy(), nu=nu.numpy(), L=L, dt=dt)
    sw = ref.BurgersSolverTorch(burgers_ns(n, b, L, nu, "WENO"))
    sf = ref.BurgersSolverTorch(burgers_ns(n, b, L, nu, "FDM"))
    out["nonlinear_weno"] = sw.calc_nonlinear_weno(u).numpy()
    out["nonlinear_fdm"] = sf.calc_nonlinear_FDM(u).numpy()
    out["linear"] = sw.calc_linear(u).numpy()
    # Euler step composed from the reference pieces (evident intent of
    # solver_1d.py:187-192 with the correction added to the source).
    out["euler_weno_s"] = (u + dt * (sw.calc_nonlinear(u) + sw.calc_linear(u) + s)).numpy()
    out["euler_fdm_s"] = (u + dt * (sf.calc_nonlinear(u) + sf.calc_linear(u) + s)).numpy()
    # 'learning' forcing, J = 2, per-trajectory t
    J = 2
    A = torch.randn(b, 1, J, generator=gen, dtype=torch.float64) * 0.5
    om = torch.randn(b, 1, J, generator=gen, dtype=torch.float64)
    ph = torch.rand(b, 1, J, generator=gen, dtype=torch.float64) * 2 * math.pi
    ll = torch.randint(1, 4, (b, 1, J), generator=gen).to(torch.float64)
    sl = ref.BurgersSolverTorch(burgers_ns(n, b, L, nu, "WENO", "learning", J, (A, om, ph, ll)))
    t = torch.tensor([[0.0], [0.3], [1.7]], dtype=torch.float64)
    sl.t = t
    out.update(A=A.numpy(), omega=om.numpy(), phi=ph.numpy(), l=ll.numpy(), t=t.numpy())
    out["forcing"] = sl._compute_source(u).numpy()
    out["euler_learning_s"] = (u + dt * (sl.calc_nonlinear(u) + sl.calc_linear(u)
                                        + sl._compute_source(u) + s)).numpy()
    # gradient of an MSE loss through 3 hand-composed Euler steps
    H = 3
    u0 = u.clone().requires_grad_(True)
    ss = [(0.2 * torch.randn(b, n, generator=gen, dtype=torch.float64)).requires_grad_(True)
          for _ in range(H)]
    tgt = torch.randn(b, H, n, generator=gen, dtype=torch.float64) * 0.5
    cur, traj = u0, []
    for k in range(H):
        cur = cur + dt * (sw.calc_nonlinear(cur) + sw.calc_linear(cur) + ss[k])
        traj.append(cur)
    loss = ((torch.stack(traj, 1) - tgt) ** 2).mean()
    loss.backward()
    out.update(grad_u0=u0.grad.numpy(), grad_s=torch.stack([x.grad for x in ss], 1).numpy(),
               grad_ss=torch.stack([x.detach() for x in ss], 1).numpy(), grad_target=tgt.numpy(),
               grad_traj=torch.stack(traj, 1).detach().numpy(), grad_loss=loss.item())
    return out


def ks_vectors(ref, gen):
    out = {}
    for n in (32, 64):
        for scheme in ("ETD1", "ETRK2"):
            for ds_name, ds in (("mixed", [22.0, 64.0, 100.0]), ("twopi", [2 * math.pi, 12.0, 2 * math.pi])):
                b = len(ds)
                key = f"n{n}_{scheme}_{ds_name}"
                dom = torch.tensor(ds, dtype=torch.float64)
                p = SimpleNamespace(resolution=n, dt=0.1, device="cpu", domain_size=dom,
                                    alpha=torch.zeros(b, dtype=torch.float64), time_scheme=scheme)
                solver = ref.KSSolverTorch(p)
                u = smooth_field(gen, b, n, kmax=5)
                s = 0.2 * torch.randn(b, n, generator=gen, dtype=torch.float64)
                cu = 0.05 * torch.randn(b, n, generator=gen, dtype=torch.float64)
                out[f"{key}_u"] = u.numpy()
                out[f"{key}_s"] = s.numpy()
                out[f"{key}_cu"] = cu.numpy()
                out[f"{key}_L"] = dom.numpy()
                solver.doublestep = False
                out[f"{key}_single"] = solver.step(u, dom).numpy()
                solver.doublestep = True
                out[f"{key}_double"] = solver.step(u, dom).numpy()
                solver.correction_s = s
                out[f"{key}_double_s"] = solver.step(u, dom).numpy()
                solver.correction_u = cu
                out[f"{key}_double_s_cu"] = solver.step(u, dom).numpy()
                solver.correction_u = None
                # gradient through a 3-call rollout with per-step corrections
                H = 3
                u0 = u.clone().requires_grad_(True)
                ss = [(0.1 * torch.randn(b, n, generator=gen, dtype=torch.float64)).requires_grad_(True)
                      for _ in range(H)]
                tgt = torch.randn(b, H, n, generator=gen, dtype=torch.float64) * 0.5
                cur, traj = u0, []
                for k in range(H):
                    solver.correction_s = ss[k]
                    cur = solver.step(cur, dom)
                    traj.append(cur)
                loss = ((torch.stack(traj, 1) - tgt) ** 2).mean()
                loss.backward()
                out[f"{key}_gss"] = torch.stack([x.detach() for x in ss], 1).numpy()
                out[f"{key}_gtarget"] = tgt.numpy()
                out[f"{key}_gu0"] = u0.grad.numpy()
                out[f"{key}_gs"] = torch.stack([x.grad for x in ss], 1).numpy()
                out[f"{key}_gtraj"] = torch.stack(traj, 1).detach().numpy()
    return out


def main():
    if not os.path.exists(REF):
        print("reference checkout not found; golden vectors unchanged", file=sys.stderr)
        return 1
    torch.set_default_dtype(torch.float64)
    ref = load_reference()
    gen = torch.Generator().manual_seed(20251203)
    np.savez_compressed(os.path.join(HERE, "weno.npz"), **weno_vectors(ref, gen))
    np.savez_compressed(os.path.join(HERE, "burgers.npz"), **burgers_vectors(ref, gen))
    np.savez_compressed(os.path.join(HERE, "ks.npz"), **ks_vectors(ref, gen))
    print("golden vectors written to", HERE)
    return 0


if __name__ == "__main__":
    sys.exit(main())
