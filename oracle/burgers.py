"""FP64 restatement of the reference Burgers solver (test infrastructure only).

Follows ``solvers/solver_1d.py:4-195`` of the reference.  Every array is
float64 on the CPU.  The WENO5 reconstruction is written table-driven: each
of the three candidate stencils is a triple of coefficient vectors
(candidate value q, the two linear forms a, b whose squares make the
smoothness indicator beta = 13/12 a^2 + 1/4 b^2) over stencil offsets
relative to the cell index i.  The same table drives the HIP kernels, so a
wrong entry here would show up as a golden-vector failure.
"""
import math
from types import SimpleNamespace

import torch

DT = torch.float64

#: epsilon of the nonlinear weights, solver_1d.py:5
WENO_EPS = 1e-6
#: linear weights, identical for both directions, solver_1d.py:22,33
WENO_GAMMA = (0.1, 0.6, 0.3)

# Offsets covered by each direction's 5-point stencil.
#   left  (solver_1d.py:14-22): f_{i-2} .. f_{i+2}, value at x_{i+1/2}
#   right (solver_1d.py:24-33): f_{i-1} .. f_{i+3}, value at x_{i+3/2}
#   (the reference's "right" is the left-biased formula shifted by one cell,
#    with beta_2's second linear form mirrored, solver_1d.py:32)
WENO_TABLE = {
    "left": {
        "offsets": (-2, -1, 0, 1, 2),
        # candidate k: (q coeffs, a coeffs, b coeffs) over offsets -2..2
        "cand": (
            ((1 / 3, -7 / 6, 11 / 6, 0, 0), (1, -2, 1, 0, 0), (1, -4, 3, 0, 0)),
            ((0, -1 / 6, 5 / 6, 1 / 3, 0), (0, 1, -2, 1, 0), (0, 1, 0, -1, 0)),
            ((0, 0, 1 / 3, 5 / 6, -1 / 6), (0, 0, 1, -2, 1), (0, 0, 3, -4, 1)),
        ),
    },
    "right": {
        "offsets": (-1, 0, 1, 2, 3),
        # candidate k over offsets -1..3
        "cand": (
            ((1 / 3, -7 / 6, 11 / 6, 0, 0), (1, -2, 1, 0, 0), (1, -4, 3, 0, 0)),
            ((0, -1 / 6, 5 / 6, 1 / 3, 0), (0, 1, -2, 1, 0), (0, -1, 0, 1, 0)),
            ((0, 0, 1 / 3, 5 / 6, -1 / 6), (0, 0, 1, -2, 1), (0, 0, 1, -4, 3)),
        ),
    },
}


def _shift(f, off):
    """f[..., i+off] with periodic wrap (torch.roll by -off), solver_1d.py:7-12."""
    return torch.roll(f, shifts=-off, dims=-1)


def weno_reconstruction(f, direction="left"):
    """WENO5 reconstruction with the reference's stencils, solver_1d.py:4-42.

    Returns F with F[..., i] the reconstructed value at x_{i+1/2} ('left')
    or x_{i+3/2} ('right').  Raises ValueError for other directions
    (solver_1d.py:35-36).
    """
    if direction not in WENO_TABLE:
        raise ValueError("Direction must be 'left' or 'right'.")
    tab = WENO_TABLE[direction]
    taps = [_shift(f, o) for o in tab["offsets"]]

    def lin(coef):
        acc = None
        for c, t in zip(coef, taps):
            if c == 0:
                continue
            term = c * t
            acc = term if acc is None else acc + term
        return acc

    qs, alphas = [], []
    for gamma, (cq, ca, cb) in zip(WENO_GAMMA, tab["cand"]):
        a = lin(ca)
        b = lin(cb)
        beta = (13 / 12) * a ** 2 + 0.25 * b ** 2
        qs.append(lin(cq))
        alphas.append(gamma / (WENO_EPS + beta) ** 2)
    s = alphas[0] + alphas[1] + alphas[2]
    return sum((al / s) * q for al, q in zip(alphas, qs))


def burgers_params(resolution, dt, L=2 * math.pi, nu=0.01, batch_size=1,
                   convection_term="WENO", equation_type="homogeneous",
                   adaptive_CFL=None, time_scheme="Euler", beta=0.0, J=0,
                   A=None, omega=None, phi=None, l=None):
    """Constructor-parameter object with the attributes BurgersSolverTorch
    reads (solver_1d.py:46-66,76-84,106-111); x_j = j*L/N."""
    x = torch.arange(resolution, dtype=DT) * (L / resolution)
    return SimpleNamespace(resolution=resolution, dt=dt, x=x, L=L, nu=nu,
                           device="cpu", batch_size=batch_size,
                           convection_term=convection_term,
                           equation_type=equation_type,
                           adaptive_CFL=adaptive_CFL, time_scheme=time_scheme,
                           beta=beta, J=J, A=A, omega=omega, phi=phi, l=l)


class BurgersOracle:
    """Restates BurgersSolverTorch (solver_1d.py:44-195) in float64.

    Differences from the reference, all deliberate:
      * ``_apply_corrections`` is called with the arity the reference's
        caller expects but its definition lacks (solver_1d.py:116 vs
        :166,:188); the evident intent rhs = N(u)+nu*D2u+source+s is used.
      * ``time_scheme='SSPRK3'`` adds the Shu-Osher three-stage scheme with
        the correction s frozen over the stages and the forcing evaluated at
        the stage times t, t+dt, t+dt/2.
    """

    def __init__(self, p, LOG=None):
        self.param = p
        self.resolution = p.resolution
        self.dt = p.dt
        self.x = p.x.to(DT)
        self.t = 0.0
        self.adaptive_CFL = p.adaptive_CFL
        self.correction_s = None
        self.L = p.L
        self.nu = p.nu
        self.dx = p.L / p.resolution
        self.LOG = LOG
        self.time_scheme = getattr(p, "time_scheme", "Euler")
        if self.time_scheme not in ("Euler", "SSPRK3"):
            raise ValueError("time_scheme must be 'Euler' or 'SSPRK3'")
        n = self.resolution
        # solver_1d.py:95-102: dense periodic 3-point Laplacian
        d2 = torch.zeros(n, n, dtype=DT)
        for i in range(n):
            d2[i, (i - 1) % n] += 1 / self.dx ** 2
            d2[i, i] += -2 / self.dx ** 2
            d2[i, (i + 1) % n] += 1 / self.dx ** 2
        self.D2_T = d2.T
        if p.convection_term == "FDM":
            # solver_1d.py:87-93: central first difference
            d1 = torch.zeros(n, n, dtype=DT)
            for i in range(n):
                d1[i, (i - 1) % n] += -1 / (2 * self.dx)
                d1[i, (i + 1) % n] += 1 / (2 * self.dx)
            self.D1_T = d1.T
            self.calc_nonlinear = self.calc_nonlinear_FDM
        elif p.convection_term == "WENO":
            self.calc_nonlinear = self.calc_nonlinear_weno
        else:
            raise ValueError("Convection scheme not recognized")

    # -- forcing, solver_1d.py:74-85 and :104-114 ---------------------------
    def _compute_delta(self, x, t):
        p = self.param
        xr = x.reshape(1, -1).to(DT)
        out = torch.zeros(p.batch_size, x.shape[0], dtype=DT)
        if isinstance(t, (int, float)):
            t = torch.full((p.batch_size, 1), float(t), dtype=DT)
        if tuple(t.shape) != (p.batch_size, 1):
            raise ValueError("t tensor must have shape [batch_size,1]")
        for j in range(p.J):
            arg = (p.omega[..., j] * t + (2 * math.pi * p.l[..., j] / p.L) * xr
                   + p.phi[..., j])
            out = out + p.A[..., j] * torch.sin(arg)
        return out

    def initial_conditions(self, x=None, t=None):
        return self._compute_delta(self.x if x is None else x,
                                   self.t if t is None else t)

    def _compute_source(self, u, t):
        kind = self.param.equation_type
        if kind in ("homogeneous", "viscosity"):
            return 0.0
        if kind == "linear":
            return self.param.beta * u
        if kind == "learning":
            return self._compute_delta(self.x, t)
        raise ValueError("Unknown equation type.")

    # -- spatial operators, solver_1d.py:125-146 -----------------------------
    def calc_nonlinear_FDM(self, u):
        return -u * (u @ self.D1_T)

    def calc_nonlinear_weno(self, u):
        flux = 0.5 * u ** 2
        alpha = torch.max(torch.abs(u), dim=1, keepdim=True)[0]
        fp = 0.5 * (flux + alpha * u)
        fm = 0.5 * (flux - alpha * u)
        hp = weno_reconstruction(fp, "left")
        hm = weno_reconstruction(fm, "right")
        dp = (hp - torch.roll(hp, 1, -1)) / self.dx
        dm = (hm - torch.roll(hm, 1, -1)) / self.dx
        return -(dp + dm)

    def calc_linear(self, u):
        return self.nu * (u @ self.D2_T)

    def rhs(self, u, t):
        r = self.calc_nonlinear(u) + self.calc_linear(u) + self._compute_source(u, t)
        if self.correction_s is not None:
            r = r + self.correction_s
        return r

    # -- time stepping, solver_1d.py:148-195 ---------------------------------
    def _advance(self, u, dt):
        t = self.t
        if self.time_scheme == "Euler":
            return u + dt * self.rhs(u, t)
        u1 = u + dt * self.rhs(u, t)
        u2 = 0.75 * u + 0.25 * (u1 + dt * self.rhs(u1, t + dt))
        th = t + 0.5 * dt
        return u / 3 + (2 / 3) * (u2 + dt * self.rhs(u2, th))

    def step_adaptive(self, u):
        # solver_1d.py:148-179 (whole-batch max|u|, Euler/RK sub-steps)
        target = self.dt
        cap = 1000
        n_sub = 0
        max_u = None
        while target > 0:
            max_u = torch.max(torch.abs(u))
            mu = max_u.item()
            allowed = self.adaptive_CFL * self.dx / mu if mu > 0 else target
            dt_loc = min(target, allowed)
            u = self._advance(u, dt_loc)
            self.t = self.t + dt_loc
            target -= dt_loc
            if int(math.ceil(target / dt_loc)) > cap:
                if self.LOG is not None:
                    self.LOG.warning("Substeps exceeded max substeps")
                break
            n_sub += 1
        return u

    def step(self, u):
        if self.adaptive_CFL is not None:
            return self.step_adaptive(u)
        u_new = self._advance(u, self.dt)
        self.t = self.t + self.dt
        return u_new
