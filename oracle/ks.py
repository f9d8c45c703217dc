"""FP64 restatement of the reference Kuramoto-Sivashinsky solver (test
infrastructure only).

Follows ``solvers/solver_1d.py:198-288`` of the reference (ETD1, ETRK2, the
double step, correction_s / correction_u), restated in complex128 on the CPU,
including the reference's full complex FFT, its no-op "dealias" (the
pad-square-truncate of solver_1d.py:228-232 equals u**2) and its ``.real``
projection, which drops the imaginary Nyquist term that ik*fft(u^2) leaves.

The ETDRK4 scheme (Cox & Matthews 2002) with Kassam & Trefethen's (2005)
contour-integral coefficients and the 2/3-rule mask on the nonlinear term has
no counterpart in the reference; it is added here as scheme 'ETDRK4'.
"""
import math
from types import SimpleNamespace

import torch

DT = torch.float64
CT = torch.complex128


def ks_params(domain_size, resolution=64, dt=1e-2, time_scheme="ETRK2",
              alpha=None, dealias=True):
    """Constructor-parameter object with the attributes KSSolverTorch reads
    (config/config_1d.py:16-41, solver_1d.py:199-212)."""
    ds = torch.as_tensor(domain_size, dtype=DT).reshape(-1)
    if alpha is None:
        alpha = torch.zeros(ds.numel(), dtype=DT)
    return SimpleNamespace(resolution=resolution, dt=dt, device="cpu",
                           domain_size=ds, alpha=torch.as_tensor(alpha, dtype=DT),
                           time_scheme=time_scheme, dealias=dealias,
                           batch_size=ds.numel())


def etd_coefficients(lop, dt, scheme):
    """E, phi1, phi2 of the reference's exponential integrators in float64.

    ETD1 (solver_1d.py:245-246): phi1 = dt where Lop == 0, else (E-1)/Lop.
    ETRK2 (solver_1d.py:259-267): phi1 = dt, phi2 = dt/2 where |Lop| < 1e-6,
    else phi1 = (E-1)/Lop, phi2 = (E-1-Lop*dt)/(Lop^2*dt).
    ``lop`` is real (k^2-k^4); returns real tensors.
    """
    lop = lop.to(DT)
    e = torch.exp(lop * dt)
    if scheme == "ETD1":
        safe = torch.where(lop == 0, torch.ones_like(lop), lop)
        phi1 = torch.where(lop == 0, torch.full_like(lop, dt), (e - 1) / safe)
        return e, phi1, torch.zeros_like(lop)
    small = lop.abs() < 1e-6
    safe = torch.where(small, torch.ones_like(lop), lop)
    phi1 = torch.where(small, torch.full_like(lop, dt), (e - 1) / safe)
    phi2 = torch.where(small, torch.full_like(lop, dt / 2),
                       (e - 1 - safe * dt) / (safe ** 2 * dt))
    return e, phi1, phi2


def etdrk4_coefficients(lop, h, m=64):
    """Cox-Matthews ETDRK4 coefficients by the Kassam-Trefethen contour
    integral, float64.

    For each real linear-operator value Lop the functions are averaged over
    ``m`` points z = h*Lop + exp(i*pi*(j-1/2)/m), j = 1..m, on the upper half
    of the unit circle centred at h*Lop; the real part of the mean is the
    value at the centre (the integrands are entire).  Returns
    (E, E2, Q, f1, f2, f3), each with the factor h folded in where the
    scheme multiplies a nonlinear term:
        E = e^{hL}, E2 = e^{hL/2}, Q = h*mean((e^{z/2}-1)/z),
        f1 = h*mean((-4-z+e^z(4-3z+z^2))/z^3),
        f2 = h*mean((2+z+e^z(z-2))/z^3),
        f3 = h*mean((-4-3z-z^2+e^z(4-z))/z^3).
    """
    lop = lop.to(DT).reshape(-1)
    j = torch.arange(1, m + 1, dtype=DT)
    r = torch.exp(1j * math.pi * (j - 0.5) / m).to(CT)
    z = (h * lop).to(CT)[:, None] + r[None, :]
    ez = torch.exp(z)
    q = h * ((torch.exp(z / 2) - 1) / z).mean(-1).real
    f1 = h * ((-4 - z + ez * (4 - 3 * z + z ** 2)) / z ** 3).mean(-1).real
    f2 = h * ((2 + z + ez * (z - 2)) / z ** 3).mean(-1).real
    f3 = h * ((-4 - 3 * z - z ** 2 + ez * (4 - z)) / z ** 3).mean(-1).real
    return (torch.exp(h * lop), torch.exp(h * lop / 2), q, f1, f2, f3)


class KSOracle:
    """Restates KSSolverTorch (solver_1d.py:198-288) in float64 / complex128,
    plus the 'ETDRK4' scheme.  ``step(u, domain_size)`` is the reference's
    bound-method dispatch (solver_1d.py:211-212) with the double step
    (solver_1d.py:202,280-288)."""

    def __init__(self, p):
        self.resolution = p.resolution
        self.dt = p.dt
        self.doublestep = True
        self.domain_size = p.domain_size
        self.alpha = p.alpha
        self.correction_s = None
        self.correction_u = None
        self.dealias = getattr(p, "dealias", True)
        n = self.resolution
        # solver_1d.py:208-210: integer wavenumbers, index n/2 holds -n/2
        self.fft_indices = torch.fft.fftfreq(n, d=1.0, dtype=DT) * n
        schemes = {"ETRK2": self.etrk2, "ETD1": self.etd1, "ETDRK4": self.etdrk4}
        if p.time_scheme not in schemes:
            raise KeyError(p.time_scheme)
        self.time_scheme = p.time_scheme
        self.step = schemes[p.time_scheme]

    def initial_conditions(self, generator=None):
        # solver_1d.py:214-223 (endpoint-inclusive linspace, global RNG)
        ds = self.domain_size.unsqueeze(1)
        al = self.alpha.unsqueeze(1)
        if len(al) != len(ds):
            raise ValueError("Length of alpha and domain_size must be the same")
        x = torch.linspace(0, 1, self.resolution, dtype=DT)
        u = torch.cos(2 * math.pi * x) + al * torch.sin(2 * math.pi * x)
        eps = torch.randn(u.shape, dtype=DT, generator=generator)
        return u + eps / torch.linalg.norm(eps)

    # solver_1d.py:235-240
    def calc_linear(self, domain_size):
        ds = domain_size.to(DT).unsqueeze(-1)
        ik = (2j * math.pi * self.fft_indices) / ds
        lop = -ik ** 2 - ik ** 4
        return lop, ik

    # solver_1d.py:225-233 (the pad/truncate there is the identity on u**2)
    def calc_nonlinear(self, u, ik, mask=None):
        g = -0.5 * ik
        if mask is not None:
            g = g * mask
        return g * torch.fft.fft(u ** 2)

    def _shat(self):
        return torch.fft.fft(self.correction_s) if self.correction_s is not None else 0.0

    def etd1_step(self, u, domain_size):
        lop, ik = self.calc_linear(domain_size)
        e, phi1, _ = etd_coefficients(lop.real, self.dt, "ETD1")
        vhat = e * torch.fft.fft(u) + phi1 * (self.calc_nonlinear(u, ik) + self._shat())
        return torch.fft.ifft(vhat).real

    def etrk2_step(self, u, domain_size):
        lop, ik = self.calc_linear(domain_size)
        e, phi1, phi2 = etd_coefficients(lop.real, self.dt, "ETRK2")
        n1 = self.calc_nonlinear(u, ik)
        mid_hat = e * torch.fft.fft(u) + phi1 * (n1 + self._shat())
        mid = torch.fft.ifft(mid_hat).real
        n2 = self.calc_nonlinear(mid, ik)
        return torch.fft.ifft(mid_hat + phi2 * (n2 - n1)).real

    def dealias_mask(self):
        n = self.resolution
        return (self.fft_indices.abs() < n / 3).to(DT)

    def etdrk4_coefficients(self, domain_size):
        lop, ik = self.calc_linear(domain_size)
        b, n = lop.shape
        co = etdrk4_coefficients(lop.real.reshape(-1), self.dt)
        return [c.reshape(b, n) for c in co], ik

    def etdrk4_step(self, u, domain_size):
        (e, e2, q, f1, f2, f3), ik = self.etdrk4_coefficients(domain_size)
        mask = self.dealias_mask() if self.dealias else None
        sh = self._shat()

        def nl(w):
            return self.calc_nonlinear(w, ik, mask) + sh

        v = torch.fft.fft(u)
        nv = nl(u)
        a = e2 * v + q * nv
        na = nl(torch.fft.ifft(a).real)
        b = e2 * v + q * na
        nb = nl(torch.fft.ifft(b).real)
        c = e2 * a + q * (2 * nb - nv)
        nc = nl(torch.fft.ifft(c).real)
        v1 = e * v + f1 * nv + 2 * f2 * (na + nb) + f3 * nc
        return torch.fft.ifft(v1).real

    def _wrap(self, sub):
        def run(u, domain_size):
            if self.correction_u is not None:
                u = u + self.correction_u
            u1 = sub(u, domain_size)
            return sub(u1, domain_size) if self.doublestep else u1
        return run

    def etd1(self, u, domain_size):
        return self._wrap(self.etd1_step)(u, domain_size)

    def etrk2(self, u, domain_size):
        return self._wrap(self.etrk2_step)(u, domain_size)

    def etdrk4(self, u, domain_size):
        return self._wrap(self.etdrk4_step)(u, domain_size)
