"""Constructor-parameter objects of the 1D solvers.

``KSSimParams`` keeps the reference's constructor (config/config_1d.py:16-41)
and adds ``dealias`` (2/3-rule mask of the ETDRK4 scheme) and ``pack_pairs``
(two trajectories per complex FFT).  ``BGSimParams``
is imported by the reference trainer (trainer/trainer_1d.py:344-355) but is
missing from its config module; this one carries every attribute
BurgersSolverTorch reads (solver_1d.py:46-66, 76-84, 106-111), with the
trainer's keyword arguments.
"""
import math

import torch


class KSSimParams:
    """Parameters of KSSolverTorch (config/config_1d.py:16-41)."""

    def __init__(self, domain_size, resolution=64, dt=1e-2, random_ratio=1, batch_size=None,
                 device="cuda", time_scheme="ETRK2", total_time=50, dealias=True, pack_pairs=True):
        self.dtype = torch.float64
        self.resolution = resolution
        self.dt = dt
        self.random_ratio = random_ratio
        self.device = device
        self.time_scheme = time_scheme
        self.total_time = total_time
        self.warm_up = 2000
        self.dealias = dealias
        # two real trajectories share one complex FFT (default); False runs one
        # trajectory per FFT, which keeps a trajectory finite when its FFT
        # partner holds a NaN or Inf, as the reference's per-row torch.fft does
        self.pack_pairs = pack_pairs
        self.set_domain_size(domain_size)
        self.batch_size = len(self.domain_size) if batch_size is None else batch_size
        self.init_alpha()

    def set_domain_size(self, domain_size):
        # config_1d.py:30-38
        if isinstance(domain_size, (list, tuple)):
            domain_size = torch.tensor(domain_size, device=self.device)
        elif isinstance(domain_size, torch.Tensor):
            domain_size = domain_size.clone().detach()
        else:
            raise ValueError("domain_size must be a list, tuple, or tensor, but got:", type(domain_size))
        self.domain_size = domain_size.repeat_interleave(self.random_ratio)

    def init_alpha(self):
        self.alpha = torch.rand(self.batch_size, device=self.device) - 0.5


class BGSimParams:
    """Parameters of BurgersSolverTorch.

    ``nu`` may be a float or a [batch, 1] tensor; the forcing parameters
    ``A, omega, phi, l`` are [batch, 1, J] tensors ('learning' equation type,
    data/data_loader_1d.py:23-29).  ``time_scheme`` selects 'Euler' (the
    reference's update, solver_1d.py:192) or 'SSPRK3'.
    """

    def __init__(self, dt=0.01, resolution=128, L=2 * math.pi, nu=0.01, equation_type="homogeneous",
                 batch_size=1, adaptive_CFL=None, convection_term="WENO", device="cuda", beta=0.0,
                 time_scheme="Euler", A=None, omega=None, phi=None, l=None, adaptive_mode="batch",
                 max_substeps=64):
        self.dt = dt
        self.resolution = resolution
        self.L = L
        self.nu = nu
        self.equation_type = equation_type
        self.batch_size = batch_size
        self.adaptive_CFL = adaptive_CFL
        # 'batch': the reference's sub-stepping (one dt for the whole batch from
        # max|u| over it, read back every sub-step); 'trajectory': each
        # trajectory sub-steps by its own max|u| on the GPU, at most
        # max_substeps per step, without host synchronisation
        self.adaptive_mode = adaptive_mode
        self.max_substeps = max_substeps
        self.convection_term = convection_term
        self.device = device
        self.beta = beta
        self.time_scheme = time_scheme
        self.A, self.omega, self.phi, self.l = A, omega, phi, l
        self.x = torch.arange(resolution, device=device, dtype=torch.float64) * (L / resolution)

    @property
    def J(self):
        return 0 if self.A is None else int(self.A.shape[-1])
