"""FP64 CPU oracle for the INC 1D solver step -- TEST INFRASTRUCTURE ONLY.

This package restates, in float64 PyTorch on the CPU, the algorithm of the
reference's differentiable 1D solvers (``solvers/solver_1d.py``) and of the
rollout driver (``trainer/trainer_1d.py:154-183``), plus the schemes the
MI355X build adds on the same path (SSP-RK3 for Burgers, ETDRK4 with 2/3-rule
dealiasing for Kuramoto-Sivashinsky).  Gradients come from PyTorch autograd
through the restated op sequence, exactly as the reference gets its gradients.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or, for the
CPU baseline, as the thing timed on the host).  The product path
(``inc_amd``) never imports it: when the HIP library is missing, the product
raises instead of falling back here.

Pinning (what checks each function against something other than itself):

* ``weno_reconstruction`` (solver_1d.py:4-42): golden vectors written from the
  reference (tests/golden), closed forms (linear-exactness, equal-beta ->
  linear weights) in tests/test_oracle_burgers.py.
* Burgers RHS / Euler step / FDM / forcing (solver_1d.py:74-195): golden
  vectors (hand-composed Euler step from reference pieces, forcing), mean
  conservation, Laplacian eigenvalues.
* SSP-RK3 (no reference counterpart): amplification factor
  1+z+z^2/2+z^3/6 on a single diffusive mode; constant-state fixed point;
  the stage times t, t+dt, t+dt/2 by Simpson's rule on a spatially constant
  state under spatially constant forcing; third order in dt with
  time-dependent forcing under transport (a wrong stage time gives order 1).
* KS ETD1 / ETRK2 / double step (solver_1d.py:198-288): golden vectors
  (including correction_s and per-trajectory L), single-mode exponential decay,
  mean conservation, constant forcing integrates to +s*dt.
* ETDRK4 + 2/3 dealias (no reference counterpart): contour-integral
  coefficients checked against the direct formulas (large |hL|) and Taylor
  series (small |hL|); f1+4f2+f3 = (e^{hL}-1)/L identity; fourth-order
  convergence in dt on a smooth solution; constant forcing exact.  The 2/3
  mask: a single mode n/6 < k < n/3 evolves by the exact linear map with the
  mask (its product is cut) and not without it; two modes whose products
  reach floor(n/3) and floor(n/3)+1 show the cut between them (n = 256, 512).
* Adjoints: autograd through this oracle versus FP64 central differences and
  versus reference autograd gradients in tests/golden.
* ``initial_conditions`` (solver_1d.py:214-223): the deterministic part
  cos(2 pi x) + alpha sin(2 pi x) on the endpoint-inclusive grid, and the
  noise's unit Frobenius norm over the whole batch.  The random draws
  themselves come from a different generator on the GPU and are not
  compared value by value.
"""
from .burgers import (BurgersOracle, weno_reconstruction, burgers_params,
                      WENO_TABLE, WENO_EPS, WENO_GAMMA)
from .ks import KSOracle, ks_params, etdrk4_coefficients, etd_coefficients
from .rollout import simulate_rollout

__all__ = [
    "BurgersOracle", "weno_reconstruction", "burgers_params", "WENO_TABLE",
    "WENO_EPS", "WENO_GAMMA", "KSOracle", "ks_params", "etdrk4_coefficients",
    "etd_coefficients", "simulate_rollout",
]
