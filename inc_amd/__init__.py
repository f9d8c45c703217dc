"""inc_amd -- MI355X-native differentiable 1D solver step of INC.

Drop-in replacements for the reference's ``solvers.solver_1d`` classes and
``trainer.trainer_1d.simulate_rollout``, running hand-written HIP kernels
for gfx950 (libinc_amd.so, C ABI in include/inc_amd.h).
"""
from .batched import RaggedBurgers
from .config_1d import BGSimParams, KSSimParams
from .correction import CorrectionCNN, ResNet1D, UNet1D
from .rollout import simulate_rollout, simulate_rollout_mse
from .solver_1d import BurgersSolverTorch, KSSolverTorch, weno_reconstruction

__all__ = ["BurgersSolverTorch", "KSSolverTorch", "weno_reconstruction", "simulate_rollout",
           "simulate_rollout_mse", "BGSimParams", "KSSimParams", "RaggedBurgers", "CorrectionCNN", "ResNet1D", "UNet1D"]
__version__ = "0.1.0"
