"""Rollout driver with the reference's signature (trainer/trainer_1d.py:154-183).

``simulate_rollout(model, solver, u_initial, mstep, correction_term, L)``
returns (results [B, mstep, D], corrections [B, mstep, D] or None) exactly
as the reference does.  Two MI355X additions:

* with ``correction_term=None`` and an inc_amd solver on a fixed time step,
  the whole rollout is ONE kernel launch (state resident in registers/LDS,
  every step's output written once) and its gradient ONE adjoint launch;
* with ``correction_term='INC'``, a Burgers solver and a CorrectionCNN the
  kernel runs, the correction network is evaluated inside that one launch
  as well, and with a gradient wanted (training) its transpose inside the
  one adjoint launch (float32, width 32, up to 4 layers, N <= 128);
* with ``correction_term='INC'``, a KS solver and a CorrectionCNN, the
  rollout is one library call (the network on this library's convolution
  kernels between the ETD steps) and its gradient one more;
* ``checkpoint_every=k`` recomputes the INC rollout in segments of k steps
  during the backward pass (segment-start states are the only activations
  kept), for long unrolled horizons.

``simulate_rollout_mse(solver, u_initial, target, L)`` is a plain rollout
followed by the trainer's MSE loss (trainer/trainer_1d.py:723) with the loss
fused into the kernels.
"""
import torch
from torch.utils.checkpoint import checkpoint

from .correction import CorrectionCNN
from .solver_1d import BurgersSolverTorch, KSSolverTorch, inc_train_supported


def _clone_t(t):
    return t.clone() if isinstance(t, torch.Tensor) else t


def _fusable(solver, correction_term):
    if correction_term is not None:
        return False
    if isinstance(solver, BurgersSolverTorch):
        return solver.adaptive_CFL is None or solver.adaptive_mode == "trajectory"
    if isinstance(solver, KSSolverTorch):
        return solver.correction_u is None
    return False


def _fusable_ks_inc(model, solver, u):
    """A KS INC rollout the library runs in one call (inc_ks_inc_fwd / _adj):
    a CorrectionCNN its convolution kernels run, no correction_u."""
    return (isinstance(solver, KSSolverTorch) and isinstance(model, CorrectionCNN) and u.is_cuda
            and solver.correction_u is None and model.ragged_desc(u.dtype) is not None
            and 2 * u.shape[-1] >= model.convs[0].k)


def _fusable_inc(model, solver, u):
    """An INC rollout the Burgers kernel runs with the correction network
    inside (fixed time step, WENO, a CorrectionCNN of a supported shape,
    within the kernel's size limits; with a gradient wanted, within the
    training adjoint's)."""
    if not isinstance(solver, BurgersSolverTorch) or not isinstance(model, CorrectionCNN):
        return False
    if solver.adaptive_CFL is not None or solver.param.convection_term != "WENO" or not u.is_cuda:
        return False
    n = u.shape[-1]
    if torch.is_grad_enabled() and (u.requires_grad or any(p.requires_grad for p in model.parameters())):
        return model.fused_desc(u.dtype) is not None and inc_train_supported(model, n, u.dtype)
    if u.dtype not in (torch.float32, torch.float64) or n > (512 if u.dtype == torch.float32 else 256):
        return False
    fd = model.fused_desc(u.dtype)
    return fd is not None and (fd[0].channels + 2) * n * u.element_size() <= 150 * 1024 and fd[0].kernel <= n


def simulate_rollout(model, solver, u_initial, mstep, correction_term, L=None, checkpoint_every=None, fused=True):
    """Unroll ``mstep`` solver steps (trainer_1d.py:154-183).  ``fused=False``
    keeps the step-by-step loop (one network evaluation and one solver launch
    per step) where a fused rollout kernel would otherwise run."""
    b, d = u_initial.shape
    if correction_term not in (None, "INC"):
        raise ValueError(f"Invalid correction term: {correction_term}")
    solver.correction_s = None
    if mstep <= 0:
        empty = u_initial.new_zeros(b, 0, d)
        return empty, (empty.clone() if correction_term else None)
    if fused and correction_term == "INC" and _fusable_inc(model, solver, u_initial):
        return solver.rollout_inc(model, u_initial, mstep)
    # KS INC with a CorrectionCNN: one call per direction.  Its adjoint reads
    # every step's stored hidden activations (solver.keep_activations), or,
    # with checkpoint_every given, keeps none and evaluates the network
    # again per step
    if fused and correction_term == "INC" and _fusable_ks_inc(model, solver, u_initial):
        return solver.rollout_inc(model, u_initial, mstep, L, keep_activations=False if checkpoint_every else None)
    # checkpoint_every=k on a fused rollout: its adjoint recomputes the stage
    # states k steps at a time instead of keeping them for the whole rollout
    # (fixed time step; the adaptive rollout keeps the step loop for it)
    adaptive = isinstance(solver, BurgersSolverTorch) and solver.adaptive_CFL is not None
    if fused and _fusable(solver, correction_term) and not (checkpoint_every and adaptive):
        if isinstance(solver, KSSolverTorch):
            return solver.rollout(u_initial, mstep, L, checkpoint_every=checkpoint_every), None
        return solver.rollout(u_initial, mstep, checkpoint_every=checkpoint_every), None

    model_fn = (lambda u: model(u)) if L is None else (lambda u: model(u, L))
    step_fn = (lambda u: solver.step(u)) if L is None else (lambda u: solver.step(u, L))

    def run(u, nsteps):
        outs, corr = [], []
        for _ in range(nsteps):
            if correction_term == "INC":
                s = model_fn(u)
                solver.correction_s = s
                corr.append(s)
            u = step_fn(u)
            outs.append(u)
        res = torch.stack(outs, 1)
        return res, (torch.stack(corr, 1) if corr else res.new_zeros(b, 0, d))

    k = checkpoint_every
    if not k or k >= mstep or not torch.is_grad_enabled():
        res, corr = run(u_initial, mstep)
        return res, (corr if correction_term == "INC" else None)

    results, corrections = [], []
    u = u_initial
    for k0 in range(0, mstep, k):
        n_seg = min(k, mstep - k0)
        t_start = _clone_t(getattr(solver, "t", None))
        state = {"first": True}

        def seg(u_in, n_seg=n_seg, t_start=t_start, state=state):
            t_after = getattr(solver, "t", None)
            if t_start is not None:
                solver.t = _clone_t(t_start)
            out = run(u_in, n_seg)
            if not state["first"] and t_start is not None:
                solver.t = t_after  # recomputation must not move the solver clock
            state["first"] = False
            return out

        res, corr = checkpoint(seg, u, use_reentrant=False)
        results.append(res)
        corrections.append(corr)
        u = res[:, -1]
    res = torch.cat(results, 1)
    return res, (torch.cat(corrections, 1) if correction_term == "INC" else None)


def simulate_rollout_mse(solver, u_initial, target, L=None, model=None):
    """``simulate_rollout(model, solver, u_initial, target.shape[1], ...)``
    followed by ``torch.nn.MSELoss()(results, target)`` (trainer_1d.py:154-183
    and :716-723), with the loss computed in the forward kernel and its
    gradient seeded inside the adjoint kernel.  Without ``model`` (no
    correction) returns (loss, results) and needs a fused rollout: a fixed
    time step (Burgers) or no correction_u (KS).  With a CorrectionCNN
    ``model`` (the trainer's 'INC' rollout) the network runs inside the
    Burgers rollout and adjoint kernels too, and (loss, results,
    corrections) is returned."""
    solver.correction_s = None
    if model is not None:
        if not isinstance(solver, BurgersSolverTorch) or not isinstance(model, CorrectionCNN):
            raise ValueError("the fused INC loss needs a BurgersSolverTorch and a CorrectionCNN")
        return solver.rollout_inc_mse(model, u_initial, target)
    if isinstance(solver, KSSolverTorch):
        return solver.rollout_mse(u_initial, target, L)
    if isinstance(solver, BurgersSolverTorch):
        return solver.rollout_mse(u_initial, target)
    raise ValueError(f"simulate_rollout_mse needs an inc_amd solver (got {type(solver).__name__})")
