"""Drop-in MI355X solvers: BurgersSolverTorch, KSSolverTorch,
weno_reconstruction.

Same constructors, attributes and method signatures as the reference's
``solvers/solver_1d.py`` (weno_reconstruction :4-42, BurgersSolverTorch
:44-195, KSSolverTorch :198-288).  Every step runs in libinc_amd (HIP,
gfx950) through the C ABI of include/inc_amd.h; gradients come from the
library's hand-written adjoint kernels behind ``torch.autograd.Function``s.
Tensors must live on the GPU; nothing falls back to the CPU.
"""
import ctypes
import math

import torch

from . import _lib as L
from . import autotune

__all__ = ["BurgersSolverTorch", "KSSolverTorch", "weno_reconstruction"]


# ---------------------------------------------------------------------------
# weno_reconstruction (solver_1d.py:4-42)
# ---------------------------------------------------------------------------
class _Weno(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, direction):
        fc = f.contiguous()
        out = torch.empty_like(fc)
        n = fc.shape[-1]
        rows = fc.numel() // n if n else 0
        L.check(L.load().inc_weno_fwd(L.dtype_code(fc.dtype), direction, L.ptr(fc), L.ptr(out), rows, n,
                                      L.stream()), "weno_reconstruction")
        ctx.save_for_backward(fc)
        ctx.direction = direction
        return out

    @staticmethod
    def backward(ctx, g):
        (f,) = ctx.saved_tensors
        g = g.contiguous()
        gf = torch.empty_like(f)
        n = f.shape[-1]
        rows = f.numel() // n if n else 0
        L.check(L.load().inc_weno_adj(L.dtype_code(f.dtype), ctx.direction, L.ptr(f), L.ptr(g), L.ptr(gf),
                                      rows, n, L.stream()), "weno_reconstruction backward")
        return gf, None


def weno_reconstruction(f, direction="left"):
    """WENO5 reconstruction along the last axis with the reference's stencils
    ('left': value at x_{i+1/2}; 'right': the reference's shifted stencil)."""
    if direction not in ("left", "right"):
        raise ValueError("Direction must be 'left' or 'right'.")
    L.require_gpu(f)
    return _Weno.apply(f, 0 if direction == "left" else 1)


# ---------------------------------------------------------------------------
# Burgers
# ---------------------------------------------------------------------------
_SOURCES = {"homogeneous": L.INC_SRC_NONE, "viscosity": L.INC_SRC_NONE,
            "linear": L.INC_SRC_LINEAR, "learning": L.INC_SRC_LEARNING}


def _burgers_correction(s, u):
    """correction_s as the [B, N] array the kernels read.  The reference adds
    it by broadcasting (source + correction_s, solver_1d.py:116-123), so a
    scalar, [N] or [B, 1] correction is expanded to the state's shape here;
    autograd sums its gradient back to the given shape."""
    if s is None:
        return None
    if not isinstance(s, torch.Tensor):
        s = torch.as_tensor(s, device=u.device)
    L.require_gpu(s)
    s = s.to(dtype=u.dtype)
    if s.shape != u.shape:
        try:
            s = s.expand(u.shape)
        except RuntimeError:
            raise ValueError(f"correction_s of shape {tuple(s.shape)} does not broadcast to the state's "
                             f"{tuple(u.shape)}") from None
    return s


def _ks_correction(s, u):
    """correction_s as the [B, N] array the kernels read.  The reference
    transforms it along its last axis and adds the spectrum by broadcasting
    (solver_1d.py:249-253), so [N] (one correction for every trajectory) and
    [B, N] are the shapes with that meaning; [N] is expanded here."""
    if s is None:
        return None
    L.require_gpu(s)
    s = s.to(dtype=u.dtype)
    if s.shape != u.shape:
        if s.dim() == 0 or s.shape[-1] != u.shape[-1]:
            raise ValueError(f"correction_s must be [N] or [batch, N] = {tuple(u.shape)} (got {tuple(s.shape)})")
        try:
            s = s.expand(u.shape)
        except RuntimeError:
            raise ValueError(f"correction_s of shape {tuple(s.shape)} does not broadcast to the state's "
                             f"{tuple(u.shape)}") from None
    return s


def _checkpoint_k(k):
    if k is None:
        return 0
    k = int(k)
    if k < 1:
        raise ValueError(f"checkpoint_every must be a positive number of steps (got {k})")
    return k


def _nu_input(solver):
    """The solver's viscosity when autograd is to differentiate it (a tensor
    that requires grad), else None: passed to the autograd Functions as an
    input so that its cotangent reaches it."""
    nu = solver.nu
    if torch.is_grad_enabled() and isinstance(nu, torch.Tensor) and nu.requires_grad:
        return nu
    return None


def _nu_cotangent(nu, g_nu):
    """Per-trajectory cotangent [batch] -> the gradient in nu's own shape,
    dtype and device (a one-element nu, broadcast over the batch, gets the
    sum)."""
    g = g_nu.sum() if nu.numel() == 1 else g_nu
    return g.reshape(nu.shape).to(dtype=nu.dtype, device=nu.device)


def _nu_buffer(ctx, index, b, like):
    return torch.empty(b, dtype=like.dtype, device=like.device) if ctx.needs_input_grad[index] else None


def _checkpointed_adjoint(g_traj, u0, traj, k, nst, fwd, adj, need_s, need_nu=False):
    """Reverse sweep of a fused rollout whose forward kept no stage states
    (checkpoint-and-recompute every k steps): for each segment of k steps,
    last first, the stages are recomputed from the segment's start state
    (u0, or the previous segment's last output, which the trajectory holds)
    into one buffer reused by every segment, then that segment's adjoint
    runs with the cotangent of its last output increased by the cotangent
    carried back from the segment after it.  Memory O(B k stages N) instead
    of O(B steps stages N).

    fwd(k0, start, out, m, stages) and adj(k0, stages, g_seg, g_u, g_s,
    g_nu, m) launch the kernels for the segment starting at step k0 with m
    steps.  Returns (g_u0, g_s summed over segments, g_nu summed)."""
    b, steps, n = g_traj.shape
    buf = g_traj.new_empty(b * min(k, steps) * nst * n)
    out = torch.empty_like(u0)
    g_s = torch.zeros_like(u0) if need_s else None
    g_s_seg = torch.empty_like(u0) if need_s else None
    g_nu = u0.new_zeros(b) if need_nu else None
    g_nu_seg = u0.new_empty(b) if need_nu else None
    lam = None
    for k0 in reversed(range(0, steps, k)):
        m = min(k, steps - k0)
        start = u0 if k0 == 0 else traj[:, k0 - 1].contiguous()
        stages = buf[: b * m * nst * n].view(b, m, nst, n)
        fwd(k0, start, out, m, stages)
        g_seg = g_traj[:, k0:k0 + m].contiguous()
        if lam is not None:
            g_seg[:, -1] += lam
        lam = torch.empty_like(u0)
        adj(k0, stages, g_seg, lam, g_s_seg, g_nu_seg, m)
        if need_s:
            g_s += g_s_seg
        if need_nu:
            g_nu += g_nu_seg
    return lam, g_s, g_nu


class _BurgersAdvance(torch.autograd.Function):
    """`steps` steps of the Burgers map; returns the final state, or the
    [B, steps, N] trajectory when ``want_traj``."""

    @staticmethod
    def forward(ctx, u, s, solver, dt, steps, t, want_traj, need_stages, nu, ckpt):
        u = u.contiguous()
        s = s.contiguous() if s is not None else None
        b, n = u.shape
        desc, keep = solver._desc(u, dt, t)
        ctx.nu = nu
        out = torch.empty_like(u)
        traj = u.new_empty(b, steps, n) if want_traj else None
        ns = 3 if solver.time_scheme == "SSPRK3" else 1
        # checkpoint every ckpt steps: no stage states now, recomputed per segment in backward
        ctx.ckpt = ckpt if (ckpt and need_stages and want_traj and ckpt < steps) else 0
        stages = u.new_empty(b, steps, ns, n) if need_stages and not ctx.ckpt else None
        rc = L.load().inc_burgers_fwd(ctypes.byref(desc), L.ptr(u), L.ptr(s), L.ptr(out), steps, L.ptr(traj),
                                      L.ptr(stages), L.stream())
        L.check(rc, "BurgersSolverTorch.step")
        if ctx.ckpt:
            ctx.save_for_backward(u, s, traj)
            ctx.ns, ctx.dt = ns, float(dt)
            ctx.t = t.detach().reshape(-1).to(u.dtype).clone() if isinstance(t, torch.Tensor) else float(t)
        else:
            ctx.save_for_backward(stages)
        ctx.desc, ctx.keep, ctx.steps, ctx.want_traj = desc, keep, steps, want_traj
        ctx.shape = (b, n)
        return traj if want_traj else out

    @staticmethod
    def _checkpointed_backward(ctx, g):
        u0, s, traj = ctx.saved_tensors
        lib = L.load()
        descs = {}

        def seg_desc(k0):
            # the segment's forcing clock starts k0 steps later
            if k0 not in descs:
                d = L.BurgersDesc.from_buffer_copy(ctx.desc)
                hold = None
                if isinstance(ctx.t, torch.Tensor):
                    hold = ctx.t + k0 * ctx.dt
                    d.t_ptr = hold.data_ptr()
                else:
                    d.t0 = ctx.t + k0 * ctx.dt
                descs.clear()
                descs[k0] = (d, hold)
            return descs[k0][0]

        def fwd(k0, start, out, m, stages):
            rc = lib.inc_burgers_fwd(ctypes.byref(seg_desc(k0)), L.ptr(start), L.ptr(s), L.ptr(out), m, None,
                                     L.ptr(stages), L.stream())
            L.check(rc, "BurgersSolverTorch rollout (checkpoint recompute)")

        def adj(k0, stages, g_seg, g_u, g_s, g_nu, m):
            rc = lib.inc_burgers_adj_nu(ctypes.byref(seg_desc(k0)), L.ptr(stages), L.ptr(g_seg), L.ptr(g_u),
                                        L.ptr(g_s), L.ptr(g_nu), m, L.stream())
            L.check(rc, "BurgersSolverTorch rollout backward (checkpointed)")

        g_u, g_s, g_nu = _checkpointed_adjoint(g.contiguous(), u0, traj, ctx.ckpt, ctx.ns, fwd, adj,
                                               ctx.needs_input_grad[1], ctx.needs_input_grad[8])
        g_nu = _nu_cotangent(ctx.nu, g_nu) if g_nu is not None else None
        return g_u, g_s, None, None, None, None, None, None, g_nu, None

    @staticmethod
    def backward(ctx, g):
        if ctx.ckpt:
            return _BurgersAdvance._checkpointed_backward(ctx, g)
        (stages,) = ctx.saved_tensors
        if stages is None:
            raise RuntimeError("Burgers step was run without saved stages; enable grad before the step")
        b, n = ctx.shape
        if ctx.want_traj:
            g_traj = g.contiguous()
        elif ctx.steps == 1:
            g_traj = g.contiguous().view(b, 1, n)
        else:
            g_traj = g.new_zeros(b, ctx.steps, n)
            g_traj[:, -1] = g
        g_u = torch.empty(b, n, dtype=g.dtype, device=g.device)
        g_s = torch.empty_like(g_u) if ctx.needs_input_grad[1] else None
        g_nu = _nu_buffer(ctx, 8, b, g_u)
        rc = L.load().inc_burgers_adj_nu(ctypes.byref(ctx.desc), L.ptr(stages), L.ptr(g_traj), L.ptr(g_u),
                                         L.ptr(g_s), L.ptr(g_nu), ctx.steps, L.stream())
        L.check(rc, "BurgersSolverTorch.step backward")
        g_nu = _nu_cotangent(ctx.nu, g_nu) if g_nu is not None else None
        return g_u, g_s, None, None, None, None, None, None, g_nu, None


def _mse_target(target, u, steps):
    b, n = u.shape
    if tuple(target.shape) != (b, steps, n):
        raise ValueError(f"target must be [batch, steps, n] = {(b, steps, n)} (got {tuple(target.shape)})")
    L.require_gpu(target)
    return target.detach().to(dtype=u.dtype).contiguous()


def _mse_backward_common(ctx, g_loss, g_traj, traj, target):
    """Device scale 2 g_loss / M for the adjoint, the optional extra
    trajectory cotangent, and the target's gradient when it is needed."""
    m = traj.numel()
    scale = (g_loss.detach().to(traj.dtype) * (2.0 / m)).reshape(1).contiguous()
    g_traj = g_traj.contiguous() if g_traj is not None else None
    g_target = -scale * (traj - target) if ctx.needs_input_grad[2] else None
    return scale, g_traj, g_target


class _BurgersRolloutMSE(torch.autograd.Function):
    """`steps` steps of the Burgers map with the MSE to a target trajectory
    fused into the kernels (inc_burgers_fwd_mse / _adj_mse); returns
    (loss, [B, steps, N] trajectory).  The loss is the mean over all of
    batch, steps and points, torch.nn.MSELoss's default."""

    @staticmethod
    def forward(ctx, u, s, target, solver, dt, steps, t, need_stages, nu):
        ctx.set_materialize_grads(False)
        ctx.nu = nu
        u = u.contiguous()
        s = s.contiguous() if s is not None else None
        b, n = u.shape
        tgt = _mse_target(target, u, steps)
        desc, keep = solver._desc(u, dt, t)
        out = torch.empty_like(u)
        traj = u.new_empty(b, steps, n)
        ns = 3 if solver.time_scheme == "SSPRK3" else 1
        stages = u.new_empty(b, steps, ns, n) if need_stages else None
        parts = u.new_empty(b, 4)
        rc = L.load().inc_burgers_fwd_mse(ctypes.byref(desc), L.ptr(u), L.ptr(s), L.ptr(out), steps, L.ptr(traj),
                                          L.ptr(stages), L.ptr(tgt), L.ptr(parts), L.stream())
        L.check(rc, "BurgersSolverTorch.rollout_mse")
        ctx.save_for_backward(stages, traj, tgt)
        ctx.desc, ctx.keep, ctx.steps, ctx.shape = desc, keep, steps, (b, n)
        return parts.sum() / traj.numel(), traj

    @staticmethod
    def backward(ctx, g_loss, g_traj):
        stages, traj, tgt = ctx.saved_tensors
        if stages is None:
            raise RuntimeError("Burgers rollout was run without saved stages; enable grad before the rollout")
        if g_loss is None:
            g_loss = traj.new_zeros(())
        scale, g_traj, g_target = _mse_backward_common(ctx, g_loss, g_traj, traj, tgt)
        b, n = ctx.shape
        g_u = torch.empty(b, n, dtype=traj.dtype, device=traj.device)
        g_s = torch.empty_like(g_u) if ctx.needs_input_grad[1] else None
        g_nu = _nu_buffer(ctx, 8, b, g_u)
        rc = L.load().inc_burgers_adj_mse_nu(ctypes.byref(ctx.desc), L.ptr(stages), L.ptr(traj), L.ptr(tgt),
                                             L.ptr(scale), L.ptr(g_traj), L.ptr(g_u), L.ptr(g_s), L.ptr(g_nu),
                                             ctx.steps, L.stream())
        L.check(rc, "BurgersSolverTorch.rollout_mse backward")
        g_nu = _nu_cotangent(ctx.nu, g_nu) if g_nu is not None else None
        return g_u, g_s, g_target, None, None, None, None, None, g_nu


class _BurgersIncRollout(torch.autograd.Function):
    """simulate_rollout(model, solver, u, steps, 'INC') (trainer_1d.py:154-183)
    for a CorrectionCNN with the network evaluated inside the rollout kernel
    and its transpose inside the adjoint kernel (inc_burgers_inc_train_fwd /
    _adj); returns (results, corrections), both [B, steps, N],
    differentiable in u and in the network's parameters (``params``, in the
    order [w0, b0, w1, b1, ...] of its Conv1d layers)."""

    @staticmethod
    def forward(ctx, u, solver, model, steps, t, *params):
        ctx.set_materialize_grads(False)
        cdesc, keep_w = model.fused_desc(u.dtype)
        u = u.contiguous()
        b, n = u.shape
        desc, keep = solver._desc(u, solver.dt, t)
        out = torch.empty_like(u)
        traj = u.new_empty(b, steps, n)
        corr = u.new_empty(b, steps, n)
        ns = 3 if solver.time_scheme == "SSPRK3" else 1
        stages = u.new_empty(b, steps, ns, n)
        rc = L.load().inc_burgers_inc_train_fwd(ctypes.byref(desc), ctypes.byref(cdesc), L.ptr(u), L.ptr(out),
                                                steps, L.ptr(traj), L.ptr(corr), L.ptr(stages), L.stream())
        L.check(rc, "BurgersSolverTorch.rollout_inc")
        ctx.save_for_backward(stages)
        ctx.desc, ctx.keep, ctx.cdesc, ctx.keep_w = desc, keep, cdesc, keep_w
        ctx.steps, ctx.shape, ctx.param_shapes = steps, (b, n), [p.shape for p in params]
        return traj, corr

    @staticmethod
    def backward(ctx, g_traj, g_corr):
        (stages,) = ctx.saved_tensors
        b, n = ctx.shape
        g_traj = g_traj.contiguous() if g_traj is not None else None
        g_corr = g_corr.contiguous() if g_corr is not None else None
        g_u = stages.new_empty(b, n)
        sizes = [math.prod(sh) for sh in ctx.param_shapes]
        g_w = stages.new_empty(b, sum(sizes))
        rc = L.load().inc_burgers_inc_train_adj(ctypes.byref(ctx.desc), ctypes.byref(ctx.cdesc), L.ptr(stages),
                                                L.ptr(g_traj), L.ptr(g_corr), L.ptr(g_u), L.ptr(g_w), ctx.steps,
                                                L.stream())
        L.check(rc, "BurgersSolverTorch.rollout_inc backward")
        g_rows = g_w.sum(0).split(sizes)
        grads = [g.view(sh) for g, sh in zip(g_rows, ctx.param_shapes)]
        return (g_u, None, None, None, None, *grads)


class _BurgersIncRolloutMSE(torch.autograd.Function):
    """The INC training rollout of _BurgersIncRollout with the trainer's MSE
    to a target trajectory fused into its kernels (inc_burgers_inc_train_fwd_mse
    / _adj_mse, trainer_1d.py:716-723): returns (loss, results, corrections);
    the backward seeds the adjoint with 2 (u - target) / M inside the kernel."""

    @staticmethod
    def forward(ctx, u, target, solver, model, steps, t, *params):
        ctx.set_materialize_grads(False)
        cdesc, keep_w = model.fused_desc(u.dtype)
        u = u.contiguous()
        b, n = u.shape
        tgt = _mse_target(target, u, steps)
        desc, keep = solver._desc(u, solver.dt, t)
        out = torch.empty_like(u)
        traj = u.new_empty(b, steps, n)
        corr = u.new_empty(b, steps, n)
        ns = 3 if solver.time_scheme == "SSPRK3" else 1
        stages = u.new_empty(b, steps, ns, n)
        parts = u.new_empty(b, 4)
        rc = L.load().inc_burgers_inc_train_fwd_mse(ctypes.byref(desc), ctypes.byref(cdesc), L.ptr(u), L.ptr(out),
                                                    steps, L.ptr(traj), L.ptr(corr), L.ptr(stages), L.ptr(tgt),
                                                    L.ptr(parts), L.stream())
        L.check(rc, "BurgersSolverTorch.rollout_inc_mse")
        ctx.save_for_backward(stages, traj, tgt)
        ctx.desc, ctx.keep, ctx.cdesc, ctx.keep_w = desc, keep, cdesc, keep_w
        ctx.steps, ctx.shape, ctx.param_shapes = steps, (b, n), [p.shape for p in params]
        return parts.sum() / traj.numel(), traj, corr

    @staticmethod
    def backward(ctx, g_loss, g_traj, g_corr):
        stages, traj, tgt = ctx.saved_tensors
        b, n = ctx.shape
        if g_loss is None:
            g_loss = traj.new_zeros(())
        m = traj.numel()
        scale = (g_loss.detach().to(traj.dtype) * (2.0 / m)).reshape(1).contiguous()
        g_traj = g_traj.contiguous() if g_traj is not None else None
        g_corr = g_corr.contiguous() if g_corr is not None else None
        g_target = -scale * (traj - tgt) if ctx.needs_input_grad[1] else None
        g_u = stages.new_empty(b, n)
        sizes = [math.prod(sh) for sh in ctx.param_shapes]
        g_w = stages.new_empty(b, sum(sizes))
        rc = L.load().inc_burgers_inc_train_adj_mse(ctypes.byref(ctx.desc), ctypes.byref(ctx.cdesc), L.ptr(stages),
                                                    L.ptr(traj), L.ptr(tgt), L.ptr(scale), L.ptr(g_traj),
                                                    L.ptr(g_corr), L.ptr(g_u), L.ptr(g_w), ctx.steps, L.stream())
        L.check(rc, "BurgersSolverTorch.rollout_inc_mse backward")
        g_rows = g_w.sum(0).split(sizes)
        grads = [g.view(sh) for g, sh in zip(g_rows, ctx.param_shapes)]
        return (g_u, g_target, None, None, None, None, *grads)


def inc_train_shape_supported(channels, kernel, layers, n):
    """The shapes inc_burgers_inc_train_adj runs: width 32, kernel size 5 or
    7, 2..4 layers, 8 <= n <= 128, and the block's LDS (hidden weights,
    layers + 1 activation arrays of [n][C + 4] and 2 (K // 2) wrapped rows,
    two [n] arrays; the final
    reduction's [4][C/4][V] partials) within 160 KiB less the static part.
    The library's own check, restated so that a rollout can choose its path
    before launching."""
    c, k = channels, kernel
    if c != 32 or k not in (5, 7) or not 2 <= layers <= 4 or not max(k, 8) <= n <= 128:
        return False
    nh = layers - 2
    ws = nh * c * (c * k + 4)
    body = ws + (nh + 3) * n * (c + 4) + (k // 2) * 2 * (c + 4) + 2 * n
    red = ws + 4 * (c // 4) * ((4 * k + 4) + (4 * k + 1) + 4 * nh)
    return 4 * max(body, red) <= 160 * 1024 - 256


def inc_train_supported(model, n, dtype):
    """Whether the fused INC training rollout runs this CorrectionCNN at
    resolution n: float32 state and parameters on the GPU and a shape
    inc_train_shape_supported accepts."""
    convs = list(model.convs)
    if dtype != torch.float32 or any(p.dtype != dtype or not p.is_cuda for p in model.parameters()):
        return False  # the gradients are returned in the parameters' own dtype
    return inc_train_shape_supported(convs[0].weight.shape[0], convs[0].k, len(convs), n)


class _BurgersAdaptive(torch.autograd.Function):
    """`steps` steps of dt, each sub-stepped per trajectory to the CFL limit
    on the GPU (inc_burgers_adaptive_fwd); returns the final state or the
    [B, steps, N] trajectory.  The sub-step sizes are constants for the
    gradient, as with the reference's .item()."""

    @staticmethod
    def forward(ctx, u, s, solver, steps, t, want_traj, need_stages, nu):
        u = u.contiguous()
        s = s.contiguous() if s is not None else None
        b, n = u.shape
        desc, keep = solver._desc(u, solver.dt, t)
        ctx.nu = nu
        ms = solver.max_substeps
        ns = 3 if solver.time_scheme == "SSPRK3" else 1
        out = torch.empty_like(u)
        traj = u.new_empty(b, steps, n) if want_traj else None
        nsub = torch.empty(b, steps, dtype=torch.int32, device=u.device)
        dts = u.new_empty(b, steps, ms)
        stages = u.new_empty(b, steps, ms, ns, n) if need_stages else None
        rc = L.load().inc_burgers_adaptive_fwd(ctypes.byref(desc), float(solver.adaptive_CFL), ms, L.ptr(u),
                                               L.ptr(s), L.ptr(out), steps, L.ptr(traj), L.ptr(nsub), L.ptr(dts),
                                               L.ptr(stages), L.stream())
        L.check(rc, "BurgersSolverTorch.step (adaptive)")
        ctx.save_for_backward(stages, nsub, dts)
        ctx.desc, ctx.keep, ctx.steps, ctx.want_traj, ctx.ms = desc, keep, steps, want_traj, ms
        ctx.shape = (b, n)
        solver.last_substeps = nsub  # [B, steps] sub-steps taken (device tensor, read without a sync)
        return traj if want_traj else out

    @staticmethod
    def backward(ctx, g):
        stages, nsub, dts = ctx.saved_tensors
        if stages is None:
            raise RuntimeError("Burgers step was run without saved stages; enable grad before the step")
        b, n = ctx.shape
        if ctx.want_traj:
            g_traj = g.contiguous()
        elif ctx.steps == 1:
            g_traj = g.contiguous().view(b, 1, n)
        else:
            g_traj = g.new_zeros(b, ctx.steps, n)
            g_traj[:, -1] = g
        g_u = torch.empty(b, n, dtype=g.dtype, device=g.device)
        g_s = torch.empty_like(g_u) if ctx.needs_input_grad[1] else None
        g_nu = _nu_buffer(ctx, 7, b, g_u)
        rc = L.load().inc_burgers_adaptive_adj_nu(ctypes.byref(ctx.desc), ctx.ms, L.ptr(nsub), L.ptr(dts),
                                                  L.ptr(stages), L.ptr(g_traj), L.ptr(g_u), L.ptr(g_s),
                                                  L.ptr(g_nu), ctx.steps, L.stream())
        L.check(rc, "BurgersSolverTorch.step (adaptive) backward")
        g_nu = _nu_cotangent(ctx.nu, g_nu) if g_nu is not None else None
        return g_u, g_s, None, None, None, None, None, g_nu


class _BurgersRHS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, solver, parts, t, nu):
        u = u.contiguous()
        desc, keep = solver._desc(u, solver.dt, t)
        out = torch.empty_like(u)
        rc = L.load().inc_burgers_rhs(ctypes.byref(desc), parts, L.ptr(u), None, L.ptr(out), L.stream())
        L.check(rc, "BurgersSolverTorch right-hand side")
        ctx.save_for_backward(u)
        ctx.desc, ctx.keep, ctx.parts, ctx.nu = desc, keep, parts, nu
        return out

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        g = g.contiguous()
        g_u = torch.empty_like(u)
        g_nu = _nu_buffer(ctx, 4, u.shape[0], u)
        rc = L.load().inc_burgers_rhs_adj_nu(ctypes.byref(ctx.desc), ctx.parts, L.ptr(u), L.ptr(g), L.ptr(g_u),
                                             L.ptr(g_nu), L.stream())
        L.check(rc, "BurgersSolverTorch right-hand side backward")
        g_nu = _nu_cotangent(ctx.nu, g_nu) if g_nu is not None else None
        return g_u, None, None, None, g_nu


class BurgersSolverTorch:
    """Burgers u_t + (u^2/2)_x = nu u_xx + source + s on a periodic grid
    (solver_1d.py:44-195).

    Differences from the reference, deliberate:
      * the correction s (``correction_s``) enters the right-hand side as the
        reference's callers intend (its ``_apply_corrections`` has the wrong
        arity, solver_1d.py:116 vs :166,:188);
      * ``sim_params.time_scheme`` may select 'SSPRK3' (default 'Euler', the
        reference's update); the forcing is then evaluated at the stage times;
      * the diffusion is the 3-point stencil, not a dense N x N product.
    """

    def __init__(self, sim_params, LOG=None):
        self.param = sim_params
        self.resolution = sim_params.resolution
        self.dt = sim_params.dt
        self.x = sim_params.x
        self.t = 0.0
        self.adaptive_CFL = sim_params.adaptive_CFL
        self.adaptive_mode = getattr(sim_params, "adaptive_mode", "batch")
        self.max_substeps = int(getattr(sim_params, "max_substeps", 64))
        if self.adaptive_mode not in ("batch", "trajectory"):
            raise ValueError("adaptive_mode must be 'batch' or 'trajectory'")
        self.last_substeps = None
        self.correction_s = None
        self.L = sim_params.L
        self.nu = sim_params.nu
        self.device = sim_params.device
        self.dx = sim_params.L / sim_params.resolution
        self.LOG = LOG
        self.time_scheme = getattr(sim_params, "time_scheme", "Euler")
        if self.time_scheme not in ("Euler", "SSPRK3"):
            raise ValueError("time_scheme must be 'Euler' or 'SSPRK3'")
        team = getattr(sim_params, "team", 0) or 0
        self._team_auto = team == "auto" or (team == 0 and autotune.enabled())
        self.team = 0 if team == "auto" else int(team)
        if sim_params.convection_term == "FDM":
            self.calc_nonlinear = self.calc_nonlinear_FDM
            self._conv = L.INC_FDM
        elif sim_params.convection_term == "WENO":
            self.calc_nonlinear = self.calc_nonlinear_weno
            self._conv = L.INC_WENO
        else:
            raise ValueError("Convection scheme not recognized")
        self._cache = {}

    # -- descriptor ------------------------------------------------------------
    def _per_traj(self, name, v, b, dtype, device):
        """float -> (value, None); tensor with 1 or b entries -> (0, [b] tensor)."""
        if not isinstance(v, torch.Tensor):
            return float(v), None
        # the cache holds v itself, so the identity test below cannot be
        # fooled by a new tensor at a recycled address or id; in-place edits
        # bump _version
        key = (v._version, dtype, b)
        hit = self._cache.get(name)
        if hit is not None and hit[0] is v and hit[1] == key:
            return 0.0, hit[2]
        flat = v.detach().reshape(-1).to(device=device, dtype=dtype)
        if flat.numel() == 1:
            flat = flat.expand(b)
        if flat.numel() != b:
            raise ValueError(f"{name} must be a scalar or have one value per trajectory ([batch_size,1])")
        flat = flat.contiguous()
        self._cache[name] = (v, key, flat)
        return 0.0, flat

    def _forcing(self, b, dtype, device):
        p = self.param
        srcs = (p.A, p.omega, p.phi, p.l)
        # held by the cache entry (identity, not id(): no recycled ids) with
        # their versions (in-place edits of the forcing are seen)
        key = (tuple(getattr(v, "_version", None) for v in srcs), dtype, b)
        hit = self._cache.get("forcing")
        if hit is not None and all(a is c for a, c in zip(hit[0], srcs)) and hit[1] == key:
            return hit[2]
        J = p.J if not callable(getattr(p, "J", None)) else p.J()
        arrs = []
        for v in (p.A, p.omega, p.phi, p.l):
            a = torch.as_tensor(v, device=device).to(dtype).reshape(b, -1)[:, :J].contiguous()
            arrs.append(a)
        self._cache["forcing"] = (srcs, key, (J, arrs))
        return J, arrs

    def _desc(self, u, dt, t, x=None):
        if u.dim() != 2:
            raise ValueError(f"state must be [batch, resolution] (got {tuple(u.shape)})")
        b, n = u.shape
        if n != self.resolution:
            # dx and the diffusion stencil belong to the solver's grid (the
            # reference's dense D2 product fails on a mismatched state)
            raise ValueError(f"state has {n} points, solver resolution is {self.resolution}")
        dtype, dev = u.dtype, u.device
        d = L.BurgersDesc()
        d.dtype = L.dtype_code(dtype)
        d.scheme = L.INC_SSPRK3 if self.time_scheme == "SSPRK3" else L.INC_EULER
        d.convection = self._conv
        kind = self.param.equation_type
        if kind not in _SOURCES:
            raise ValueError("Unknown equation type.")
        d.source = _SOURCES[kind]
        d.batch, d.n, d.team = b, n, self.team
        d.dt, d.dx, d.L = float(dt), float(self.dx), float(self.L)
        keep = []
        d.nu, nu_t = self._per_traj("nu", self.nu, b, dtype, dev)
        if nu_t is not None:
            keep.append(nu_t)
            d.nu_ptr = nu_t.data_ptr()
        if d.source == L.INC_SRC_LINEAR:
            d.beta, be_t = self._per_traj("beta", self.param.beta, b, dtype, dev)
            if be_t is not None:
                keep.append(be_t)
                d.beta_ptr = be_t.data_ptr()
        if isinstance(t, torch.Tensor):
            if tuple(t.shape) not in ((b, 1), (b,)):
                raise ValueError("t tensor must have shape [batch_size,1]")
            tt = t.detach().reshape(-1).to(device=dev, dtype=dtype).contiguous()
            keep.append(tt)
            d.t_ptr = tt.data_ptr()
        else:
            d.t0 = float(t)
        if d.source == L.INC_SRC_LEARNING:
            xg = self.x if x is None else x
            xg = torch.as_tensor(xg, device=dev).to(dtype).reshape(-1).contiguous()
            if xg.numel() != n:
                raise ValueError("x must have one coordinate per grid point")
            J, arrs = self._forcing(b, dtype, dev)
            keep += [xg] + arrs
            d.J = J
            d.x = xg.data_ptr()
            d.A, d.omega, d.phi, d.l = (a.data_ptr() for a in arrs)
        return d, keep

    # -- reference API -----------------------------------------------------------
    def initial_conditions(self, x=None, t=None):
        """Compute the forcing field from the parameters (solver_1d.py:68-72)."""
        return self._compute_delta(self.x if x is None else x, self.t if t is None else t)

    def _compute_delta(self, x, t):
        p = self.param
        x = torch.as_tensor(x)
        dev = x.device if x.is_cuda else torch.device(self.device)
        dtype = x.dtype if x.dtype in (torch.float32, torch.float64) else torch.float32
        if isinstance(t, torch.Tensor) and tuple(t.shape) != (p.batch_size, 1):
            raise ValueError("t tensor must have shape [batch_size,1]")
        z = torch.zeros(p.batch_size, x.numel(), device=dev, dtype=dtype)
        L.require_gpu(z)
        saved = p.equation_type
        try:
            p.equation_type = "learning"
            desc, keep = self._desc(z, self.dt, t, x=x.to(dev))
        finally:
            p.equation_type = saved
        out = torch.empty_like(z)
        L.check(L.load().inc_burgers_rhs(ctypes.byref(desc), L.INC_PART_SOURCE, L.ptr(z), None, L.ptr(out),
                                         L.stream()), "_compute_delta")
        return out

    def _compute_source(self, u):
        kind = self.param.equation_type
        if kind in ("homogeneous", "viscosity"):
            return 0.0
        if kind not in _SOURCES:
            raise ValueError("Unknown equation type.")
        L.require_gpu(u)
        return _BurgersRHS.apply(u, self, L.INC_PART_SOURCE, self.t, _nu_input(self))

    def _apply_corrections(self, source):
        return source + self.correction_s if self.correction_s is not None else source

    def calc_nonlinear_weno(self, u):
        L.require_gpu(u)
        if self._conv != L.INC_WENO:
            raise ValueError("solver was built with convection_term='FDM'")
        return _BurgersRHS.apply(u, self, L.INC_PART_CONVECTION, self.t, _nu_input(self))

    def calc_nonlinear_FDM(self, u):
        L.require_gpu(u)
        if self._conv != L.INC_FDM:
            raise ValueError("solver was built with convection_term='WENO'")
        return _BurgersRHS.apply(u, self, L.INC_PART_CONVECTION, self.t, _nu_input(self))

    def calc_linear(self, u):
        L.require_gpu(u)
        return _BurgersRHS.apply(u, self, L.INC_PART_DIFFUSION, self.t, _nu_input(self))

    def _advance(self, u, dt, steps=1, want_traj=False, checkpoint_every=None):
        L.require_gpu(u)
        if self._team_auto:
            self._team_auto = False
            self.team = autotune.tune_burgers(u.shape[-1], u.shape[0], u.dtype, self.time_scheme,
                                              self.param.convection_term)
        s = _burgers_correction(self.correction_s, u)
        nu = _nu_input(self)
        need = torch.is_grad_enabled() and (u.requires_grad or (s is not None and s.requires_grad) or nu is not None)
        return _BurgersAdvance.apply(u, s, self, dt, steps, self.t, want_traj, need, nu,
                                     _checkpoint_k(checkpoint_every))

    def step_adaptive(self, u):
        """Sub-step to CFL * dx / max|u| over one dt (solver_1d.py:148-179):
        a whole-batch max with a host sync per sub-step, capped at 1000."""
        target = self.dt
        n_sub = 0
        max_u = None
        while target > 0:
            max_u = torch.max(torch.abs(u))
            mu = max_u.item()
            allowed = self.adaptive_CFL * self.dx / mu if mu > 0 else target
            dt_loc = min(target, allowed)
            u = self._advance(u, dt_loc)
            self.t += dt_loc
            target -= dt_loc
            if int(math.ceil(target / dt_loc)) > 1000:
                if self.LOG is not None:
                    self.LOG.warning(f"Substeps {int(math.ceil(target / dt_loc))} exceeded max substeps (1000)")
                break
            n_sub += 1
        if self.LOG is not None:
            self.LOG.info(f"CFL: {self.adaptive_CFL}, using substeps: {n_sub}, time_step: {self.dt}, max_u: {max_u}")
        return u

    def _advance_adaptive(self, u, steps=1, want_traj=False):
        """Per-trajectory adaptive sub-stepping on the GPU (adaptive_mode
        'trajectory'): no host synchronisation; sub-step counts land in
        ``self.last_substeps``."""
        L.require_gpu(u)
        s = _burgers_correction(self.correction_s, u)
        nu = _nu_input(self)
        need = torch.is_grad_enabled() and (u.requires_grad or (s is not None and s.requires_grad) or nu is not None)
        return _BurgersAdaptive.apply(u, s, self, steps, self.t, want_traj, need, nu)

    def step(self, u):
        """Advance by one dt (solver_1d.py:181-195)."""
        if self.adaptive_CFL is not None:
            if self.adaptive_mode == "trajectory":
                u_new = self._advance_adaptive(u)
                self.t += self.dt
                return u_new
            return self.step_adaptive(u)
        u_new = self._advance(u, self.dt)
        self.t += self.dt
        if self.LOG is not None:
            self.LOG.info(f"CFL = {torch.max(torch.abs(u)) * self.dt / self.dx:.2f}, time_step: {self.dt}")
        return u_new

    # Largest nu dt / dx^2 with a stable step on the 3-point Laplacian: Euler's
    # 1 + z needs z >= -2, SSP-RK3's 1 + z + z^2/2 + z^3/6 needs z >= -2.5127,
    # with z = -4 nu dt / dx^2 at the highest mode.
    VISCOUS_LIMIT = {"Euler": 0.5, "SSPRK3": 2.5127 / 4}

    def stability_numbers(self, u, dt=None):
        """Per-trajectory (Courant, viscous) numbers of a step of size dt (the
        solver's by default) from u: max_i |u_i| dt / dx and nu dt / dx^2,
        [batch, 2] on the GPU with no host synchronisation (inc_burgers_cfl).
        A NaN in a trajectory makes its Courant number NaN.  The reference
        reads the batch-wide max |u| back to the host to size its adaptive
        sub-steps (solver_1d.py:158-161)."""
        L.require_gpu(u)
        u = u.detach().contiguous()
        d, keep = self._desc(u, self.dt if dt is None else dt, 0.0)
        out = u.new_empty(u.shape[0], 2)
        rc = L.load().inc_burgers_cfl(ctypes.byref(d), L.ptr(u), L.ptr(out), L.stream())
        L.check(rc, "Burgers stability numbers")
        return out

    def unstable(self, u, courant=1.0, viscous=None, dt=None):
        """Bool [batch]: the trajectories whose step from u leaves the stable
        region: Courant number above ``courant``, viscous number above
        ``viscous`` (by default the time scheme's limit, VISCOUS_LIMIT), or a
        NaN in the state.  On the GPU, no host synchronisation."""
        c = self.stability_numbers(u, dt)
        vl = self.VISCOUS_LIMIT[self.time_scheme] if viscous is None else viscous
        return ~((c[:, 0] <= courant) & (c[:, 1] <= vl))

    def rollout(self, u, mstep, checkpoint_every=None):
        """``mstep`` steps in one launch with the correction held fixed;
        returns the [B, mstep, N] trajectory (the fused path of
        simulate_rollout).  With adaptive_CFL this needs adaptive_mode
        'trajectory' (the batch mode reads max|u| back every sub-step).
        ``checkpoint_every=k`` (fixed time step): the forward keeps no stage
        states and the backward recomputes them k steps at a time from the
        trajectory, so the adjoint's memory is O(k) steps of stages instead
        of O(mstep)."""
        if self.adaptive_CFL is not None:
            if self.adaptive_mode != "trajectory":
                raise ValueError("fused rollout with adaptive_CFL needs adaptive_mode='trajectory'")
            if checkpoint_every:
                raise ValueError("checkpoint_every needs a fixed time step; adaptive_CFL is set")
            traj = self._advance_adaptive(u, steps=mstep, want_traj=True)
        else:
            traj = self._advance(u, self.dt, steps=mstep, want_traj=True, checkpoint_every=checkpoint_every)
        for _ in range(mstep):
            self.t += self.dt
        return traj

    def rollout_inc(self, model, u, mstep):
        """simulate_rollout(model, self, u, mstep, 'INC') (trainer_1d.py:154-183)
        with the correction network evaluated inside the rollout kernel
        (inc_burgers_inc_fwd): returns (results, corrections), both
        [B, mstep, N].  ``model`` must be a CorrectionCNN whose shape the
        kernel runs (CorrectionCNN.fused_desc).  With a gradient wanted (u or
        the network's parameters require it) the rollout is differentiable:
        the adjoint kernel runs the network's transpose
        (inc_burgers_inc_train_fwd / _adj, shapes per inc_train_supported)."""
        L.require_like(u, *model.parameters(), what="correction network parameter")
        if self.adaptive_CFL is not None:
            raise ValueError("rollout_inc runs a fixed time step; adaptive_CFL is set")
        if _nu_input(self) is not None:
            raise ValueError("the in-kernel INC adjoint does not differentiate nu; use simulate_rollout's "
                             "per-step path for gradients with respect to nu")
        cnn = model.fused_desc(u.dtype)
        if cnn is None:
            raise ValueError("the correction network's shape is not one the in-kernel path runs")
        params = [t for cv in model.convs for t in (cv.weight, cv.bias)]
        if torch.is_grad_enabled() and (u.requires_grad or any(p.requires_grad for p in params)):
            if not inc_train_supported(model, u.shape[-1], u.dtype):
                raise ValueError("the correction network's shape is not one the in-kernel training adjoint runs")
            traj, corr = _BurgersIncRollout.apply(u, self, model, mstep, self.t, *params)
            for _ in range(mstep):
                self.t += self.dt
            self.correction_s = corr[:, -1] if mstep > 0 else None
            return traj, corr
        cdesc, keep_w = cnn
        u = u.detach().contiguous()
        b, n = u.shape
        desc, keep = self._desc(u, self.dt, self.t)
        out = torch.empty_like(u)
        traj = u.new_empty(b, mstep, n)
        corr = u.new_empty(b, mstep, n)
        rc = L.load().inc_burgers_inc_fwd(ctypes.byref(desc), ctypes.byref(cdesc), L.ptr(u), L.ptr(out), mstep,
                                          L.ptr(traj), L.ptr(corr), L.stream())
        L.check(rc, "BurgersSolverTorch.rollout_inc")
        for _ in range(mstep):
            self.t += self.dt
        self.correction_s = corr[:, -1] if mstep > 0 else None
        return traj, corr

    def rollout_inc_mse(self, model, u, target):
        """simulate_rollout(model, self, u, steps, 'INC') followed by the
        trainer's MSELoss(results, target) (trainer_1d.py:154-183, :716-723)
        with the network inside the rollout and adjoint kernels and the loss
        fused into them: returns (loss, results, corrections).  Shapes as for
        the training path of rollout_inc (inc_train_supported)."""
        L.require_like(u, *model.parameters(), what="correction network parameter")
        L.require_gpu(target)
        if self.adaptive_CFL is not None:
            raise ValueError("rollout_inc_mse runs a fixed time step; adaptive_CFL is set")
        if _nu_input(self) is not None:
            raise ValueError("the in-kernel INC adjoint does not differentiate nu; use simulate_rollout's "
                             "per-step path for gradients with respect to nu")
        if model.fused_desc(u.dtype) is None or not inc_train_supported(model, u.shape[-1], u.dtype):
            raise ValueError("the correction network's shape is not one the in-kernel training adjoint runs")
        steps = target.shape[1]
        params = [t for cv in model.convs for t in (cv.weight, cv.bias)]
        loss, traj, corr = _BurgersIncRolloutMSE.apply(u, target, self, model, steps, self.t, *params)
        for _ in range(steps):
            self.t += self.dt
        self.correction_s = corr[:, -1] if steps > 0 else None
        return loss, traj, corr

    def rollout_mse(self, u, target):
        """Fused rollout over ``target.shape[1]`` steps with the MSE to
        ``target`` [B, steps, N] computed in the kernels: returns (loss,
        trajectory), the loss being torch.nn.MSELoss()(trajectory, target)
        (trainer_1d.py:723 after simulate_rollout).  The backward seeds the
        adjoint with 2 (u - target) / M inside the kernel; no cotangent
        tensor is formed.  Fixed time step only."""
        if self.adaptive_CFL is not None:
            raise ValueError("rollout_mse runs a fixed time step; adaptive_CFL is set")
        L.require_gpu(u, target)
        if self._team_auto:
            self._team_auto = False
            self.team = autotune.tune_burgers(u.shape[-1], u.shape[0], u.dtype, self.time_scheme,
                                              self.param.convection_term)
        steps = target.shape[1]
        s = _burgers_correction(self.correction_s, u)
        nu = _nu_input(self)
        need = torch.is_grad_enabled() and (u.requires_grad or (s is not None and s.requires_grad) or nu is not None)
        loss, traj = _BurgersRolloutMSE.apply(u, s, target, self, self.dt, steps, self.t, need, nu)
        for _ in range(steps):
            self.t += self.dt
        return loss, traj


# ---------------------------------------------------------------------------
# Kuramoto-Sivashinsky
# ---------------------------------------------------------------------------
_KS_SCHEMES = {"ETD1": L.INC_ETD1, "ETRK2": L.INC_ETRK2, "ETDRK4": L.INC_ETDRK4}
_KS_NS = {L.INC_ETD1: 1, L.INC_ETRK2: 2, L.INC_ETDRK4: 4}


class _KSAdvance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, s, solver, scheme, substeps, steps, domain_size, want_traj, need_stages, ckpt):
        u = u.contiguous()
        s = s.contiguous() if s is not None else None
        b, n = u.shape
        desc, keep = solver._desc(u, scheme, substeps, domain_size)
        out = torch.empty_like(u)
        traj = u.new_empty(b, steps, n) if want_traj else None
        nst = _KS_NS[scheme] * substeps
        # checkpoint every ckpt steps: no stage states now, recomputed per segment in backward
        ctx.ckpt = ckpt if (ckpt and need_stages and want_traj and ckpt < steps) else 0
        stages = u.new_empty(b, steps, nst, n) if need_stages and not ctx.ckpt else None
        rc = L.load().inc_ks_fwd(ctypes.byref(desc), L.ptr(u), L.ptr(s), L.ptr(out), steps, L.ptr(traj),
                                 L.ptr(stages), L.stream())
        L.check(rc, "KSSolverTorch.step")
        if ctx.ckpt:
            ctx.save_for_backward(u, s, traj)
            ctx.nst = nst
        else:
            ctx.save_for_backward(stages)
        ctx.desc, ctx.keep, ctx.steps, ctx.want_traj, ctx.shape = desc, keep, steps, want_traj, (b, n)
        return traj if want_traj else out

    @staticmethod
    def _checkpointed_backward(ctx, g):
        u0, s, traj = ctx.saved_tensors
        lib, d = L.load(), ctypes.byref(ctx.desc)

        def fwd(k0, start, out, m, stages):
            rc = lib.inc_ks_fwd(d, L.ptr(start), L.ptr(s), L.ptr(out), m, None, L.ptr(stages), L.stream())
            L.check(rc, "KSSolverTorch rollout (checkpoint recompute)")

        def adj(k0, stages, g_seg, g_u, g_s, g_nu, m):
            rc = lib.inc_ks_adj(d, L.ptr(stages), L.ptr(g_seg), L.ptr(g_u), L.ptr(g_s), m, L.stream())
            L.check(rc, "KSSolverTorch rollout backward (checkpointed)")

        g_u, g_s, _ = _checkpointed_adjoint(g.contiguous(), u0, traj, ctx.ckpt, ctx.nst, fwd, adj,
                                            ctx.needs_input_grad[1])
        return g_u, g_s, None, None, None, None, None, None, None, None

    @staticmethod
    def backward(ctx, g):
        if ctx.ckpt:
            return _KSAdvance._checkpointed_backward(ctx, g)
        (stages,) = ctx.saved_tensors
        if stages is None:
            raise RuntimeError("KS step was run without saved stages; enable grad before the step")
        b, n = ctx.shape
        if ctx.want_traj:
            g_traj = g.contiguous()
        elif ctx.steps == 1:
            g_traj = g.contiguous().view(b, 1, n)
        else:
            g_traj = g.new_zeros(b, ctx.steps, n)
            g_traj[:, -1] = g
        g_u = torch.empty(b, n, dtype=g.dtype, device=g.device)
        g_s = torch.empty_like(g_u) if ctx.needs_input_grad[1] else None
        rc = L.load().inc_ks_adj(ctypes.byref(ctx.desc), L.ptr(stages), L.ptr(g_traj), L.ptr(g_u), L.ptr(g_s),
                                 ctx.steps, L.stream())
        L.check(rc, "KSSolverTorch.step backward")
        return g_u, g_s, None, None, None, None, None, None, None, None


class _KSIncRollout(torch.autograd.Function):
    """simulate_rollout(model, solver, u, steps, 'INC', L) (trainer_1d.py:154-183)
    for a KS solver and a CorrectionCNN in one library call per direction
    (inc_ks_inc_fwd / _adj): per step the network as one convolution launch
    per layer over the batch and one ETD step with its output as the
    correction.  With a gradient wanted the forward keeps every step's
    hidden activations for the backward as ``keep_acts`` says
    (CorrectionCNN.activation_store); otherwise the backward evaluates the
    network again at every step's input.  Returns (results, corrections),
    both [B, steps, N], differentiable in u and in the network's parameters
    (``params`` in the order [w0, b0, w1, b1, ...])."""

    @staticmethod
    def forward(ctx, u, solver, model, scheme, substeps, steps, domain_size, need_grad, keep_acts, *params):
        ctx.set_materialize_grads(False)
        u = u.contiguous()
        b, n = u.shape
        desc, keep = solver._desc(u, scheme, substeps, domain_size)
        cdesc, keep_w = model.ragged_desc(u.dtype)
        lib = L.load()
        elems = ctypes.c_int64(0)
        L.check(lib.inc_ks_inc_workspace(ctypes.byref(desc), ctypes.byref(cdesc), 0, ctypes.byref(elems)),
                "KSSolverTorch.rollout_inc workspace")
        work = u.new_empty(elems.value)
        traj = u.new_empty(steps, b, n)
        corr = u.new_empty(steps, b, n)
        nst = _KS_NS[scheme] * substeps
        stages = u.new_empty(steps, b, nst, n) if need_grad else None
        acts = model.activation_store(steps, b * n, u, keep_acts) if need_grad else None
        rc = lib.inc_ks_inc_fwd(ctypes.byref(desc), ctypes.byref(cdesc), L.ptr(u), steps, L.ptr(traj), L.ptr(corr),
                                L.ptr(stages), L.ptr(acts), L.ptr(work), work.numel(), L.stream())
        L.check(rc, "KSSolverTorch.rollout_inc")
        if need_grad:
            # the parameters too: an in-place update before the backward is then an autograd error
            ctx.save_for_backward(u, traj, stages, acts, *params)
        ctx.desc, ctx.keep, ctx.cdesc, ctx.keep_w = desc, keep, cdesc, keep_w
        ctx.steps, ctx.param_shapes = steps, [p.shape for p in params]
        return traj.transpose(0, 1), corr.transpose(0, 1)

    @staticmethod
    def backward(ctx, g_traj, g_corr):
        u0, traj, stages, acts = ctx.saved_tensors[:4]
        g_traj = g_traj.transpose(0, 1).contiguous() if g_traj is not None else torch.zeros_like(traj)
        g_corr = g_corr.transpose(0, 1).contiguous() if g_corr is not None else None
        lib = L.load()
        elems = ctypes.c_int64(0)
        kind = 1 if acts is None else 2
        L.check(lib.inc_ks_inc_workspace(ctypes.byref(ctx.desc), ctypes.byref(ctx.cdesc), kind, ctypes.byref(elems)),
                "KSSolverTorch.rollout_inc workspace")
        work = u0.new_empty(elems.value)
        g_u0 = torch.empty_like(u0)
        sizes = [math.prod(sh) for sh in ctx.param_shapes]
        g_w = u0.new_empty(sum(sizes))
        rc = lib.inc_ks_inc_adj(ctypes.byref(ctx.desc), ctypes.byref(ctx.cdesc), L.ptr(u0), L.ptr(traj),
                                L.ptr(stages), L.ptr(acts), L.ptr(g_traj), L.ptr(g_corr), L.ptr(g_u0), L.ptr(g_w), ctx.steps,
                                L.ptr(work), work.numel(), L.stream())
        L.check(rc, "KSSolverTorch.rollout_inc backward")
        grads = [g.view(sh) for g, sh in zip(g_w.split(sizes), ctx.param_shapes)]
        return (g_u0, None, None, None, None, None, None, None, None, *grads)


class _KSRolloutMSE(torch.autograd.Function):
    """Fused KS rollout with the MSE to a target trajectory
    (inc_ks_fwd_mse / _adj_mse); returns (loss, [B, steps, N] trajectory)."""

    @staticmethod
    def forward(ctx, u, s, target, solver, scheme, substeps, steps, domain_size, need_stages):
        ctx.set_materialize_grads(False)
        u = u.contiguous()
        s = s.contiguous() if s is not None else None
        b, n = u.shape
        tgt = _mse_target(target, u, steps)
        desc, keep = solver._desc(u, scheme, substeps, domain_size)
        out = torch.empty_like(u)
        traj = u.new_empty(b, steps, n)
        nst = _KS_NS[scheme] * substeps
        stages = u.new_empty(b, steps, nst, n) if need_stages else None
        parts = u.new_empty(b, 4)
        rc = L.load().inc_ks_fwd_mse(ctypes.byref(desc), L.ptr(u), L.ptr(s), L.ptr(out), steps, L.ptr(traj),
                                     L.ptr(stages), L.ptr(tgt), L.ptr(parts), L.stream())
        L.check(rc, "KSSolverTorch.rollout_mse")
        ctx.save_for_backward(stages, traj, tgt)
        ctx.desc, ctx.keep, ctx.steps, ctx.shape = desc, keep, steps, (b, n)
        return parts.sum() / traj.numel(), traj

    @staticmethod
    def backward(ctx, g_loss, g_traj):
        stages, traj, tgt = ctx.saved_tensors
        if stages is None:
            raise RuntimeError("KS rollout was run without saved stages; enable grad before the rollout")
        if g_loss is None:
            g_loss = traj.new_zeros(())
        scale, g_traj, g_target = _mse_backward_common(ctx, g_loss, g_traj, traj, tgt)
        b, n = ctx.shape
        g_u = torch.empty(b, n, dtype=traj.dtype, device=traj.device)
        g_s = torch.empty_like(g_u) if ctx.needs_input_grad[1] else None
        rc = L.load().inc_ks_adj_mse(ctypes.byref(ctx.desc), L.ptr(stages), L.ptr(traj), L.ptr(tgt), L.ptr(scale),
                                     L.ptr(g_traj), L.ptr(g_u), L.ptr(g_s), ctx.steps, L.stream())
        L.check(rc, "KSSolverTorch.rollout_mse backward")
        return g_u, g_s, g_target, None, None, None, None, None, None


class KSSolverTorch:
    """Kuramoto-Sivashinsky u_t + u u_x + u_xx + u_xxxx = s on [0, L)
    (solver_1d.py:198-288), pseudo-spectral exponential integrators.

    ``time_scheme`` 'ETD1' / 'ETRK2' reproduce the reference (including the
    double step, ``doublestep``); 'ETDRK4' adds Cox-Matthews ETDRK4 with
    contour-integral coefficients and, when ``sim_params.dealias`` (default
    True), the 2/3-rule mask on the nonlinear term.  ``resolution`` is any
    size in [4, 1024]: powers of two from 16 run the register-transpose FFT,
    the others a mixed-radix transform through LDS.
    """

    def __init__(self, sim_params):
        self.resolution = sim_params.resolution
        self.dt = sim_params.dt
        self.doublestep = True
        self.device = sim_params.device
        self.domain_size = sim_params.domain_size
        self.alpha = sim_params.alpha
        self.correction_s = None
        self.correction_u = None
        self.dealias = bool(getattr(sim_params, "dealias", True))
        # two real trajectories per complex FFT (default) or one per FFT; the
        # unpacked form keeps a trajectory finite when its FFT partner is not
        self.pack_pairs = bool(getattr(sim_params, "pack_pairs", True))
        # rollout_inc: keep every step's hidden activations for the adjoint
        # (True, False or "auto", CorrectionCNN.activation_store)
        self.keep_activations = getattr(sim_params, "keep_activations", "auto")
        radix = getattr(sim_params, "radix", 0) or 0
        self._radix_auto = radix == "auto" or (radix == 0 and autotune.enabled())
        self.radix = 0 if radix == "auto" else int(radix)
        self.fft_indices = torch.fft.fftfreq(self.resolution, d=1.0, device=self.device) * self.resolution
        time_scheme_map = {"ETRK2": self.etrk2, "ETD1": self.etd1, "ETDRK4": self.etdrk4}
        self.step = time_scheme_map[sim_params.time_scheme]
        self.time_scheme = sim_params.time_scheme
        self._tab = {}

    def initial_conditions(self):
        """cos(2 pi x) + alpha sin(2 pi x) + eps/||eps|| (solver_1d.py:214-223)."""
        domain_size = self.domain_size.unsqueeze(1)
        alpha = self.alpha.unsqueeze(1)
        if len(alpha) != len(domain_size):
            raise ValueError("Length of alpha and domain_size must be the same")
        x = torch.linspace(0, 1, self.resolution, device=self.device)
        u = torch.cos(2 * torch.pi * x) + alpha * torch.sin(2 * torch.pi * x)
        eps = torch.randn_like(u)
        return u + eps / torch.linalg.norm(eps)

    def calc_linear(self, domain_size):
        """(Lop, ik) spectra on [B, N] (solver_1d.py:235-240); a helper of the
        reference API, not used by the HIP step."""
        domain_size = torch.as_tensor(domain_size, device=self.device).unsqueeze(-1)
        wavenumbers = (2j * torch.pi * self.fft_indices) / domain_size
        return -wavenumbers ** 2 - wavenumbers ** 4, wavenumbers

    def calc_nonlinear(self, u, wavenumbers):
        """-ik/2 * fft(u^2) (solver_1d.py:225-233); a helper of the reference
        API, not used by the HIP step."""
        return -0.5 * wavenumbers * torch.fft.fft(u ** 2)

    # -- tables and descriptor -----------------------------------------------------
    def _tables(self, domain_size, b, dtype, dev, scheme, dt):
        conf = (dtype, scheme, float(dt), self.dealias, self.pack_pairs, b)
        if not isinstance(domain_size, torch.Tensor) or not domain_size.is_cuda:
            # host values: keyed by value (no device synchronisation involved)
            ds = torch.as_tensor(domain_size, dtype=torch.float64)
            key = ("values", tuple(ds.reshape(-1).tolist())) + conf
            hit = self._tab.get(key)
            if hit is not None:
                return hit[1]
        else:
            # a device tensor: keyed by storage, shape and version.  The
            # entry holds the tensor, so the caching allocator cannot hand
            # its block to a new tensor while the entry lives, and an
            # in-place edit bumps _version (shared by every view).
            ds = domain_size
            key = ("device", ds.data_ptr(), ds.storage_offset(), tuple(ds.shape), tuple(ds.stride()), ds.dtype,
                   ds._version) + conf
            hit = self._tab.get(key)
            if hit is not None and hit[0] is not None and hit[0].data_ptr() == ds.data_ptr():
                return hit[1]
        Lf = ds.detach().reshape(-1).to(device=dev, dtype=torch.float64)
        if Lf.numel() not in (1, b):
            raise ValueError("domain_size must have one value or one per trajectory")
        uniq, inv = torch.unique(Lf, return_inverse=True)
        count = int(uniq.numel())
        n = self.resolution
        rows = n // 2 + 1
        table = torch.empty(count * rows * L.INC_KS_TABLE_STRIDE + 2 * n, device=dev, dtype=dtype)
        uniq = uniq.contiguous()
        rc = L.load().inc_ks_tables(L.dtype_code(dtype), scheme, n, float(dt), int(self.dealias), L.ptr(uniq),
                                    count, L.ptr(table), L.stream())
        L.check(rc, "KSSolverTorch coefficient tables")
        if count == 1 and self.pack_pairs:
            idx = None
        else:
            idx = inv.to(torch.int32).expand(b).contiguous() if inv.numel() == 1 else inv.to(torch.int32).contiguous()
        entry = (table, idx, uniq)
        if len(self._tab) > 8:
            self._tab.clear()
        self._tab[key] = (ds if ds.is_cuda else None, entry)
        return entry

    def _desc(self, u, scheme, substeps, domain_size):
        b, n = u.shape
        if n != self.resolution:
            raise ValueError(f"state has {n} points, solver resolution is {self.resolution}")
        table, idx, uniq = self._tables(domain_size, b, u.dtype, u.device, scheme, self.dt)
        d = L.KSDesc()
        d.dtype = L.dtype_code(u.dtype)
        d.scheme = scheme
        d.batch, d.n = b, n
        d.substeps = substeps
        d.dealias = int(self.dealias)
        d.radix = self.radix
        d.dt = float(self.dt)
        d.table = table.data_ptr()
        d.twiddle = table.data_ptr() + (table.numel() - 2 * n) * table.element_size()
        d.table_idx = idx.data_ptr() if idx is not None else None
        return d, [table, idx, uniq]

    def _advance(self, u, domain_size, scheme, substeps, steps=1, want_traj=False, checkpoint_every=None):
        L.require_gpu(u)
        if self._radix_auto:
            self._radix_auto = False
            self.radix = autotune.tune_ks(u.shape[-1], u.shape[0], u.dtype, self.time_scheme, dt=self.dt)
        s = _ks_correction(self.correction_s, u)
        if domain_size is None:
            domain_size = self.domain_size
        need = torch.is_grad_enabled() and (u.requires_grad or (s is not None and s.requires_grad))
        return _KSAdvance.apply(u, s, self, scheme, substeps, steps, domain_size, want_traj, need,
                                _checkpoint_k(checkpoint_every))

    def _run(self, u, domain_size, scheme):
        if self.correction_u is not None:
            u = u + self.correction_u
        return self._advance(u, domain_size, scheme, 2 if self.doublestep else 1)

    # single sub-steps (solver_1d.py:242-278)
    def etd1_step(self, u, domain_size):
        return self._advance(u, domain_size, L.INC_ETD1, 1)

    def etrk2_step(self, u, domain_size):
        return self._advance(u, domain_size, L.INC_ETRK2, 1)

    def etdrk4_step(self, u, domain_size):
        return self._advance(u, domain_size, L.INC_ETDRK4, 1)

    # step() bindings with correction_u and the double step (solver_1d.py:280-288)
    def etd1(self, u, domain_size):
        return self._run(u, domain_size, L.INC_ETD1)

    def etrk2(self, u, domain_size):
        return self._run(u, domain_size, L.INC_ETRK2)

    def etdrk4(self, u, domain_size):
        return self._run(u, domain_size, L.INC_ETDRK4)

    def rollout(self, u, mstep, domain_size=None, checkpoint_every=None):
        """``mstep`` step() calls in one launch with the correction held
        fixed; returns the [B, mstep, N] trajectory.  ``checkpoint_every=k``:
        the forward keeps no stage states and the backward recomputes them
        k steps at a time from the trajectory (adjoint memory O(k) steps of
        stages instead of O(mstep))."""
        if self.correction_u is not None:
            raise ValueError("fused rollout does not apply correction_u; use step()")
        scheme = _KS_SCHEMES[self.time_scheme]
        return self._advance(u, domain_size, scheme, 2 if self.doublestep else 1, steps=mstep, want_traj=True,
                             checkpoint_every=checkpoint_every)

    def rollout_inc(self, model, u, mstep, domain_size=None, keep_activations=None):
        """simulate_rollout(model, self, u, mstep, 'INC', L) (trainer_1d.py:154-183)
        for a CorrectionCNN (any width up to 64, odd kernel size up to 9,
        2..8 layers, parameters on the GPU in the state's dtype) in one
        library call (inc_ks_inc_fwd), and its gradient in one more
        (inc_ks_inc_adj): returns (results, corrections), both [B, mstep,
        N].  The network runs on this library's convolution kernels.  For
        the adjoint the forward keeps the step states and stages, and every
        step's hidden activations as ``keep_activations`` says (True, False,
        or "auto": when they take at most half of the free device memory;
        without them the adjoint evaluates the network again per step);
        ``keep_activations`` given here overrides it for this call."""
        if self.correction_u is not None:
            raise ValueError("rollout_inc does not apply correction_u; use step()")
        L.require_like(u, *model.parameters(), what="correction network parameter")
        if model.ragged_desc(u.dtype) is None:
            raise ValueError("the correction network's shape is not one the library's convolutions run")
        if self._radix_auto:
            self._radix_auto = False
            self.radix = autotune.tune_ks(u.shape[-1], u.shape[0], u.dtype, self.time_scheme, dt=self.dt)
        scheme = _KS_SCHEMES[self.time_scheme]
        substeps = 2 if self.doublestep else 1
        if domain_size is None:
            domain_size = self.domain_size
        params = [t for cv in model.convs for t in (cv.weight, cv.bias)]
        need = torch.is_grad_enabled() and (u.requires_grad or any(p.requires_grad for p in params))
        keep = self.keep_activations if keep_activations is None else keep_activations
        traj, corr = _KSIncRollout.apply(u, self, model, scheme, substeps, mstep, domain_size, need, keep, *params)
        self.correction_s = corr[:, -1] if mstep > 0 else None
        return traj, corr

    def rollout_mse(self, u, target, domain_size=None):
        """Fused rollout over ``target.shape[1]`` step() calls with the MSE
        to ``target`` [B, steps, N] computed in the kernels: returns (loss,
        trajectory) as BurgersSolverTorch.rollout_mse does."""
        if self.correction_u is not None:
            raise ValueError("fused rollout does not apply correction_u; use step()")
        L.require_gpu(u, target)
        scheme = _KS_SCHEMES[self.time_scheme]
        substeps = 2 if self.doublestep else 1
        if domain_size is None:
            domain_size = self.domain_size
        if self._radix_auto:
            self._radix_auto = False
            self.radix = autotune.tune_ks(u.shape[-1], u.shape[0], u.dtype, self.time_scheme, dt=self.dt)
        s = _ks_correction(self.correction_s, u)
        need = torch.is_grad_enabled() and (u.requires_grad or (s is not None and s.requires_grad))
        return _KSRolloutMSE.apply(u, s, target, self, scheme, substeps, target.shape[1], domain_size, need)
