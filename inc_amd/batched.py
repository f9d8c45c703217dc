"""Uneven batches: many Burgers trajectories of different resolution,
viscosity, time step and horizon advanced by ONE kernel launch per step.

The reference trains on batches drawn from one dataset resolution and steps
them with one ``BurgersSolverTorch`` per batch (trainer/trainer_1d.py:338-358,
solvers/solver_1d.py:181-195).  ``RaggedBurgers`` runs a heterogeneous set in a
single launch: trajectories are packed end to end, grouped by resolution (so
a correction network can run on each group as a contiguous [count, n] view)
and issued longest-first (n * horizon), so the in-order block dispatcher acts
as the work queue.  A trajectory whose horizon is reached keeps its state for
the remaining steps of the rollout.
"""
import ctypes
import math

import torch

from . import _lib
from . import autotune


RaggedDesc = _lib.RaggedDesc


def stable_dt(n, nu, L=2 * math.pi, cfl=0.4, umax=1.0, viscous=0.3):
    """min(CFL * dx / max|u|, viscous * dx^2 / nu).  The Laplacian's spectral
    radius is 4 nu / dx^2, so the viscous factor puts its eigenvalue at
    z = -4 * viscous: 0.3 gives z = -1.2, inside SSP-RK3's real-axis limit
    -2.5127 with room for the advective eigenvalues.  0.628 (the real-axis
    limit itself) left none: with CFL 0.4 advection on top, 285 of the 2048
    trajectories of the uneven INC batch blew up; 0.4 and below kept all
    bounded by max|u0|."""
    dx = L / n
    return min(cfl * dx / umax, viscous * dx * dx / nu)


class _RaggedAdvance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, s, rb, steps, step0, want_traj, need_stages):
        u = u.contiguous()
        s = s.contiguous() if s is not None else None
        d = rb._desc(step0)
        out = torch.empty_like(u)
        traj = u.new_empty(steps, rb.total) if want_traj else None
        ns = 3 if rb.scheme == _lib.INC_SSPRK3 else 1
        stages = u.new_empty(steps, ns, rb.total) if need_stages else None
        rc = _lib.load().inc_burgers_ragged_fwd(ctypes.byref(d), _lib.ptr(u), _lib.ptr(s), _lib.ptr(out), steps,
                                                _lib.ptr(traj), _lib.ptr(stages), _lib.stream())
        _lib.check(rc, "RaggedBurgers.step")
        ctx.save_for_backward(stages)
        ctx.rb, ctx.steps, ctx.step0, ctx.want_traj = rb, steps, step0, want_traj
        return traj if want_traj else out

    @staticmethod
    def backward(ctx, g):
        (stages,) = ctx.saved_tensors
        if stages is None:
            raise RuntimeError("ragged step was run without saved stages; enable grad before the step")
        rb = ctx.rb
        if ctx.want_traj:
            g_traj = g.contiguous()
        elif ctx.steps == 1:
            g_traj = g.contiguous().view(1, rb.total)
        else:
            g_traj = g.new_zeros(ctx.steps, rb.total)
            g_traj[-1] = g
        d = rb._desc(ctx.step0)
        g_u = torch.empty(rb.total, dtype=g.dtype, device=g.device)
        g_s = torch.empty_like(g_u) if ctx.needs_input_grad[1] else None
        rc = _lib.load().inc_burgers_ragged_adj(ctypes.byref(d), _lib.ptr(stages), _lib.ptr(g_traj), _lib.ptr(g_u),
                                                _lib.ptr(g_s), ctx.steps, _lib.stream())
        _lib.check(rc, "RaggedBurgers.step backward")
        return g_u, g_s, None, None, None, None, None


class _RaggedIncRollout(torch.autograd.Function):
    """The INC rollout of a ragged batch with a CorrectionCNN
    (inc_ragged_inc_fwd / _adj): per step one ragged convolution per network
    layer over the trajectories still inside their horizon and one ragged
    solver step.  With a gradient wanted the forward keeps every step's
    hidden activations for the backward when rb.keep_activations allows
    (CorrectionCNN.activation_store); otherwise the backward evaluates the
    network again at every step's input state.  Returns the [steps, total]
    trajectory, differentiable in u0 and in the network's parameters
    (``params`` in the order [w0, b0, w1, b1, ...])."""

    @staticmethod
    def forward(ctx, u0, rb, model, steps, need_grad, *params):
        ctx.set_materialize_grads(False)
        u0 = u0.contiguous()
        cdesc, keep_w = model.ragged_desc(u0.dtype)
        d = rb._desc(0)
        rows, active = rb._rows(steps)
        lib = _lib.load()
        elems = ctypes.c_int64(0)
        _lib.check(lib.inc_ragged_inc_workspace(ctypes.byref(d), ctypes.byref(cdesc), 0, ctypes.byref(elems)),
                   "RaggedBurgers INC rollout workspace")
        work = u0.new_empty(elems.value)
        traj = u0.new_empty(steps, rb.total)
        ns = 3 if rb.scheme == _lib.INC_SSPRK3 else 1
        stages = u0.new_empty(steps, ns, rb.total) if need_grad else None
        acts = model.activation_store(steps, rb.total, u0, rb.keep_activations) if need_grad else None
        rc = lib.inc_ragged_inc_fwd(ctypes.byref(d), ctypes.byref(cdesc), ctypes.byref(rows), active, _lib.ptr(u0),
                                    steps, _lib.ptr(traj), None, _lib.ptr(stages), _lib.ptr(acts), _lib.ptr(work),
                                    work.numel(), _lib.stream())
        _lib.check(rc, "RaggedBurgers INC rollout")
        if need_grad:
            ctx.save_for_backward(u0, traj, stages, acts)
        ctx.rb, ctx.model, ctx.steps = rb, model, steps
        ctx.param_shapes = [p.shape for p in params]
        ctx.keep = (keep_w, active)
        return traj

    @staticmethod
    def backward(ctx, g_traj):
        u0, traj, stages, acts = ctx.saved_tensors
        rb, steps = ctx.rb, ctx.steps
        if g_traj is None:
            g_traj = traj.new_zeros(steps, rb.total)
        g_traj = g_traj.contiguous()
        cdesc, keep_w = ctx.model.ragged_desc(u0.dtype)
        d = rb._desc(0)
        rows, active = rb._rows(steps)
        lib = _lib.load()
        elems = ctypes.c_int64(0)
        kind = 1 if acts is None else 2
        _lib.check(lib.inc_ragged_inc_workspace(ctypes.byref(d), ctypes.byref(cdesc), kind, ctypes.byref(elems)),
                   "RaggedBurgers INC rollout workspace")
        work = u0.new_empty(elems.value)
        g_u0 = torch.empty_like(u0)
        sizes = [math.prod(sh) for sh in ctx.param_shapes]
        g_w = u0.new_empty(sum(sizes))
        rc = lib.inc_ragged_inc_adj(ctypes.byref(d), ctypes.byref(cdesc), ctypes.byref(rows), active, _lib.ptr(u0),
                                    _lib.ptr(traj), _lib.ptr(stages), _lib.ptr(acts), _lib.ptr(g_traj),
                                    _lib.ptr(g_u0), _lib.ptr(g_w), steps, _lib.ptr(work), work.numel(),
                                    _lib.stream())
        _lib.check(rc, "RaggedBurgers INC rollout backward")
        grads = [g.view(sh) for g, sh in zip(g_w.split(sizes), ctx.param_shapes)]
        return (g_u0, None, None, None, None, *grads)


class RaggedBurgers:
    """A heterogeneous batch of periodic Burgers problems on [0, L).

    ``n``, ``nu``, ``dt``, ``horizon``: one entry per trajectory (user order).
    Packed tensors are [total] in the internal order (grouped by n);
    ``pack``/``unpack`` convert from/to per-trajectory lists.
    """

    def __init__(self, n, nu, dt, horizon=None, L=2 * math.pi, scheme="SSPRK3", convection="WENO",
                 device="cuda", dtype=torch.float32, team=0, keep_activations="auto"):
        n = [int(v) for v in n]
        self.count = len(n)
        if not (len(nu) == len(dt) == self.count) or (horizon is not None and len(horizon) != self.count):
            raise ValueError("n, nu, dt (and horizon) need one entry per trajectory")
        if self.count and min(n) < 8:
            raise ValueError("every trajectory needs at least 8 points")
        self.scheme = {"Euler": _lib.INC_EULER, "SSPRK3": _lib.INC_SSPRK3}[scheme]
        self.convection = {"WENO": _lib.INC_WENO, "FDM": _lib.INC_FDM}[convection]
        self.dtype, self.device, self.L = dtype, torch.device(device), float(L)
        self._team_auto = team == "auto" or (team == 0 and autotune.enabled())
        self.team = 0 if team == "auto" else int(team)
        # the one-call INC rollout keeps every step's hidden activations for
        # its adjoint (True, False or "auto", CorrectionCNN.activation_store)
        self.keep_activations = keep_activations
        # internal order: grouped by n (stable), so each resolution is contiguous
        self.perm = sorted(range(self.count), key=lambda b: n[b])
        self.n = [n[b] for b in self.perm]
        self.horizon = None if horizon is None else [int(horizon[b]) for b in self.perm]
        offs, o = [], 0
        for v in self.n:
            offs.append(o)
            o += v
        self.offsets, self.total = offs, o
        # resolution groups: (n, first trajectory, count, first point)
        self.groups = []
        for i, v in enumerate(self.n):
            if self.groups and self.groups[-1][0] == v:
                g = self.groups[-1]
                self.groups[-1] = (v, g[1], g[2] + 1, g[3])
            else:
                self.groups.append((v, i, 1, offs[i]))
        cost = [self.n[i] * (self.horizon[i] if self.horizon else 1) for i in range(self.count)]
        order = sorted(range(self.count), key=lambda i: -cost[i])
        dev = self.device
        self.n_d = torch.tensor(self.n, dtype=torch.int32, device=dev)
        self.off_d = torch.tensor(offs, dtype=torch.int64, device=dev)
        self.order_d = torch.tensor(order, dtype=torch.int32, device=dev)
        self.horizon_d = (torch.tensor(self.horizon, dtype=torch.int32, device=dev) if self.horizon else None)
        self.nu_d = torch.tensor([float(nu[b]) for b in self.perm], dtype=dtype, device=dev)
        self.dt_d = torch.tensor([float(dt[b]) for b in self.perm], dtype=dtype, device=dev)
        self.dx_d = torch.tensor([self.L / v for v in self.n], dtype=dtype, device=dev)
        self.nmin = min(self.n) if self.n else 8
        self.nmax = max(self.n) if self.n else 8
        # the correction network's work: 128-position tiles (row, first
        # position), trajectories with the longest horizon first, so that the
        # rows still inside their horizon at a step are a prefix of the list
        T = _lib.INC_CONV_ROW_TILE
        hz = self.horizon or [1 << 30] * self.count
        by_h = sorted(range(self.count), key=lambda i: -hz[i])
        tiles, self._tile_h = [], []
        for i in by_h:
            for i0 in range(0, self.n[i], T):
                tiles += [i, i0]
                self._tile_h.append(hz[i])
        self.ntiles = len(self._tile_h)
        self.tiles_d = torch.tensor(tiles or [0, 0], dtype=torch.int32, device=dev)
        self._active = {}

    def _desc(self, step0=0):
        d = RaggedDesc()
        d.dtype = _lib.dtype_code(self.dtype)
        d.scheme, d.convection = self.scheme, self.convection
        d.count, d.total = self.count, self.total
        d.nmin, d.nmax, d.team, d.step0 = self.nmin, self.nmax, self.team, int(step0)
        d.n, d.offset, d.order = self.n_d.data_ptr(), self.off_d.data_ptr(), self.order_d.data_ptr()
        d.horizon = self.horizon_d.data_ptr() if self.horizon_d is not None else None
        d.nu, d.dt, d.dx = self.nu_d.data_ptr(), self.dt_d.data_ptr(), self.dx_d.data_ptr()
        return d

    def _rows(self, steps):
        """(inc_conv_rows, host int32 array [steps] of the tiles still inside
        their horizon at each step)."""
        act = self._active.get(steps)
        if act is None:
            counts = [sum(1 for h in self._tile_h if h > k) for k in range(steps)]
            act = (ctypes.c_int32 * max(steps, 1))(*counts)
            self._active[steps] = act
        r = _lib.ConvRows(count=self.count, njobs=self.ntiles, tile=self.tiles_d.data_ptr(),
                          n=self.n_d.data_ptr(), offset=self.off_d.data_ptr())
        return r, act

    # -- packing --------------------------------------------------------------------
    def pack(self, fields):
        """List of per-trajectory 1D tensors (user order) -> packed [total]."""
        if len(fields) != self.count:
            raise ValueError("need one field per trajectory")
        return torch.cat([fields[b].reshape(-1).to(self.device, self.dtype) for b in self.perm])

    def unpack(self, packed):
        """Packed [..., total] -> list of [..., n_b] views in user order."""
        out = [None] * self.count
        for i, b in enumerate(self.perm):
            out[b] = packed[..., self.offsets[i]:self.offsets[i] + self.n[i]]
        return out

    def group_views(self, packed):
        """[count_k, n_k] views of each resolution group (internal order)."""
        return [packed[o:o + c * v].view(c, v) for (v, _, c, o) in self.groups]

    def valid_mask(self, steps):
        """[steps, total] 1.0 where the trajectory is still inside its horizon."""
        m = torch.ones(steps, self.total, device=self.device, dtype=self.dtype)
        if self.horizon:
            for i, h in enumerate(self.horizon):
                if h < steps:
                    m[h:, self.offsets[i]:self.offsets[i] + self.n[i]] = 0
        return m

    def stability_numbers(self, u):
        """Per-trajectory (Courant, viscous) numbers of each trajectory's step
        from the packed state u: max |u| dt_b / dx_b and nu_b dt_b / dx_b^2,
        [count, 2] in the user's order, on the GPU (inc_burgers_ragged_cfl).
        A NaN in a trajectory makes its Courant number NaN."""
        _lib.require_gpu(u)
        if u.numel() != self.total or u.dtype != self.dtype:
            raise ValueError(f"packed state must be [{self.total}] of {self.dtype}")
        u = u.detach().contiguous()
        d = self._desc()
        out = u.new_empty(self.count, 2)
        rc = _lib.load().inc_burgers_ragged_cfl(ctypes.byref(d), _lib.ptr(u), _lib.ptr(out), _lib.stream())
        _lib.check(rc, "ragged stability numbers")
        res = torch.empty_like(out)
        res[torch.tensor(self.perm, device=out.device)] = out
        return res

    def unstable(self, u, courant=1.0, viscous=None):
        """Bool [count] (user order): trajectories whose step from u leaves the
        stable region (see BurgersSolverTorch.unstable; the viscous limit is
        the batch's time scheme's)."""
        c = self.stability_numbers(u)
        vl = (0.5 if self.scheme == _lib.INC_EULER else 2.5127 / 4) if viscous is None else viscous
        return ~((c[:, 0] <= courant) & (c[:, 1] <= vl))

    def _fusable(self, model, u0):
        from .correction import CorrectionCNN
        return (isinstance(model, CorrectionCNN) and u0.is_cuda and model.ragged_desc(u0.dtype) is not None
                and 2 * self.nmin >= model.convs[0].k)

    # -- stepping -------------------------------------------------------------------
    def advance(self, u, s=None, steps=1, step0=0, want_traj=False):
        """``steps`` steps from rollout step ``step0`` with a fixed correction."""
        _lib.require_gpu(u, s)
        if u.numel() != self.total:
            raise ValueError(f"packed state has {u.numel()} values, batch has {self.total}")
        if s is not None and s.numel() != self.total:
            raise ValueError(f"packed correction has {s.numel()} values, batch has {self.total}")
        for name, t in (("state", u), ("correction", s)):
            if t is not None and t.dtype != self.dtype:
                raise TypeError(f"packed {name} is {t.dtype}, the ragged batch was built for {self.dtype}")
        if self._team_auto:
            self._team_auto = False
            self.team = autotune.tune_ragged(self)
        need = torch.is_grad_enabled() and (u.requires_grad or (s is not None and s.requires_grad))
        return _RaggedAdvance.apply(u, s, self, steps, step0, want_traj, need)

    def rollout(self, u0, steps, model=None, fused=True):
        """[steps, total] trajectory.  Without ``model`` the whole rollout is
        one launch.  With a CorrectionCNN whose parameters live on the GPU in
        the state's dtype, the INC rollout runs as one call of the library
        (inc_ragged_inc_fwd; its adjoint inc_ragged_inc_adj): per step one
        ragged convolution per layer over every resolution at once, skipping
        trajectories past their horizon, and one ragged solver step.  Any
        other correction network (a callable [count, n] -> [count, n]), or
        ``fused=False``, is evaluated per resolution group before every step."""
        if model is None:
            return self.advance(u0, None, steps=steps, step0=0, want_traj=True)
        if fused and self._fusable(model, u0):
            if u0.numel() != self.total:
                raise ValueError(f"packed state has {u0.numel()} values, batch has {self.total}")
            if u0.dtype != self.dtype:
                raise TypeError(f"packed state is {u0.dtype}, the ragged batch was built for {self.dtype}")
            if self._team_auto:
                self._team_auto = False
                self.team = autotune.tune_ragged(self)
            params = [t for cv in model.convs for t in (cv.weight, cv.bias)]
            need = torch.is_grad_enabled() and (u0.requires_grad or any(p.requires_grad for p in params))
            return _RaggedIncRollout.apply(u0.reshape(-1), self, model, steps, need, *params)
        if torch.cuda.is_current_stream_capturing():
            # the per-group path under hipGraph capture is not validated (a
            # capture of it, then with library convolutions, crashed in
            # hipStreamEndCapture with no cause shown); the one-call path is
            raise RuntimeError("RaggedBurgers.rollout with a per-group correction network cannot be captured; "
                               "use a CorrectionCNN (the one-call path) inside a hipGraph")
        outs, u = [], u0
        for k in range(steps):
            s = torch.cat([model(v).reshape(-1) for v in self.group_views(u)])
            u = self.advance(u, s, steps=1, step0=k)
            outs.append(u)
        return torch.stack(outs, 0)
