"""Correction network s(u) used by the INC workloads.

A periodic 1D CNN: ``depth`` circular Conv1d layers of ``width`` channels and
kernel ``k`` with GELU in between, He-normal weights and zero biases (the
reference's correction networks take (u, L) and return [B, N],
models/models_1d/resnet.py:22-52; this one keeps that call shape).

The convolutions run in libinc_amd (``inc_conv1d_*``: LDS-tiled direct
convolution with register blocking, deterministic two-phase weight gradient)
by default; ``backend='torch'`` uses torch.nn.Conv1d instead.
"""
import torch
from torch import nn
from torch.nn import functional as F

from . import _lib


class _CircConv1d(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b):
        x = x.contiguous()
        w = w.contiguous()
        bsz, ci, n = x.shape
        co, _, k = w.shape
        y = x.new_empty(bsz, co, n)
        bias = b.contiguous() if b is not None else None
        rc = _lib.load().inc_conv1d_fwd(_lib.dtype_code(x.dtype), bsz, ci, co, n, k, _lib.ptr(x), _lib.ptr(w),
                                        _lib.ptr(bias), _lib.ptr(y), _lib.stream())
        _lib.check(rc, "circular conv1d")
        ctx.save_for_backward(x, w)
        ctx.has_bias = b is not None
        return y

    @staticmethod
    def backward(ctx, g):
        x, w = ctx.saved_tensors
        g = g.contiguous()
        bsz, ci, n = x.shape
        co, _, k = w.shape
        code = _lib.dtype_code(x.dtype)
        gx = gw = gb = None
        if ctx.needs_input_grad[0]:
            gx = torch.empty_like(x)
            rc = _lib.load().inc_conv1d_bwd_data(code, bsz, ci, co, n, k, _lib.ptr(g), _lib.ptr(w), _lib.ptr(gx),
                                                 _lib.stream())
            _lib.check(rc, "circular conv1d input gradient")
        if ctx.needs_input_grad[1] or (ctx.has_bias and ctx.needs_input_grad[2]):
            width = co * ci * k + co
            tiles = bsz * ((n + 127) // 128)
            work = x.new_empty(min(tiles, 512) * width)
            gw = torch.empty_like(w)
            gb = x.new_empty(co) if ctx.has_bias else None
            rc = _lib.load().inc_conv1d_bwd_weight(code, bsz, ci, co, n, k, _lib.ptr(g), _lib.ptr(x), _lib.ptr(gw),
                                                   _lib.ptr(gb), _lib.ptr(work), work.numel(), _lib.stream())
            _lib.check(rc, "circular conv1d weight gradient")
        return gx, gw, gb


def circular_conv1d(x, weight, bias=None):
    """Periodic 'same' convolution of x [B, Ci, N] with weight [Co, Ci, K]
    (torch.nn.Conv1d(padding=K//2, padding_mode='circular') semantics)."""
    _lib.require_gpu(x, weight, bias)
    return _CircConv1d.apply(x, weight, bias)


class CircularConv1d(nn.Module):
    def __init__(self, cin, cout, k, backend="hip"):
        super().__init__()
        self.weight = nn.Parameter(torch.empty(cout, cin, k))
        self.bias = nn.Parameter(torch.zeros(cout))
        self.k, self.backend = k, backend

    def forward(self, x):
        if self.backend == "hip":
            return circular_conv1d(x, self.weight, self.bias)
        p = self.k // 2
        return F.conv1d(F.pad(x, (p, p), mode="circular"), self.weight, self.bias)


class CorrectionCNN(nn.Module):
    def __init__(self, depth=4, width=32, k=5, generator=None, backend="hip"):
        super().__init__()
        layers = []
        cin = 1
        for i in range(depth):
            cout = 1 if i == depth - 1 else width
            conv = CircularConv1d(cin, cout, k, backend)
            with torch.no_grad():
                std = (2.0 / (cin * k)) ** 0.5
                conv.weight.copy_(torch.randn(conv.weight.shape, generator=generator) * std)
            layers.append(conv)
            cin = cout
        self.convs = nn.ModuleList(layers)
        self.act = nn.GELU()

    def forward(self, u, L=None):
        x = u.unsqueeze(1)
        for i, conv in enumerate(self.convs):
            x = conv(x)
            if i < len(self.convs) - 1:
                x = self.act(x)
        return x.squeeze(1)
