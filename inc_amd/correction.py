"""Correction network s(u) used by the INC workloads.

A periodic 1D CNN: ``depth`` circular Conv1d layers of ``width`` channels and
kernel ``k`` with GELU in between, He-normal weights and zero biases (the
reference's correction networks take (u, L) and return [B, N],
models/models_1d/resnet.py:22-52; this one keeps that call shape).

The convolutions run in libinc_amd by default (``inc_conv1d_*_act``:
LDS-tiled direct convolution with register blocking; weight gradients on
the matrix cores in float32, reduced deterministically; the GELU between
layers applied while each convolution stages its input, so post-activations
are never stored); ``backend='torch'`` uses torch.nn.Conv1d and nn.GELU.
"""
import math

import torch
from torch import nn
from torch.nn import functional as F

from . import _lib


def _conv_fwd(x, w, b, act=_lib.INC_ACT_NONE, need_dy=True):
    """(act(conv(x) + b), act'(conv(x) + b) or None); need_dy=False skips the
    derivative (no gradient wanted)."""
    bsz, ci, n = x.shape
    co, _, k = w.shape
    y = x.new_empty(bsz, co, n)
    dy = x.new_empty(bsz, co, n) if act != _lib.INC_ACT_NONE and need_dy else None
    rc = _lib.load().inc_conv1d_fwd_act(_lib.dtype_code(x.dtype), act, bsz, ci, co, n, k, _lib.ptr(x), _lib.ptr(w),
                                        _lib.ptr(b), _lib.ptr(y), _lib.ptr(dy), _lib.stream())
    _lib.check(rc, "circular conv1d")
    return y, dy


def _conv_bwd_data(g, w, ci, m=None):
    """(conv^T g) * m, the gradient w.r.t. the layer input (m: the derivative
    of the activation that formed that input, or None)."""
    bsz, co, n = g.shape
    k = w.shape[-1]
    gx = g.new_empty(bsz, ci, n)
    rc = _lib.load().inc_conv1d_bwd_data_mul(_lib.dtype_code(g.dtype), bsz, ci, co, n, k, _lib.ptr(g), _lib.ptr(w),
                                             _lib.ptr(m), _lib.ptr(gx), _lib.stream())
    _lib.check(rc, "circular conv1d input gradient")
    return gx


def _conv_bwd_weight(g, x, w, has_bias):
    bsz, ci, n = x.shape
    co, _, k = w.shape
    width = co * ci * k + co
    work = x.new_empty(min(bsz * ((n + 127) // 128), 512) * width)
    gw = torch.empty_like(w)
    gb = x.new_empty(co) if has_bias else None
    rc = _lib.load().inc_conv1d_bwd_weight(_lib.dtype_code(x.dtype), bsz, ci, co, n, k, _lib.ptr(g), _lib.ptr(x),
                                           _lib.ptr(gw), _lib.ptr(gb), _lib.ptr(work), work.numel(), _lib.stream())
    _lib.check(rc, "circular conv1d weight gradient")
    return gw, gb


class _CircConv1d(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b):
        x, w = x.contiguous(), w.contiguous()
        b = b.contiguous() if b is not None else None
        y, _ = _conv_fwd(x, w, b)
        ctx.save_for_backward(x, w)
        ctx.has_bias = b is not None
        return y

    @staticmethod
    def backward(ctx, g):
        x, w = ctx.saved_tensors
        g = g.contiguous()
        gx = gw = gb = None
        if ctx.needs_input_grad[0]:
            gx = _conv_bwd_data(g, w, x.shape[1])
        if ctx.needs_input_grad[1] or (ctx.has_bias and ctx.needs_input_grad[2]):
            gw, gb = _conv_bwd_weight(g, x, w, ctx.has_bias)
        return gx, gw, gb


def circular_conv1d(x, weight, bias=None):
    """Periodic 'same' convolution of x [B, Ci, N] with weight [Co, Ci, K]
    (torch.nn.Conv1d(padding=K//2, padding_mode='circular') semantics)."""
    _lib.require_like(x, weight, bias, what="circular_conv1d weight/bias")
    return _CircConv1d.apply(x, weight, bias)


def _cnn_eval(u, ws, bs):
    """The network without gradient: every layer's GELU fused into its
    convolution, no derivative computed or stored."""
    x = u.unsqueeze(1).contiguous()
    depth = len(ws)
    for i in range(depth):
        act = _lib.INC_ACT_GELU if i < depth - 1 else _lib.INC_ACT_NONE
        x, _ = _conv_fwd(x, ws[i].contiguous(), bs[i].contiguous(), act, need_dy=False)
    return x.squeeze(1)


class _CNN(torch.autograd.Function):
    """The whole stack conv -> GELU -> ... -> conv as one autograd node.  Each
    hidden layer's kernel writes gelu(v) (the next layer's input) and gelu'(v);
    backward runs per layer a weight gradient on the saved input and one input
    gradient that multiplies by the saved gelu' in its epilogue."""

    @staticmethod
    def forward(ctx, u, *params):
        ws, bs = params[0::2], params[1::2]
        depth = len(ws)
        x = u.unsqueeze(1).contiguous()
        inputs, dys = [x], []
        for i in range(depth):
            act = _lib.INC_ACT_GELU if i < depth - 1 else _lib.INC_ACT_NONE
            x, dy = _conv_fwd(x, ws[i].contiguous(), bs[i].contiguous(), act)
            if i < depth - 1:
                inputs.append(x)
                dys.append(dy)
        ctx.depth = depth
        ctx.save_for_backward(*inputs, *dys, *ws)
        return x.squeeze(1)

    @staticmethod
    def backward(ctx, gs):
        depth = ctx.depth
        saved = ctx.saved_tensors
        inputs, dys, ws = saved[:depth], saved[depth:2 * depth - 1], saved[2 * depth - 1:]
        g = gs.unsqueeze(1).contiguous()
        grads = [None] * (2 * depth)
        gu = None
        for i in reversed(range(depth)):
            if ctx.needs_input_grad[1 + 2 * i] or ctx.needs_input_grad[2 + 2 * i]:
                grads[2 * i], grads[2 * i + 1] = _conv_bwd_weight(g, inputs[i], ws[i], True)
            if i > 0:
                g = _conv_bwd_data(g, ws[i].contiguous(), inputs[i].shape[1], dys[i - 1])
            elif ctx.needs_input_grad[0]:
                gu = _conv_bwd_data(g, ws[0].contiguous(), 1).squeeze(1)
        return (gu, *grads)


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
    """s = scale * conv_{depth-1}(gelu(... gelu(conv_0(u)) ...)): circular
    Conv1d layers 1 -> width -> ... -> width -> 1 with exact GELU between them,
    He-normal weights and zero biases.  ``scale`` multiplies the output (a
    correction kept small against the dynamics, as in ``0.01 * model(u)``);
    the fused rollout kernels apply it themselves."""

    def __init__(self, depth=4, width=32, k=5, generator=None, backend="hip", scale=1.0):
        super().__init__()
        self.scale = float(scale)
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

    def fused_desc(self, dtype):
        """(inc_cnn_desc, tensors to keep alive) for evaluating this network
        inside a rollout kernel (inc_burgers_inc_fwd), or None when its shape
        is not one the kernel runs (width 32 or 48, kernel size 5 or 7, 2..8
        layers, the HIP backend, every parameter on the GPU in ``dtype``: the
        kernels read the weights through pointers of the state's element
        type, and a mismatched model takes the per-step path, which raises)."""
        convs = list(self.convs)
        c = convs[0].weight.shape[0]
        k = convs[0].k
        if (convs[0].backend != "hip" or not 2 <= len(convs) <= 8 or c not in (32, 48) or k not in (5, 7)
                or any(cv.k != k for cv in convs)):
            return None
        if any(p.dtype != dtype or not p.is_cuda for p in self.parameters()):
            return None
        d = _lib.CnnDesc(layers=len(convs), channels=c, kernel=k, scale=self.scale)
        keep = []
        for i, cv in enumerate(convs):
            w = cv.weight.detach().contiguous()
            b = cv.bias.detach().contiguous()
            keep += [w, b]
            d.weight[i] = w.data_ptr()
            d.bias[i] = b.data_ptr()
        return d, keep

    def ragged_desc(self, dtype):
        """(inc_cnn_desc, tensors to keep alive) for the ragged INC rollout
        (inc_ragged_inc_fwd / _adj: any width up to 64, odd kernel size up to
        9, 2..8 layers, the HIP backend, every parameter on the GPU in
        ``dtype``), or None."""
        convs = list(self.convs)
        c, k = convs[0].weight.shape[0], convs[0].k
        if (convs[0].backend != "hip" or not 2 <= len(convs) <= 8 or not 1 <= c <= 64 or k % 2 != 1 or k > 9
                or any(cv.k != k for cv in convs)):
            return None
        if any(p.dtype != dtype or not p.is_cuda for p in self.parameters()):
            return None
        d = _lib.CnnDesc(layers=len(convs), channels=c, kernel=k, scale=self.scale)
        keep = []
        for i, cv in enumerate(convs):
            w = cv.weight.detach().contiguous()
            b = cv.bias.detach().contiguous()
            keep += [w, b]
            d.weight[i] = w.data_ptr()
            d.bias[i] = b.data_ptr()
        return d, keep

    def activation_store(self, steps, total, like, keep="auto"):
        """The buffer in which a one-call INC rollout over ``total`` state
        values keeps every step's hidden activations and GELU derivatives
        for its adjoint ([steps, layers - 1, 2, channels * total], ``like``'s
        dtype and device; inc_ragged_inc_* / inc_ks_inc_* ``acts``), or None:
        the adjoint then evaluates the network's hidden layers again at every
        step.  ``keep``: True, False, or "auto" (keep when the buffer takes at
        most half of the device's free memory)."""
        if not (keep is True or keep is False or keep == "auto"):
            raise ValueError(f"keep_activations must be True, False or 'auto' (got {keep!r})")
        if keep is False:
            return None
        convs = list(self.convs)
        shape = (steps, len(convs) - 1, 2, convs[0].weight.shape[0] * total)
        if keep == "auto":
            # free device memory plus what torch's caching allocator holds
            # unused (the previous step's buffer, so that every training step
            # decides alike)
            free, _ = torch.cuda.mem_get_info(like.device)
            free += torch.cuda.memory_reserved(like.device) - torch.cuda.memory_allocated(like.device)
            if math.prod(shape) * like.element_size() > free // 2:
                return None
        return like.new_empty(shape)

    def forward(self, u, L=None):
        if self.convs[0].backend == "hip":
            params = [t for conv in self.convs for t in (conv.weight, conv.bias)]
            _lib.require_like(u, *params, what="CorrectionCNN parameter")
            if torch.is_grad_enabled() and (u.requires_grad or any(p.requires_grad for p in params)):
                s = _CNN.apply(u, *params)
            else:
                s = _cnn_eval(u, params[0::2], params[1::2])
        else:
            x = u.unsqueeze(1)
            for i, conv in enumerate(self.convs):
                x = conv(x)
                if i < len(self.convs) - 1:
                    x = self.act(x)
            s = x.squeeze(1)
        return s if self.scale == 1.0 else self.scale * s


class _CircConv1dAct(torch.autograd.Function):
    """act(conv(x) + b) as one autograd node: the convolution's epilogue
    writes act'(v) beside the output (none for act = NONE), and the backward
    multiplies the output's gradient by it before the convolution's input
    and weight gradients."""

    @staticmethod
    def forward(ctx, x, w, b, act):
        x, w = x.contiguous(), w.contiguous()
        b = b.contiguous() if b is not None else None
        y, dy = _conv_fwd(x, w, b, act, need_dy=any(ctx.needs_input_grad[:3]))
        ctx.save_for_backward(x, w, dy)
        ctx.has_bias = b is not None
        return y

    @staticmethod
    def backward(ctx, g):
        x, w, dy = ctx.saved_tensors
        g = g.contiguous()
        if dy is not None:
            g = g * dy
        gx = gw = gb = None
        if ctx.needs_input_grad[0]:
            gx = _conv_bwd_data(g, w, x.shape[1])
        if ctx.needs_input_grad[1] or (ctx.has_bias and ctx.needs_input_grad[2]):
            gw, gb = _conv_bwd_weight(g, x, w, ctx.has_bias)
        return gx, gw, gb, None


def _conv_act(x, conv, act):
    """A circular nn.Conv1d module's weights through this library's kernels
    ('same' periodic padding: padding = K // 2, as the modules are built)."""
    _lib.require_like(x, conv.weight, conv.bias, what="Conv1d weight/bias")
    return _CircConv1dAct.apply(x, conv.weight, conv.bias, act)


def get_grid(x, L):
    """Grid coordinates [B, D] of models/models_1d/utils.py:3-30: linspace(0,
    L, D) for a number L, L_b * linspace(0, 1, D) for a tensor of one L per
    trajectory."""
    b, d = x.shape[:2]
    if isinstance(L, (float, int)):
        return torch.linspace(0, L, d, dtype=x.dtype, device=x.device).expand(b, d)
    if isinstance(L, torch.Tensor):
        return L.to(dtype=x.dtype, device=x.device).reshape(b, 1) * torch.linspace(0, 1, d, dtype=x.dtype,
                                                                                   device=x.device)
    raise ValueError("L must be a float, int, or a torch.Tensor")


class _ResidualBlock1D(nn.Module):
    """relu(conv2(relu(conv1(x))) + shortcut(x)), 5-tap circular convolutions
    (models/models_1d/resnet.py:5-20, the same submodules and parameter
    names)."""

    def __init__(self, in_channels, out_channels):
        super().__init__()
        self.conv1 = nn.Conv1d(in_channels, out_channels, kernel_size=5, padding=2, padding_mode="circular")
        self.conv2 = nn.Conv1d(out_channels, out_channels, kernel_size=5, padding=2, padding_mode="circular")
        self.relu = nn.ReLU(inplace=True)
        self.shortcut = nn.Conv1d(in_channels, out_channels, 1) if in_channels != out_channels else nn.Identity()


class ResNet1D(nn.Module):
    """The reference's ResNet1D correction network (models/models_1d/resnet.py:22-52):
    the state and a grid channel -> circular 5-tap conv + ReLU -> residual
    blocks -> 1x1 conv, returning [B, D].  The module tree, parameter names
    and initialisation (torch's nn.Conv1d defaults, created in the same
    order) are the reference's, so its state dicts load and a seeded
    construction gives the same weights.  With ``backend='hip'`` (default)
    every convolution runs in this library's kernels, ReLU fused into the
    convolution that produces it; ``backend='torch'`` runs the modules
    themselves."""

    def __init__(self, L=16, in_channels=1 + 1, out_channels=1, init_features=32, num_blocks=6, backend="hip"):
        super().__init__()
        self.L = L
        self.backend = backend
        f = init_features
        self.initial = nn.Sequential(nn.Conv1d(in_channels, f, kernel_size=5, padding=2, padding_mode="circular"),
                                     nn.ReLU(inplace=True))
        self.blocks = nn.Sequential(*[_ResidualBlock1D(f, f) for _ in range(num_blocks)])
        self.final = nn.Conv1d(f, out_channels, kernel_size=1)

    def forward(self, x, L=None):
        L = L if L is not None else self.L
        h = torch.stack((x, get_grid(x, L)), 1)  # [B, 2, D]
        if self.backend != "hip":
            h = self.initial(h)
            for blk in self.blocks:
                r = blk.shortcut(h)
                h = blk.relu(blk.conv2(blk.relu(blk.conv1(h))) + r)
            return self.final(h).squeeze(1)
        relu, none = _lib.INC_ACT_RELU, _lib.INC_ACT_NONE
        h = _conv_act(h, self.initial[0], relu)
        for blk in self.blocks:
            r = h if isinstance(blk.shortcut, nn.Identity) else _conv_act(h, blk.shortcut, none)
            h = torch.relu(_conv_act(_conv_act(h, blk.conv1, relu), blk.conv2, none) + r)
        return _conv_act(h, self.final, none).squeeze(1)


def _require_even(n, what):
    if n % 2:
        raise ValueError(f"{what}: the library's 2-to-1 max pooling takes an even number of points (got {n})")


class _MaxPool2(torch.autograd.Function):
    """torch.nn.MaxPool1d(2, 2) on [B, C, n] (n even) in this library's
    kernels; the backward recomputes each pair's pick from the saved input."""

    @staticmethod
    def forward(ctx, x):
        x = x.contiguous()
        b, c, n = x.shape
        _require_even(n, "max pooling")
        y = x.new_empty(b, c, n // 2)
        rc = _lib.load().inc_maxpool2_fwd(_lib.dtype_code(x.dtype), b * c, n, _lib.ptr(x), _lib.ptr(y), _lib.stream())
        _lib.check(rc, "max pooling")
        ctx.save_for_backward(x)
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        gy = gy.contiguous()
        b, c, n = x.shape
        gx = torch.empty_like(x)
        rc = _lib.load().inc_maxpool2_bwd(_lib.dtype_code(x.dtype), b * c, n, _lib.ptr(x), _lib.ptr(gy),
                                          _lib.ptr(gx), _lib.stream())
        _lib.check(rc, "max pooling gradient")
        return gx


def max_pool2(x):
    """torch.nn.MaxPool1d(kernel_size=2, stride=2) of x [B, C, n], n even."""
    _lib.require_gpu(x)
    return _MaxPool2.apply(x)


class _UpsampleCat(torch.autograd.Function):
    """torch.cat((up, skip), 1) where up [B, C, 2n] interleaves the two
    channel halves of y2 [B, 2C, n] (up[:, c, 2i + t] = y2[:, t C + c, i])."""

    @staticmethod
    def forward(ctx, y2, skip):
        y2, skip = y2.contiguous(), skip.contiguous()
        b, c2, n = y2.shape
        c, cs = c2 // 2, skip.shape[1]
        out = y2.new_empty(b, c + cs, 2 * n)
        rc = _lib.load().inc_upsample2_cat(_lib.dtype_code(y2.dtype), b, c, n, cs, _lib.ptr(y2), _lib.ptr(skip),
                                           _lib.ptr(out), _lib.stream())
        _lib.check(rc, "upsampling concatenation")
        ctx.dims = (b, c, n, cs)
        return out

    @staticmethod
    def backward(ctx, g):
        b, c, n, cs = ctx.dims
        g = g.contiguous()
        gy2 = g.new_empty(b, 2 * c, n)
        gskip = g.new_empty(b, cs, 2 * n) if ctx.needs_input_grad[1] else None
        rc = _lib.load().inc_upsample2_cat_bwd(_lib.dtype_code(g.dtype), b, c, n, cs, _lib.ptr(g), _lib.ptr(gy2),
                                               _lib.ptr(gskip), _lib.stream())
        _lib.check(rc, "upsampling concatenation gradient")
        return gy2, gskip


def upconv2_cat(x, up, skip):
    """torch.cat((up(x), skip), 1) for up = nn.ConvTranspose1d(Ci, C, 2, stride=2):
    the transposed convolution as a 1x1 convolution with 2C outputs
    (W'[t C + c] = W[:, c, t], the bias repeated) on the matrix cores, then
    one pass that interleaves its halves and appends skip's channels."""
    ci, c, k = up.weight.shape
    if k != 2 or up.stride != (2,) or up.padding != (0,) or up.output_padding != (0,) or up.dilation != (1,) \
            or up.groups != 1:
        raise ValueError("upconv2_cat takes ConvTranspose1d(kernel_size=2, stride=2) without padding or groups")
    if skip.shape[0] != x.shape[0] or skip.shape[2] != 2 * x.shape[2]:
        raise RuntimeError(f"Sizes of tensors must match except in dimension 1: upsampled {2 * x.shape[2]} points, "
                           f"skip connection {skip.shape[2]}")
    _lib.require_like(x, up.weight, up.bias, skip, what="ConvTranspose1d weight/bias or skip connection")
    w2 = up.weight.permute(2, 1, 0).reshape(2 * c, ci, 1)
    b2 = up.bias.repeat(2) if up.bias is not None else None
    return _UpsampleCat.apply(_CircConv1dAct.apply(x, w2, b2, _lib.INC_ACT_NONE), skip)


def _unet_block(in_channels, features):
    """conv5 -> ReLU -> conv5 -> ReLU, circular padding (unet.py:44-53)."""
    return nn.Sequential(
        nn.Conv1d(in_channels=in_channels, out_channels=features, kernel_size=5, padding=2, padding_mode="circular"),
        nn.ReLU(inplace=True),
        nn.Conv1d(in_channels=features, out_channels=features, kernel_size=5, padding=2, padding_mode="circular"),
        nn.ReLU(inplace=True),
    )


class UNet1D(nn.Module):
    """The reference's UNet1D correction network (models/models_1d/unet.py:5-53),
    the trainer's default model type (scripts/Train_1D.py:82): the state and
    a grid channel -> two encoder blocks with 2-to-1 max pooling -> a
    bottleneck block -> two decoder blocks, each after a kernel-2 stride-2
    transposed convolution concatenated with the encoder output of the same
    resolution -> 1x1 convolution, returning [B, D].  Blocks are two
    circular 5-tap convolutions with ReLU.  The module tree, parameter names
    and initialisation order are the reference's, so its state dicts load
    and a seeded construction gives the same weights.  With
    ``backend='hip'`` (default) every convolution runs in this library's
    kernels with its ReLU fused, and pooling, upsampling and concatenation
    run in the library's resolution-change kernels; ``backend='torch'`` runs
    the modules themselves.  D must be a multiple of 4 (the reference's
    concatenations fail otherwise)."""

    def __init__(self, L=16, in_channels=1 + 1, out_channels=1, init_features=32, backend="hip"):
        super().__init__()
        f = init_features
        self.L = L
        self.backend = backend
        self.encoder1 = _unet_block(in_channels, f)
        self.pool1 = nn.MaxPool1d(kernel_size=2, stride=2)
        self.encoder2 = _unet_block(f, f * 2)
        self.pool2 = nn.MaxPool1d(kernel_size=2, stride=2)
        self.bottleneck = _unet_block(f * 2, f * 4)
        self.upconv2 = nn.ConvTranspose1d(f * 4, f * 2, kernel_size=2, stride=2)
        self.decoder2 = _unet_block(f * 4, f * 2)
        self.upconv1 = nn.ConvTranspose1d(f * 2, f, kernel_size=2, stride=2)
        self.decoder1 = _unet_block(f * 2, f)
        self.conv = nn.Conv1d(in_channels=f, out_channels=out_channels, kernel_size=1)

    def forward(self, x, L=None):
        L = L if L is not None else self.L
        h = torch.stack((x, get_grid(x, L)), 1)  # [B, 2, D]
        if self.backend != "hip":
            enc1 = self.encoder1(h)
            enc2 = self.encoder2(self.pool1(enc1))
            bottleneck = self.bottleneck(self.pool2(enc2))
            dec2 = self.decoder2(torch.cat((self.upconv2(bottleneck), enc2), dim=1))
            dec1 = self.decoder1(torch.cat((self.upconv1(dec2), enc1), dim=1))
            return self.conv(dec1).squeeze(1)
        if x.shape[-1] % 4:
            raise RuntimeError(f"UNet1D needs a multiple of 4 points for its skip concatenations (got {x.shape[-1]})")
        relu = _lib.INC_ACT_RELU

        def block(t, seq):
            return _conv_act(_conv_act(t, seq[0], relu), seq[2], relu)

        enc1 = block(h, self.encoder1)
        enc2 = block(max_pool2(enc1), self.encoder2)
        bottleneck = block(max_pool2(enc2), self.bottleneck)
        dec2 = block(upconv2_cat(bottleneck, self.upconv2, enc2), self.decoder2)
        dec1 = block(upconv2_cat(dec2, self.upconv1, enc1), self.decoder1)
        return _conv_act(dec1, self.conv, _lib.INC_ACT_NONE).squeeze(1)
