"""The float32 GELU of the convolution epilogues (gelu and gelu' from one
exp(-v^2/2) and a fitted erfc polynomial) against the exact function in
float64: identity convolutions (K = 1, unit weights, zero bias) through the
vector-ALU kernel (1 -> 1 channel) and the matrix-core kernel (16 -> 16
channels) put v itself into the epilogue, over v in [-12, 12].

Bars (a float32 restatement of the same arithmetic reaches 7 ulp for gelu
at v >= -1 and 52 ulp for gelu' on [-6, 12] away from its root; the
hardware exp2 and reciprocal add up to an ulp): gelu within 8 ulp of
max(|exact|, 2^-20) for v >= -1, gelu' within 64 ulp for v >= -6 more than
0.05 from its root, and both within 3e-7 * (1 + |v|) absolute everywhere.
Non-finite inputs follow the erf form.
"""
import math

import pytest
import torch

import inc_amd._lib as L

pytestmark = pytest.mark.gpu


def _exact(v):
    v = v.double()
    phi = 0.5 * (1 + torch.special.erf(v / math.sqrt(2)))
    return v * phi, phi + v * torch.exp(-0.5 * v * v) / math.sqrt(2 * math.pi)


def _conv_gelu(v, channels):
    n = v.numel() // channels
    x = v.reshape(1, channels, n).contiguous()
    w = torch.eye(channels, device=v.device).reshape(channels, channels, 1).contiguous()
    b = torch.zeros(channels, device=v.device)
    y, dy = torch.empty_like(x), torch.empty_like(x)
    L.check(L.load().inc_conv1d_fwd_act(L.INC_F32, L.INC_ACT_GELU, 1, channels, channels, n, 1, L.ptr(x), L.ptr(w),
                                        L.ptr(b), L.ptr(y), L.ptr(dy), L.stream()), "gelu conv")
    y2 = torch.empty_like(x)
    L.check(L.load().inc_conv1d_fwd_act(L.INC_F32, L.INC_ACT_GELU, 1, channels, channels, n, 1, L.ptr(x), L.ptr(w),
                                        L.ptr(b), L.ptr(y2), None, L.stream()), "gelu conv, no derivative")
    return y.reshape(-1), dy.reshape(-1), y2.reshape(-1)


def _ulp(ref):
    r32 = ref.float().abs().clamp_min(2.0 ** -20)
    return (torch.nextafter(r32, torch.full_like(r32, math.inf)) - r32).double()


@pytest.mark.parametrize("channels", [1, 16])
def test_conv_gelu_accuracy(cuda, channels):
    v = torch.linspace(-12, 12, 16 * 65536, device=cuda)
    y, dy, y2 = _conv_gelu(v, channels)
    assert torch.equal(y, y2)  # gelu alone is the same value
    g, dg = _exact(v)
    vd = v.double()
    err_g, err_d = (y.double() - g).abs(), (dy.double() - dg).abs()
    assert bool((err_g <= 3e-7 * (1 + vd.abs())).all()), float(err_g.max())
    assert bool((err_d <= 3e-7 * (1 + vd.abs())).all()), float(err_d.max())
    m = vd >= -1
    assert float((err_g[m] / _ulp(g[m])).max()) <= 8
    away = (vd - (-0.7518)).abs() > 0.05  # gelu' has a root near -0.7518
    assert float((err_d[away & (vd >= -6)] / _ulp(dg[away & (vd >= -6)])).max()) <= 64


def test_conv_gelu_non_finite(cuda):
    v = torch.tensor([math.nan, math.inf, -math.inf, 0.0, -0.0] + [0.0] * 11, device=cuda)
    y, dy, _ = _conv_gelu(v, 1)
    assert math.isnan(float(y[0])) and math.isnan(float(dy[0]))
    assert float(y[1]) == math.inf and math.isnan(float(dy[1]))
    assert math.isnan(float(y[2])) and math.isnan(float(dy[2]))
    assert float(y[3]) == 0.0 and float(dy[3]) == 0.5
