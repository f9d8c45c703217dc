"""GPU parity of the circular conv1d kernels of the correction network against
torch's FP64 CPU convolution (the same definition as
torch.nn.Conv1d(padding_mode='circular')).  float64: 1e-12 relative;
float32: 1e-5 relative."""
import pytest
import torch
from torch.nn import functional as F

from inc_amd._lib import INC_ACT_GELU
from inc_amd.correction import CorrectionCNN, _conv_bwd_data, _conv_fwd, circular_conv1d
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def ref_conv(x, w, b):
    p = w.shape[-1] // 2
    return F.conv1d(F.pad(x, (p, p), mode="circular"), w, b)


# (batch, in channels, out channels, n, kernel size).  float32 weight
# gradients run on the matrix cores when Co <= 64 and Ci*K + 1 <= 512: one
# output-channel tile (Co = 1), thin layers whose waves split positions
# (Ci = 1), one workspace slot (B * tiles = 1, idle waves), a partial last
# position tile (n = 130); (1, 64, 64, 256, 9) exceeds the column limit and
# takes the vector-ALU kernel.  float32 forward and input-gradient passes run
# on the matrix cores when the pass's input channels are a multiple of 16 and
# its output channels lie in [16, 64]: one to four output tiles (Co = 16, 32,
# 48, 64), a padded tile (Co = 24), 64-, 128- and 256-position block tiles
# with partial last tiles (n = 100, 300) and n below one tile with several
# wraps of the halo (n = 16, 40), K = 1; (1, 64, 64, 256, 9) exceeds LDS and
# takes the vector-ALU kernel.  Layers of more than 64 channels run as
# channel blocks: input blocks of 64 accumulated in turn (80 = 64 + 16,
# 144 = 64 + 64 + 16), equal output blocks (72 = 2 x 36, 144 = 3 x 48), and
# weight gradients in 64 x 64 blocks.  Rows of 16 or 32 points are packed
# 64 / n to a tile (a partial last job: B = 5 at n = 32, B = 7 and B = 1 at
# n = 16).
SHAPES = [(3, 1, 32, 128, 5), (3, 32, 32, 128, 5), (2, 32, 1, 128, 5), (2, 48, 48, 100, 7), (5, 16, 8, 64, 3),
          (1, 8, 8, 9, 9), (4, 32, 32, 1024, 5), (2, 48, 1, 512, 7), (3, 1, 48, 130, 7), (1, 1, 32, 60, 5),
          (1, 64, 64, 256, 9), (2, 48, 48, 512, 7), (2, 32, 24, 40, 5), (1, 16, 16, 16, 9), (3, 32, 16, 64, 1),
          (2, 64, 64, 300, 3), (2, 80, 72, 96, 3), (2, 128, 128, 64, 5), (3, 16, 144, 70, 5), (2, 144, 16, 50, 1),
          (5, 32, 48, 32, 5), (7, 64, 64, 16, 5), (3, 128, 96, 32, 3)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_conv_forward_and_gradients(cuda, shape, dtype):
    bsz, ci, co, n, k = shape
    gen = torch.Generator().manual_seed(sum(shape))
    x = torch.randn(bsz, ci, n, generator=gen, dtype=torch.float64)
    w = torch.randn(co, ci, k, generator=gen, dtype=torch.float64) / (ci * k) ** 0.5
    b = torch.randn(co, generator=gen, dtype=torch.float64)
    g = torch.randn(bsz, co, n, generator=gen, dtype=torch.float64)
    xr, wr, br = (t.clone().requires_grad_(True) for t in (x, w, b))
    yr = ref_conv(xr, wr, br)
    (yr * g).sum().backward()
    xg, wg, bg = (t.to(cuda, dtype).requires_grad_(True) for t in (x, w, b))
    y = circular_conv1d(xg, wg, bg)
    (y * g.to(cuda, dtype)).sum().backward()
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert rel_l2(y, yr.detach()) < tol
    assert rel_l2(xg.grad, xr.grad) < tol
    assert rel_l2(wg.grad, wr.grad) < tol
    assert rel_l2(bg.grad, br.grad) < tol


@pytest.mark.parametrize("shape", [(3, 32, 32, 128, 5), (2, 48, 1, 100, 7), (2, 16, 48, 512, 7),
                                   (2, 96, 80, 128, 5), (3, 48, 32, 32, 7), (5, 16, 80, 16, 5)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_conv_gelu_epilogue_and_scaled_input_gradient(cuda, shape, dtype):
    # inc_conv1d_fwd_act: y = gelu(v), dy = gelu'(v) with v = conv(x) + b;
    # inc_conv1d_bwd_data_mul: gx = (conv^T g) * m
    bsz, ci, co, n, k = shape
    gen = torch.Generator().manual_seed(sum(shape) + 1)
    x = torch.randn(bsz, ci, n, generator=gen, dtype=torch.float64)
    w = 1.5 * torch.randn(co, ci, k, generator=gen, dtype=torch.float64) / (ci * k) ** 0.5
    b = torch.randn(co, generator=gen, dtype=torch.float64)
    g = torch.randn(bsz, co, n, generator=gen, dtype=torch.float64)
    m = torch.rand(bsz, ci, n, generator=gen, dtype=torch.float64)
    v = ref_conv(x, w, b).requires_grad_(True)
    h = F.gelu(v)
    (dv,) = torch.autograd.grad(h.sum(), v)
    xr = x.clone().requires_grad_(True)
    (gxr,) = torch.autograd.grad((ref_conv(xr, w, None) * g).sum(), xr)
    c = lambda t: t.to(cuda, dtype).contiguous()
    y, dy = _conv_fwd(c(x), c(w), c(b), act=INC_ACT_GELU)
    gx = _conv_bwd_data(c(g), c(w), ci, c(m))
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert rel_l2(y, h.detach()) < tol
    assert rel_l2(dy, dv) < tol
    assert rel_l2(gx, gxr * m) < tol
    # without the derivative (no gradient wanted): the same y, bit for bit
    y2, dy2 = _conv_fwd(c(x), c(w), c(b), act=INC_ACT_GELU, need_dy=False)
    assert dy2 is None and torch.equal(y2, y)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 1e-5)])
def test_cnn_backends_agree(cuda, dtype, tol):
    gen = torch.Generator().manual_seed(0)
    hip = CorrectionCNN(4, 32, 5, generator=gen).cuda().to(dtype)
    ref = CorrectionCNN(4, 32, 5, backend="torch").cuda().to(dtype)
    ref.load_state_dict(hip.state_dict())
    u = torch.randn(6, 128, dtype=torch.float64).to(cuda, dtype)
    with torch.no_grad():
        a0 = hip(u)  # the evaluation path: no derivative computed or stored
    a, r = hip(u), ref(u)
    assert rel_l2(a, r) < tol
    assert torch.equal(a0, a.detach())
    a.sum().backward()
    r.sum().backward()
    for p, q in zip(hip.parameters(), ref.parameters()):
        assert rel_l2(p.grad, q.grad) < 10 * tol


# The KS INC and ragged INC workloads' layers (B = 256, N = 512, 48 -> 48 k7
# and the thin 1 -> 48, 48 -> 1 layers): enough (row, position-tile) jobs that
# every block of the matrix-core forward / input-gradient kernels and of the
# weight-gradient kernel loops over several jobs, with the next job's tile
# prefetched and the staging indices computed once per block.
@pytest.mark.parametrize("shape", [(256, 48, 48, 512, 7), (256, 1, 48, 512, 7), (256, 48, 1, 512, 7)],
                         ids=lambda s: "x".join(map(str, s)))
def test_conv_multi_job_blocks_fp32(cuda, shape):
    bsz, ci, co, n, k = shape
    gen = torch.Generator().manual_seed(sum(shape) + 2)
    x = torch.randn(bsz, ci, n, generator=gen, dtype=torch.float64)
    w = torch.randn(co, ci, k, generator=gen, dtype=torch.float64) / (ci * k) ** 0.5
    b = torch.randn(co, generator=gen, dtype=torch.float64)
    g = torch.randn(bsz, co, n, generator=gen, dtype=torch.float64)
    xr, wr, br = (t.clone().requires_grad_(True) for t in (x, w, b))
    yr = ref_conv(xr, wr, br)
    (yr * g).sum().backward()
    xg, wg, bg = (t.to(cuda, torch.float32).requires_grad_(True) for t in (x, w, b))
    y = circular_conv1d(xg, wg, bg)
    (y * g.to(cuda, torch.float32)).sum().backward()
    assert rel_l2(y, yr.detach()) < 1e-5
    assert rel_l2(xg.grad, xr.grad) < 1e-5
    assert rel_l2(wg.grad, wr.grad) < 1e-5
    assert rel_l2(bg.grad, br.grad) < 1e-5
    # the fused GELU epilogue on the same multi-job launch
    if ci > 1 or co > 1:
        v = ref_conv(x, w, b)
        y2, dy2 = _conv_fwd(xg.detach().contiguous(), wg.detach().contiguous(), bg.detach().contiguous(),
                            act=INC_ACT_GELU)
        vr = v.clone().requires_grad_(True)
        (dvr,) = torch.autograd.grad(F.gelu(vr).sum(), vr)
        assert rel_l2(y2, F.gelu(v)) < 1e-5
        assert rel_l2(dy2, dvr) < 1e-5
