"""inc_conv1d_bwd_weight_noalloc (the weight gradient with its partial sums
in the library's per-GPU scratch array) against inc_conv1d_bwd_weight with a
caller workspace, through the C ABI.

With a workspace large enough for the same number of partial rows both
entry points run the same kernels with the same launch shape and reduce in
the same fixed order, so the results are bit-identical; both are also
checked against torch's FP64 weight gradient (float32: 1e-5 relative,
float64: 1e-12).  Two calls in a row, the second with a smaller input, show
that the shared array keeps no state that matters between calls.
"""
import ctypes

import pytest
import torch
from torch.nn import functional as F

import inc_amd._lib as L
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20


def _ref(x, w, b, g):
    xr, wr, br = (t.detach().double().cpu().requires_grad_(True) for t in (x, w, b))
    p = w.shape[-1] // 2
    y = F.conv1d(F.pad(xr, (p, p), mode="circular"), wr, br)
    (y * g.double().cpu()).sum().backward()
    return wr.grad, br.grad


def _with_workspace(lib, x, w, g, elems):
    bsz, ci, n = x.shape
    co, _, k = w.shape
    gw, gb = torch.empty_like(w), x.new_empty(co)
    work = x.new_empty(elems)
    rc = lib.inc_conv1d_bwd_weight(L.dtype_code(x.dtype), bsz, ci, co, n, k, L.ptr(g), L.ptr(x), L.ptr(gw),
                                   L.ptr(gb), L.ptr(work), work.numel(), L.stream())
    L.check(rc, "inc_conv1d_bwd_weight")
    return gw, gb


def _no_alloc(lib, x, w, g):
    bsz, ci, n = x.shape
    co, _, k = w.shape
    gw, gb = torch.empty_like(w), x.new_empty(co)
    rc = lib.inc_conv1d_bwd_weight_noalloc(L.dtype_code(x.dtype), bsz, ci, co, n, k, L.ptr(g), L.ptr(x),
                                           L.ptr(gw), L.ptr(gb), L.stream())
    L.check(rc, "inc_conv1d_bwd_weight_noalloc")
    return gw, gb


# (batch, in channels, out channels, n, kernel) pairs: the first call, then a
# smaller second call; float32 shapes take the matrix-core kernel, float64
# the vector-ALU kernel.
CASES = [((64, 32, 32, 256, 5), (3, 32, 32, 128, 5)), ((32, 48, 48, 512, 7), (2, 48, 1, 100, 7)),
         ((16, 1, 32, 128, 5), (1, 1, 32, 60, 5))]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_no_alloc_weight_gradient_matches_workspace_entry(cuda, dtype, case):
    lib = L.load()
    gen = torch.Generator().manual_seed(sum(case[0]))
    for shape in case:
        bsz, ci, co, n, k = shape
        x = torch.randn(bsz, ci, n, generator=gen, dtype=torch.float64)
        w = torch.randn(co, ci, k, generator=gen, dtype=torch.float64) / (ci * k) ** 0.5
        b = torch.randn(co, generator=gen, dtype=torch.float64)
        g = torch.randn(bsz, co, n, generator=gen, dtype=torch.float64)
        xd, wd, gd = (t.to(cuda, dtype).contiguous() for t in (x, w, g))
        elems = SCRATCH_BYTES // xd.element_size()
        gw_n, gb_n = _no_alloc(lib, xd, wd, gd)
        gw_w, gb_w = _with_workspace(lib, xd, wd, gd, elems)
        torch.cuda.synchronize()
        assert torch.equal(gw_n, gw_w) and torch.equal(gb_n, gb_w), shape
        rw, rb = _ref(x, w, b, g)
        tol = 1e-5 if dtype == torch.float32 else 1e-12
        assert rel_l2(gw_n, rw) < tol and rel_l2(gb_n, rb) < tol, shape


def test_no_alloc_rejects_a_partial_row_beyond_the_scratch(cuda):
    lib = L.load()
    # Co*Ci*K + Co = 2000*2000*9 + 2000 float64 elements > 32 MiB: refused
    # before anything is read (the pointers are never dereferenced)
    rc = lib.inc_conv1d_bwd_weight_noalloc(L.INC_F64, 1, 2000, 2000, 64, 9, None, None, None, None, L.stream())
    assert rc == L.INC_EINVAL and b"scratch" in lib.inc_last_error()
