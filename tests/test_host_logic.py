"""Host-side logic that needs no GPU: ragged-batch packing, grouping and issue
order, the stable time step, the autotuner cache, the KS table cache key."""
import json
import math

import torch

from inc_amd import autotune
from inc_amd.batched import RaggedBurgers, stable_dt


def test_ragged_packing_groups_and_order():
    ns = [128, 64, 1024, 64, 256]
    hz = [10, 30, 5, 30, 20]
    rb = RaggedBurgers(ns, [0.01] * 5, [0.001] * 5, horizon=hz, device="cpu", dtype=torch.float64)
    assert rb.total == sum(ns)
    assert rb.n == sorted(ns)  # grouped by resolution
    assert [(g[0], g[2]) for g in rb.groups] == [(64, 2), (128, 1), (256, 1), (1024, 1)]
    fields = [torch.full((n,), float(b)) for b, n in enumerate(ns)]
    packed = rb.pack(fields)
    for b, v in enumerate(rb.unpack(packed)):
        assert v.shape == (ns[b],) and torch.all(v == b)
    views = rb.group_views(packed)
    assert [tuple(v.shape) for v in views] == [(2, 64), (1, 128), (1, 256), (1, 1024)]
    # longest-first issue order by n * horizon (internal indices)
    cost = [rb.n[i] * rb.horizon[i] for i in range(rb.count)]
    order = rb.order_d.tolist()
    assert [cost[i] for i in order] == sorted(cost, reverse=True)
    m = rb.valid_mask(32)
    for i, h in enumerate(rb.horizon):
        seg = m[:, rb.offsets[i]:rb.offsets[i] + rb.n[i]]
        assert torch.all(seg[:min(h, 32)] == 1) and torch.all(seg[h:] == 0)


def test_stable_dt_is_min_of_advective_and_viscous():
    n, nu = 1024, 0.03
    dx = 2 * math.pi / n
    assert stable_dt(n, nu) == min(0.4 * dx, 0.3 * dx * dx / nu) == 0.3 * dx * dx / nu
    assert stable_dt(64, 0.003) == 0.4 * (2 * math.pi / 64)


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    monkeypatch.setenv("INC_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    monkeypatch.setattr(autotune, "_mem", None)
    assert autotune.lookup("k") is None
    autotune._store("k", 8)
    assert autotune.lookup("k") == 8
    assert json.loads((tmp_path / "t.json").read_text()) == {"k": 8}
    monkeypatch.setattr(autotune, "_mem", None)
    assert autotune.lookup("k") == 8  # reloaded from disk
    assert autotune.ks_radix_candidates(1024, torch.float32) == [4, 8]
    assert autotune.ks_radix_candidates(16, torch.float32) == [2, 4, 8]
    assert autotune.ks_radix_candidates(256, torch.float64) == [2, 4]
    assert autotune.ks_radix_candidates(1024, torch.float64) == [4]
    assert 16 not in autotune.burgers_team_candidates(4096)


def test_correction_cnn_output_scale():
    # CorrectionCNN(scale=s) is s * CorrectionCNN() on the same weights (torch backend on the CPU),
    # and its gradients scale with it
    from inc_amd.correction import CorrectionCNN
    gen = torch.Generator().manual_seed(0)
    a = CorrectionCNN(3, 32, 5, generator=gen, backend="torch", scale=0.05).double()
    b = CorrectionCNN(3, 32, 5, backend="torch").double()
    b.load_state_dict(a.state_dict())
    u = torch.randn(2, 24, generator=gen, dtype=torch.float64)
    ya, yb = a(u), b(u)
    assert torch.allclose(ya, 0.05 * yb, rtol=1e-14, atol=0)
    ya.sum().backward()
    yb.sum().backward()
    for pa, pb in zip(a.parameters(), b.parameters()):
        assert torch.allclose(pa.grad, 0.05 * pb.grad, rtol=1e-12, atol=1e-300)


def test_ks_initial_conditions_structure_and_batch_norm():
    # KSSolverTorch.initial_conditions (solver_1d.py:214-223) is eager torch
    # on the solver's device, no HIP call: cos(2 pi x) + alpha sin(2 pi x) on
    # the endpoint-inclusive linspace(0, 1, N), plus noise eps/||eps|| whose
    # Frobenius norm over the whole batch is 1.
    import inc_amd
    torch.manual_seed(0)
    p = inc_amd.KSSimParams([22.0, 30.0, 64.0, 100.0], resolution=40, device="cpu")
    sv = inc_amd.KSSolverTorch(p)
    u = sv.initial_conditions()
    x = torch.linspace(0, 1, 40)
    det = torch.cos(2 * torch.pi * x) + p.alpha[:, None] * torch.sin(2 * torch.pi * x)
    noise = (u - det).double()
    assert abs(float(torch.linalg.norm(noise)) - 1.0) < 1e-5
    assert float((noise.norm(dim=1) - 1).abs().max()) > 1e-2
    sv.alpha = p.alpha[:3]
    try:
        sv.initial_conditions()
    except ValueError:
        pass
    else:
        raise AssertionError("mismatched alpha / domain_size lengths must raise ValueError")


def test_resnet1d_module_tree_is_the_reference_s():
    """ResNet1D keeps models/models_1d/resnet.py:22-52's parameter names (its
    state dicts load) and builds the same weights from a seed in either
    backend; the torch backend runs on the CPU."""
    from inc_amd.correction import ResNet1D
    keys = (["initial.0.weight", "initial.0.bias"]
            + [f"blocks.{i}.{c}.{p}" for i in range(6) for c in ("conv1", "conv2") for p in ("weight", "bias")]
            + ["final.weight", "final.bias"])
    assert list(ResNet1D().state_dict().keys()) == keys
    torch.manual_seed(3)
    a = ResNet1D()
    torch.manual_seed(3)
    b = ResNet1D(backend="torch")
    assert all(torch.equal(p, q) for p, q in zip(a.parameters(), b.parameters()))
    assert b(torch.randn(2, 32, dtype=torch.float32), 16.0).shape == (2, 32)
    # a block whose widths differ gets the reference's 1x1 shortcut convolution
    from inc_amd.correction import _ResidualBlock1D
    assert list(_ResidualBlock1D(8, 16).state_dict().keys()) == [
        "conv1.weight", "conv1.bias", "conv2.weight", "conv2.bias", "shortcut.weight", "shortcut.bias"]


def test_unet1d_module_tree_is_the_reference_s():
    """UNet1D keeps models/models_1d/unet.py:5-53's module tree, parameter
    names and shapes (its state dicts load) and builds the same weights from
    a seed in either backend; the torch backend runs on the CPU.  The HIP
    backend refuses host tensors and point counts that are not a multiple of
    4 before any launch."""
    from inc_amd.correction import UNet1D
    f = 32
    blocks = {"encoder1": (2, f), "encoder2": (f, 2 * f), "bottleneck": (2 * f, 4 * f),
              "decoder2": (4 * f, 2 * f), "decoder1": (2 * f, f)}
    shapes = {}
    for name in ("encoder1", "encoder2", "bottleneck"):
        ci, co = blocks[name]
        shapes.update({f"{name}.0.weight": (co, ci, 5), f"{name}.0.bias": (co,),
                       f"{name}.2.weight": (co, co, 5), f"{name}.2.bias": (co,)})
    shapes.update({"upconv2.weight": (4 * f, 2 * f, 2), "upconv2.bias": (2 * f,)})
    ci, co = blocks["decoder2"]
    shapes.update({"decoder2.0.weight": (co, ci, 5), "decoder2.0.bias": (co,), "decoder2.2.weight": (co, co, 5),
                   "decoder2.2.bias": (co,)})
    shapes.update({"upconv1.weight": (2 * f, f, 2), "upconv1.bias": (f,)})
    ci, co = blocks["decoder1"]
    shapes.update({"decoder1.0.weight": (co, ci, 5), "decoder1.0.bias": (co,), "decoder1.2.weight": (co, co, 5),
                   "decoder1.2.bias": (co,)})
    shapes.update({"conv.weight": (1, f, 1), "conv.bias": (1,)})
    sd = UNet1D().state_dict()
    assert list(sd.keys()) == list(shapes.keys())
    assert all(tuple(sd[k].shape) == v for k, v in shapes.items())
    torch.manual_seed(3)
    a = UNet1D()
    torch.manual_seed(3)
    b = UNet1D(backend="torch")
    assert all(torch.equal(p, q) for p, q in zip(a.parameters(), b.parameters()))
    assert b(torch.randn(2, 32, dtype=torch.float32), 16.0).shape == (2, 32)
    try:
        a(torch.randn(2, 30), 16.0)
    except RuntimeError as e:
        assert "multiple of 4" in str(e)
    else:
        raise AssertionError("30 points must raise")
    try:
        a(torch.randn(2, 32), 16.0)
    except ValueError as e:
        assert "GPU" in str(e)
    else:
        raise AssertionError("host tensors must raise")
