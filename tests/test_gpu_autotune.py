"""The autotuner times every valid launch shape, caches the fastest per GPU
architecture and shape, and the tuned shapes give the same results."""
import json

import pytest
import torch

from inc_amd import autotune
from inc_amd import KSSimParams, KSSolverTorch, RaggedBurgers
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_autotune_ks_and_ragged_cache(cuda, tmp_path, monkeypatch):
    monkeypatch.setenv("INC_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_mem", None)
    rec = {}
    r = autotune.tune_ks(128, 32, torch.float32, "ETDRK4", steps=4, record=rec)
    assert r in (2, 4, 8) and set(rec) == {"2", "4", "8"}
    assert r == autotune.tune_ks(128, 32, torch.float32, "ETDRK4")  # cached
    cache = json.loads((tmp_path / "tune.json").read_text())
    assert any(k.endswith("|ks|ETDRK4|n128|b32|float32") for k in cache)
    # a solver asking for 'auto' uses the tuned radix and matches a fixed one
    p = KSSimParams([22.0], resolution=128, dt=0.25, time_scheme="ETDRK4", device="cuda")
    p.radix = "auto"
    sv = KSSolverTorch(p)
    sv.doublestep = False
    u = torch.randn(32, 128, generator=torch.Generator().manual_seed(0)).to(cuda) * 0.3
    out = sv.step(u, torch.tensor([22.0], device=cuda))
    assert sv.radix == r
    p2 = KSSimParams([22.0], resolution=128, dt=0.25, time_scheme="ETDRK4", device="cuda")
    p2.radix = 4
    sv2 = KSSolverTorch(p2)
    sv2.doublestep = False
    assert rel_l2(out, sv2.step(u, torch.tensor([22.0], device=cuda))) < 1e-6
    # ragged batch
    rb = RaggedBurgers([64, 256, 1024, 128], [0.01] * 4, [0.001] * 4, team="auto")
    u0 = rb.pack([torch.sin(torch.linspace(0, 6.28, n)) for n in (64, 256, 1024, 128)])
    rb.advance(u0, None, steps=2)
    assert rb.team in (64, 128, 256)
