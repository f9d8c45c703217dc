"""The GPU dataset generators produce the reference's dataset schema and
trajectories that are consecutive solver steps (also across chunk borders)."""
import io
import math

import pytest
import torch

from inc_amd import BGSimParams, BurgersSolverTorch, KSSimParams, KSSolverTorch
from inc_amd.datagen import generate_burgers, generate_ks
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_generate_ks_schema_and_continuity(cuda):
    d = generate_ks(6, resolution=64, domain_sizes=(22.0, 35.0), dt=0.25, steps=20, warmup=4, chunk=8,
                    dtype=torch.float64)
    tr, L = d["trajectories"], d["domain_size"]
    assert tr.shape == (6, 21, 64) and L.shape == (6, 1) and d["metadata"]["gen_dt"] == 0.25
    sv = KSSolverTorch(KSSimParams(L[:, 0].to(cuda), resolution=64, dt=0.25, time_scheme="ETDRK4", device="cuda"))
    sv.doublestep = False
    for k in (0, 7, 8, 15):  # 7 -> 8 and 15 -> 16 cross chunk borders
        nxt = sv.step(tr[:, k].to(cuda), L[:, 0].to(cuda))
        assert rel_l2(nxt, tr[:, k + 1]) < 1e-12, k
    buf = io.BytesIO()
    torch.save(d, buf)
    buf.seek(0)
    back = torch.load(buf, weights_only=True)
    assert torch.equal(back["trajectories"], tr) and back["metadata"] == d["metadata"]


def test_generate_burgers_schema_and_continuity(cuda):
    d = generate_burgers(4, resolution=64, dt=0.01, steps=10, J=3, chunk=4, dtype=torch.float64)
    tr = d["trajectories"]
    assert tr.shape == (4, 11, 64)
    for key in ("A", "omega", "phi", "l"):
        assert d[key].shape == (4, 1, 3)
    assert d["nu"].shape == (4, 1) and d["metadata"]["L"] == pytest.approx(2 * math.pi)
    p = BGSimParams(dt=0.01, resolution=64, nu=d["nu"].to(cuda), equation_type="learning", batch_size=4,
                    time_scheme="SSPRK3", device="cuda")
    p.A, p.omega, p.phi, p.l = (d[k].to(cuda) for k in ("A", "omega", "phi", "l"))
    sv = BurgersSolverTorch(p)
    x = torch.arange(64, device=cuda, dtype=torch.float64) * (2 * math.pi / 64)
    assert rel_l2(sv.initial_conditions(x, 0.0), tr[:, 0]) < 1e-14
    for k in (0, 3, 4, 8):  # 3 -> 4 and 7 -> 8 cross chunk borders
        sv.t = k * 0.01
        assert rel_l2(sv.step(tr[:, k].to(cuda)), tr[:, k + 1]) < 1e-12, k
