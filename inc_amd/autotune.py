"""Launch-shape autotuner.

Times every valid launch shape of a solver kernel on the live GPU and keeps
the fastest, cached in a JSON file keyed by GPU architecture and problem
shape (``$INC_AMD_TUNE_CACHE``, default ``~/.cache/inc_amd/tune.json``):

* KS: FFT points per lane (radix 2 / 4 / 8, float64 2 / 4; lanes per
  trajectory pair = n/radix, at most 256);
* Burgers (uniform batch): threads per trajectory (team);
* Burgers (ragged batch): threads per trajectory block (64 / 128 / 256).

Solvers ask for a tuned shape when their parameters say ``radix='auto'`` /
``team='auto'`` or the environment sets ``INC_AMD_AUTOTUNE=1``.  Tuning
launches kernels, so the first (tuning) call must not be inside a hipGraph
capture.
"""
import json
import math
import os
import threading

import torch

_lock = threading.Lock()
_mem = None


def cache_path():
    return os.environ.get("INC_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "inc_amd", "tune.json"))


def enabled():
    return os.environ.get("INC_AMD_AUTOTUNE", "0") not in ("", "0", "false", "False")


def _arch():
    try:
        return torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName.split(":")[0]
    except Exception:  # noqa: BLE001 - a missing property only changes the cache key
        return "gpu"


def _load():
    global _mem
    if _mem is None:
        try:
            with open(cache_path()) as f:
                _mem = json.load(f)
        except (OSError, ValueError):
            _mem = {}
    return _mem


def _store(key, value):
    mem = _load()
    mem[key] = value
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(mem, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # read-only home: keep the in-memory result


def lookup(key):
    with _lock:
        return _load().get(key)


def _time_ms(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    best = math.inf
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def ks_radix_candidates(n, dtype):
    """FFT radices of the register transform (power-of-two n in [16, 1024]);
    other resolutions run the mixed-radix transform, which has no radix."""
    if not (16 <= n <= 1024 and n & (n - 1) == 0):
        return []
    radices = (2, 4) if dtype == torch.float64 else (2, 4, 8)
    return [r for r in radices if r <= n and n // r <= 256]


def tune_ks(n, batch, dtype=torch.float32, scheme="ETDRK4", L=64.0, dt=0.25, steps=16, record=None):
    """Best FFT radix for a KS batch (forward + adjoint of ``steps`` steps);
    0 (no choice) for the resolutions of the mixed-radix transform."""
    if not ks_radix_candidates(n, dtype):
        return 0
    key = f"{_arch()}|ks|{scheme}|n{n}|b{batch}|{str(dtype).split('.')[-1]}"
    hit = lookup(key)
    if hit is not None:
        return hit
    from .config_1d import KSSimParams
    from .solver_1d import KSSolverTorch
    dev = torch.device("cuda")
    gen = torch.Generator().manual_seed(0)
    u0 = (0.5 * torch.randn(batch, n, generator=gen)).to(dev, dtype)
    cot = torch.randn(batch, steps, n, generator=gen).to(dev, dtype)
    Ld = torch.tensor([L], device=dev)
    results = {}
    for r in ks_radix_candidates(n, dtype):
        p = KSSimParams([L], resolution=n, dt=dt, time_scheme=scheme, device="cuda")
        p.radix = r
        sv = KSSolverTorch(p)
        sv.doublestep = False

        def run():
            u = u0.clone().requires_grad_(True)
            torch.autograd.backward(sv.rollout(u, steps, Ld), cot)
        results[r] = _time_ms(run)
    best = min(results, key=results.get)
    if record is not None:
        record.update({str(k): v for k, v in results.items()})
    with _lock:
        _store(key, best)
    return best


def burgers_team_candidates(n):
    return [t for t in (16, 32, 64, 128, 256) if 16 * t >= n and t <= max(16, 4 * n)]


def tune_burgers(n, batch, dtype=torch.float32, scheme="SSPRK3", convection="WENO", steps=8, record=None):
    """Best threads-per-trajectory for a uniform Burgers batch."""
    key = f"{_arch()}|burgers|{scheme}|{convection}|n{n}|b{batch}|{str(dtype).split('.')[-1]}"
    hit = lookup(key)
    if hit is not None:
        return hit
    from .config_1d import BGSimParams
    from .solver_1d import BurgersSolverTorch
    dev = torch.device("cuda")
    gen = torch.Generator().manual_seed(0)
    u0 = torch.sin(torch.arange(n, dtype=torch.float64) * 2 * math.pi / n)[None].repeat(batch, 1)
    u0 = (u0 + 0.01 * torch.randn(batch, n, generator=gen, dtype=torch.float64)).to(dev, dtype)
    cot = torch.randn(batch, steps, n, generator=gen).to(dev, dtype)
    results = {}
    for team in burgers_team_candidates(n):
        p = BGSimParams(dt=0.2 * 2 * math.pi / n, resolution=n, nu=0.01, batch_size=batch, time_scheme=scheme,
                        convection_term=convection, device="cuda")
        p.team = team
        sv = BurgersSolverTorch(p)

        def run():
            u = u0.clone().requires_grad_(True)
            torch.autograd.backward(sv.rollout(u, steps), cot)
        results[team] = _time_ms(run)
    best = min(results, key=results.get)
    if record is not None:
        record.update({str(k): v for k, v in results.items()})
    with _lock:
        _store(key, best)
    return best


def tune_ragged(rb, steps=4, record=None):
    """Best block size for a ragged Burgers batch (keyed by its n-histogram)."""
    hist = {}
    for v in rb.n:
        hist[v] = hist.get(v, 0) + 1
    sig = ",".join(f"{k}x{hist[k]}" for k in sorted(hist))
    key = f"{_arch()}|ragged|{rb.scheme}|{rb.convection}|{sig}|{str(rb.dtype).split('.')[-1]}"
    hit = lookup(key)
    if hit is not None:
        return hit
    gen = torch.Generator().manual_seed(0)
    u0 = (0.5 * torch.randn(rb.total, generator=gen)).to(rb.device, rb.dtype)
    cot = torch.randn(steps, rb.total, generator=gen).to(rb.device, rb.dtype)
    saved = rb.team
    results = {}
    for team in (64, 128, 256):
        if rb.nmax > 16 * team:
            continue
        rb.team = team

        def run():
            u = u0.clone().requires_grad_(True)
            torch.autograd.backward(rb.advance(u, None, steps=steps, want_traj=True), cot)
        results[team] = _time_ms(run)
    rb.team = saved
    best = min(results, key=results.get)
    if record is not None:
        record.update({str(k): v for k, v in results.items()})
    with _lock:
        _store(key, best)
    return best
