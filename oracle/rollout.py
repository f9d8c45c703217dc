"""Rollout driver restated from trainer/trainer_1d.py:154-183 (test
infrastructure only)."""
import torch


def simulate_rollout(model, solver, u_initial, mstep, correction_term, L=None):
    """Unroll ``mstep`` solver steps, evaluating s = model(u[, L]) before each
    step when correction_term == 'INC' (trainer_1d.py:169-182).  Returns
    (results [B, mstep, D], corrections [B, mstep, D] or None)."""
    b, d = u_initial.shape
    outs, corr = [], []
    solver.correction_s = None
    u = u_initial
    for _ in range(mstep):
        if correction_term == "INC":
            s = model(u) if L is None else model(u, L)
            solver.correction_s = s
            corr.append(s)
        elif correction_term is not None:
            raise ValueError(f"Invalid correction term: {correction_term}")
        u = solver.step(u) if L is None else solver.step(u, L)
        outs.append(u)
    res = torch.stack(outs, 1) if outs else u_initial.new_zeros(b, 0, d)
    return res, (torch.stack(corr, 1) if correction_term == "INC" else None)
