"""The KS kernels' register-transpose FFT plan, restated in numpy and checked
against numpy.fft for every supported (N, radix).

A team of TL = N/R lanes holds the physical points j + m*TL in its R
registers.  Decimation in frequency: pass 0 is a radix-R butterfly over the
registers; pass s >= 1 swaps the low d_s register bits with the next lane
digit (taken from the top, d_s <= log2 R bits) and runs a radix-2^d_s
butterfly; between passes the twiddle W_{M_s}^{n_rest * k_s} multiplies
register k_s.  The inverse runs the passes backwards with conjugates.  The
mode held by (register m, lane j) afterwards is the spectral index map that
the kernels' coefficient load follows.  This is the algorithm of
inc_amd/csrc/ks.hip (RPlan, spec_bits, load_reg_twiddles, reg_fft_fwd/inv);
the kernels themselves are checked against the oracle by the GPU tests.
"""
import numpy as np
import pytest


def plan(N, R):
    rho = R.bit_length() - 1
    t = N.bit_length() - 1 - rho
    d, c = [rho], [t]
    rem = t
    while rem > 0:
        ds = min(rho, rem)
        rem -= ds
        d.append(ds)
        c.append(rem)
    return rho, t, d, c


def swap_bit(v, b, c):
    """Exchange register bit b with lane bit c (v: [R, TL] complex)."""
    lane = np.arange(v.shape[1])
    hi = (lane >> c) & 1
    out = v.copy()
    for m in range(v.shape[0]):
        if m >> b & 1:
            continue
        A, B = v[m], v[m | (1 << b)]
        out[m] = np.where(hi == 1, B[lane ^ (1 << c)], A)
        out[m | (1 << b)] = np.where(hi == 1, B, A[lane ^ (1 << c)])
    return out


def butterfly(v, ds, sign):
    P = 1 << ds
    W = np.exp(sign * 2j * np.pi * np.outer(np.arange(P), np.arange(P)) / P)
    out = v.copy()
    for q in range(v.shape[0] // P):
        idx = [(q << ds) | r for r in range(P)]
        out[idx] = W @ v[idx]
    return out


def twiddle(v, N, R, s, conj):
    rho, t, d, c = plan(N, R)
    P, Ms = 1 << d[s], 1 << (c[s] + d[s])
    nrest = np.arange(v.shape[1]) & ((1 << c[s]) - 1)
    out = v.copy()
    for m in range(v.shape[0]):
        r = m & (P - 1)
        if r:
            w = np.exp(-2j * np.pi * ((nrest * r * (N // Ms)) % N) / N)
            out[m] = v[m] * (np.conj(w) if conj else w)
    return out


def exchange(v, N, R, s):
    rho, t, d, c = plan(N, R)
    for i in range(d[s]):
        v = swap_bit(v, i, c[s] + i)
    return v


def forward(v, N, R):
    S = len(plan(N, R)[2])
    for s in range(S):
        if s > 0:
            v = exchange(v, N, R, s)
        v = butterfly(v, plan(N, R)[2][s], -1)
        if s < S - 1:
            v = twiddle(v, N, R, s, False)
    return v


def inverse(v, N, R):
    S = len(plan(N, R)[2])
    for s in range(S - 1, -1, -1):
        v = butterfly(v, plan(N, R)[2][s], +1)
        if s > 0:
            v = exchange(v, N, R, s)
            v = twiddle(v, N, R, s - 1, True)
    return v


def spec_index(N, R):
    """Mode index at (register m, lane j) after the forward passes."""
    rho, t, d, c = plan(N, R)
    kbit = [0] * (rho + t)   # position p: register bit p (p < rho) or lane bit p - rho
    done = 0
    for s in range(len(d)):
        if s > 0:
            for i in range(d[s]):
                q = rho + c[s] + i
                kbit[i], kbit[q] = kbit[q], kbit[i]
        for i in range(d[s]):
            kbit[i] = done + i
        done += d[s]
    TL = N // R
    k = np.zeros((R, TL), dtype=np.int64)
    for m in range(R):
        for j in range(TL):
            for p in range(rho + t):
                bit = (m >> p) & 1 if p < rho else (j >> (p - rho)) & 1
                k[m, j] |= bit << kbit[p]
    return k


CASES = [(n, r) for r in (2, 4, 8) for n in (16, 32, 64, 128, 256, 512, 1024) if n // r <= 256]


@pytest.mark.parametrize("n,r", CASES)
def test_register_fft_plan_matches_numpy(n, r):
    rng = np.random.default_rng(n + r)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tl = n // r
    v = x.reshape(r, tl)                      # register m, lane j holds x[j + m*tl]
    X = forward(v, n, r)
    k = spec_index(n, r)
    assert sorted(k.ravel()) == list(range(n))  # a permutation of the modes
    ref = np.fft.fft(x)
    np.testing.assert_allclose(X, ref[k], rtol=0, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(inverse(X, n, r), n * v, rtol=0, atol=1e-12 * n * np.abs(v).max())


def test_plan_shapes():
    # digit sizes from the top of the lane bits; lane bits >= 6 cross wavefronts
    assert plan(256, 2)[2] == [1] * 8
    assert plan(256, 4)[2:] == ([2, 2, 2, 2], [6, 4, 2, 0])
    assert plan(128, 4)[2:] == ([2, 2, 2, 1], [5, 3, 1, 0])
    assert plan(1024, 4)[2:] == ([2, 2, 2, 2, 2], [8, 6, 4, 2, 0])
