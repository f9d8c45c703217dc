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
The mixed-radix transform of the other resolutions (MixX, mix_plan,
plan_radix, mix_pass, mix_pass_any) and the radix-P register step before
power-of-two register transforms (MixRegX) are restated at the end of this
file.
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


# ---- the mixed-radix transform (MixX in ks.hip) for every other n ----------------
# (radix, 4-bit pass code) as mix_plan encodes them and plan_radix decodes them
_ODD = [(3, 2), (5, 4), (7, 5), (11, 6), (13, 7), (17, 8), (19, 9), (23, 10), (29, 11), (31, 12)]
_DECODE = {1: 2, 2: 3, 3: 4, 4: 5, 5: 7, 6: 11, 7: 13, 8: 17, 9: 19, 10: 23, 11: 29, 12: 31}


def mix_plan(n):
    codes, m = [], n
    while m % 4 == 0:
        codes.append(3)
        m //= 4
    while m % 2 == 0:
        codes.append(1)
        m //= 2
    for p, c in _ODD:
        while m % p == 0:
            codes.append(c)
            m //= p
    if m > 1:
        codes.append(15)
    return codes


def mix_radices(n):
    out, L = [], 1
    for c in mix_plan(n):
        p = _DECODE[c] if c != 15 else n // L
        out.append(p)
        L *= p
    return out


def mix_fft(x, sign):
    """Stockham decimation in time: (L, M) -> (L p, M / p) per pass with
    X'[(k + L t) M' + r] = sum_q w_p^{q t} w_n^{q k M'} X[(k p + q) M' + r]."""
    n = len(x)
    w = np.exp(-2j * np.pi * np.arange(n) / n)
    if sign > 0:
        w = np.conj(w)
    a, L = x.astype(complex), 1
    for p in mix_radices(n):
        mp = n // (L * p)
        out = np.empty(n, complex)
        for k in range(L):
            for r in range(mp):
                v = np.array([a[(k * p + q) * mp + r] * w[q * k * mp] for q in range(p)])
                for t in range(p):
                    out[(k + L * t) * mp + r] = (v * w[(np.arange(p) * t * (n // p)) % n]).sum()
        a, L = out, L * p
    return a


def test_mixed_radix_plan_covers_every_size():
    # every n in [4, 1024] factors into at most 16 passes whose product is n,
    # with code 15 only for the one prime factor above 31
    for n in range(4, 1025):
        codes = mix_plan(n)
        assert len(codes) <= 16
        assert int(np.prod(mix_radices(n))) == n
        if 15 in codes:
            assert codes.index(15) == len(codes) - 1 and mix_radices(n)[-1] > 31


@pytest.mark.parametrize("n", [4, 6, 8, 12, 75, 77, 96, 100, 127, 243, 250, 384])
def test_mixed_radix_transform_matches_numpy(n):
    g = np.random.default_rng(n)
    x = g.standard_normal(n) + 1j * g.standard_normal(n)
    np.testing.assert_allclose(mix_fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(mix_fft(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)


# ---- radix-P register step before power-of-two register transforms (MixRegX) ----
@pytest.mark.parametrize("p,m", [(3, 32), (3, 128), (3, 256), (5, 32), (5, 64), (5, 128)])
def test_radix_p_register_step_matches_numpy(p, m):
    # N = p*m, rs = 2 registers per sub-transform, tl = m / 2 lanes: register
    # mm + rs*n2 of lane j holds x[j + mm*tl + m*n2].  Forward: the radix-p DFT
    # over n2 and the twiddles W_N^{n1 r} (n1 = j + mm*tl) in registers, then
    # the m-point register transform of registers {mm + rs*r}; the mode of
    # register mm + rs*r, lane j is p * spec_index(mm, j) + r.  The inverse
    # runs the steps backwards.
    n, rs = p * m, 2
    tl = m // rs
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v = x.reshape(p * rs, tl)                  # register mm + rs*n2, lane j
    n1 = np.arange(tl)[None, :] + np.arange(rs)[:, None] * tl     # [rs, tl]
    Wp = np.exp(-2j * np.pi * np.outer(np.arange(p), np.arange(p)) / p)
    y = np.einsum("rq,qaj->raj", Wp, v.reshape(p, rs, tl))          # radix-p over n2
    y = y * np.exp(-2j * np.pi * np.arange(p)[:, None, None] * n1[None] / n)
    X = np.stack([forward(y[r], m, rs) for r in range(p)])          # [p, rs, tl]
    k1 = spec_index(m, rs)
    modes = p * k1[None] + np.arange(p)[:, None, None]
    assert sorted(modes.ravel()) == list(range(n))
    ref = np.fft.fft(x)
    np.testing.assert_allclose(X, ref[modes], rtol=0, atol=1e-12 * np.abs(ref).max())
    # inverse: m-point inverses, conjugate twiddles, inverse radix-p
    z = np.stack([inverse(X[r], m, rs) for r in range(p)])
    z = z * np.exp(2j * np.pi * np.arange(p)[:, None, None] * n1[None] / n)
    back = np.einsum("qr,raj->qaj", np.conj(Wp), z).reshape(p * rs, tl)
    np.testing.assert_allclose(back, n * v, rtol=0, atol=1e-11 * n * np.abs(v).max())
