"""Pins the split-precision (INT8 slice) restatement in the oracle.

* Slicing is exact: x == 2^e * (sum_t q_t 2^-7t + r) with |r| < 2^-7S, digits
  share x's sign and |q| <= 127 (checked in exact rational arithmetic).
* The split GEMM obeys the closed-form truncation bound against the
  double-double oracle, and tightens as S grows.
* An independent numpy emulation (exact float64 products of small integers)
  gives the same bits as the C restatement.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc

U = 2.0 ** -53


def _operand(rng, r, c, phi):
    return (rng.random((r, c)) - 0.5) * np.exp(phi * rng.standard_normal((r, c)))


@pytest.mark.parametrize("S", [3, 5, 8, 9])
def test_slices_reconstruct_exactly(S):
    rng = np.random.default_rng(S)
    x = _operand(rng, 1, 300, 3.0).ravel()
    x[:5] = [0.0, -0.0, 5e-324, -1.0, 2.0 ** -1074 * 3]
    e = orc.split_exponent(x)
    mx = np.max(np.abs(x))
    assert 2.0 ** (e - 1) <= mx < 2.0 ** e
    q = orc.split_slices(x, e, S)
    assert np.all(np.abs(q.astype(int)) <= 127)
    for i, xv in enumerate(x):
        digits = [int(q[t, i]) for t in range(S)]
        assert all(d == 0 or (d > 0) == (xv > 0) for d in digits)
        approx = sum(Fraction(d) * Fraction(2) ** (-7 * (t + 1)) for t, d in enumerate(digits))
        scaled = Fraction(float(xv)) / Fraction(2) ** e
        r = scaled - approx
        assert abs(r) < Fraction(2) ** (-7 * S)
        assert r == 0 or (r > 0) == (xv > 0)


def test_exponent_marks_nonfinite_and_zero():
    assert orc.split_exponent(np.zeros(7)) == 0
    assert orc.split_exponent(np.array([1.0, np.inf])) == orc.SPLIT_NONFINITE
    assert orc.split_exponent(np.array([np.nan, 1.0])) == orc.SPLIT_NONFINITE
    assert orc.split_exponent(np.array([0.75])) == 0
    assert orc.split_exponent(np.array([1.0])) == 1


def _truncation_bound(S, k, ea, eb):
    # per-term: residuals 2*2^-7S (plus the product of residuals) + dropped pairs t+u >= S+2:
    # sum_{d>=S+2} (d-1) 127^2 2^-7d <= 2^-7S * (S+1)/126 * 1.01
    per = 2.0 ** (-7 * S) * (2.0 + 2.0 ** (-7 * S) + (S + 1) / 126.0 * 1.01)
    return k * per * np.ldexp(1.0, ea)[:, None] * np.ldexp(1.0, eb)[None, :]


@pytest.mark.parametrize("phi", [0.5, 2.0])
def test_split_gemm_within_closed_form_bound(phi):
    rng = np.random.default_rng(int(phi * 10))
    m, n, k = 24, 20, 700
    A = _operand(rng, m, k, phi)
    B = _operand(rng, k, n, phi)
    a, b = A.ravel().copy(), B.ravel().copy()
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    hi, lo, ab = orc.dd_entries(0, 0, k, a, k, b, n, ii.ravel(), jj.ravel())
    exact = (hi + lo).reshape(m, n)
    ea = np.array([orc.split_exponent(A[i]) for i in range(m)])
    eb = np.array([orc.split_exponent(B[:, j].copy()) for j in range(n)])
    prev = None
    for S in range(3, 10):
        c = np.zeros(m * n)
        orc.split_gemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, S)
        err = np.abs(c.reshape(m, n) - exact)
        bound = _truncation_bound(S, k, ea, eb) + 4 * S * U * np.abs(exact) + 4 * U * ab.reshape(m, n)
        assert np.all(err <= bound), (S, float(np.max(err / bound)))
        mx = float(np.max(err))
        if prev is not None and prev > 1e-300:
            assert mx <= prev * 1.0001 or mx < 1e-15 * float(np.max(np.abs(exact)))
        prev = mx


def _rigorous_bound(S, bits, k, ea, eb):
    """Worst case of the slice forms' truncation, per unit of the row / column scales 2^ea, 2^eb: each
    scaled entry (|x| < 1) is its S digits plus a residual below 2^-(bits S); the kept digit pairs
    t + u <= S - 1 miss, for each t >= 1, digit t of one operand (below 2^bits) times the other's
    digits after the first S - t, whose sum is the remaining fraction minus the residual and so below
    2 * 2^-(bits (S - t)). Per term: 2^-(bits S) (2.01 + 2 (S - 1))."""
    per = 2.0 ** (-bits * S) * (2.01 + 2.0 * (S - 1))
    return k * per * np.ldexp(1.0, ea)[:, None] * np.ldexp(1.0, eb)[None, :]


@pytest.mark.parametrize("phi", [0.5, 2.0])
def test_bf16_digit_restatement_within_closed_form_bound(phi):
    """The BF16 form's restatement (8-bit digits, exact FP32 runs, INT32 flushes and FP64 flushes every
    chunk) stays inside the worst-case truncation bound for every S, across several FP64 flushes
    (k = 9000 > 2 chunks at S = 8); an error in the digit width, the pair set or the recombination weights
    can exceed it (one dropped diagonal costs up to 2^8 times the unit the bound is counted in)."""
    rng = np.random.default_rng(int(phi * 10) + 5)
    m, n, k = 16, 12, 9000
    A = _operand(rng, m, k, phi)
    B = _operand(rng, k, n, phi)
    a, b = A.ravel().copy(), B.ravel().copy()
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    hi, lo, ab = orc.dd_entries(0, 0, k, a, k, b, n, ii.ravel(), jj.ravel())
    exact = (hi + lo).reshape(m, n)
    ea = np.array([orc.split_exponent(A[i]) for i in range(m)])
    eb = np.array([orc.split_exponent(B[:, j].copy()) for j in range(n)])
    assert orc.split_bf16_chunk_k(8, k) < k // 2
    for S in range(3, 9):
        c = np.zeros(m * n)
        orc.split_gemm_bf16(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, S)
        err = np.abs(c.reshape(m, n) - exact)
        chunks = -(-k // orc.split_bf16_chunk_k(S, k))
        bound = _rigorous_bound(S, 8, k, ea, eb) + 4 * (S + chunks) * U * (np.abs(exact) + ab.reshape(m, n))
        assert np.all(err <= bound), (S, float(np.max(err / bound)))


def test_int8_digit_restatement_within_rigorous_bound():
    """The same worst-case bound with 7-bit digits, for the INT8 slice form's restatement."""
    rng = np.random.default_rng(77)
    m, n, k = 16, 12, 3000
    A = _operand(rng, m, k, 1.0)
    B = _operand(rng, k, n, 1.0)
    a, b = A.ravel().copy(), B.ravel().copy()
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    hi, lo, ab = orc.dd_entries(0, 0, k, a, k, b, n, ii.ravel(), jj.ravel())
    exact = (hi + lo).reshape(m, n)
    ea = np.array([orc.split_exponent(A[i]) for i in range(m)])
    eb = np.array([orc.split_exponent(B[:, j].copy()) for j in range(n)])
    for S in range(3, 10):
        c = np.zeros(m * n)
        orc.split_gemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, S)
        bound = _rigorous_bound(S, 7, k, ea, eb) + 4 * (S + 1) * U * (np.abs(exact) + ab.reshape(m, n))
        assert np.all(np.abs(c.reshape(m, n) - exact) <= bound), S


def _numpy_split(A, B, S, alpha, beta, C0, chunk_k):
    """Independent emulation: slices via numpy, diagonal sums by float64 matmul of
    small integers (exact: |sum| < 2^53), same recombination order."""
    def slices_rows(X):
        mx = np.max(np.abs(X), axis=1)
        e = np.array([np.frexp(v)[1] if v > 0 else 0 for v in mx])
        v = X * np.ldexp(1.0, -e)[:, None]
        out = []
        for _ in range(S):
            v = v * 128.0
            q = np.trunc(v)
            v = v - q
            out.append(q)
        return out, e
    As, ea = slices_rows(A)
    Bs, eb = slices_rows(B.T)
    Bs = [x.T for x in Bs]
    m, k = A.shape
    n = B.shape[1]
    C = C0.copy()
    first = True
    for k0 in range(0, k, chunk_k):
        k1 = min(k, k0 + chunk_k)
        acc = [np.zeros((m, n)) for _ in range(S)]
        for t in range(S):
            for u in range(S - t):
                acc[t + u] = acc[t + u] + As[t][:, k0:k1] @ Bs[u][k0:k1, :]
        s = np.zeros((m, n))
        for d in range(S - 1, -1, -1):
            s = s + acc[d] * 2.0 ** (-7 * (d + 2))
        v = s * np.ldexp(1.0, ea)[:, None] * np.ldexp(1.0, eb)[None, :]
        if first:
            C = alpha * v if beta == 0.0 else np.array([[np.fma(alpha, v[i, j], beta * C[i, j]) if hasattr(np, "fma")
                                                         else _fma(alpha, v[i, j], beta * C[i, j]) for j in range(n)]
                                                        for i in range(m)])
        else:
            C = np.array([[_fma(alpha, v[i, j], C[i, j]) for j in range(n)] for i in range(m)])
        first = False
    return C


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


@pytest.mark.parametrize("S,chunk", [(4, None), (7, None), (9, 64)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.5, 0.5)])
def test_c_restatement_matches_numpy_emulation(S, chunk, alpha, beta):
    rng = np.random.default_rng(S * 7 + int(beta * 4))
    m, n, k = 9, 7, 150
    A = _operand(rng, m, k, 1.0)
    B = _operand(rng, k, n, 1.0)
    C0 = rng.uniform(-1, 1, (m, n))
    ck = chunk or orc.split_chunk_k(S, k)
    c = C0.ravel().copy()
    orc.split_gemm(0, 0, m, n, k, alpha, A.ravel().copy(), k, B.ravel().copy(), n, beta, c, n, S, ck)
    ref = _numpy_split(A, B, S, alpha, beta, C0, ck)
    assert np.array_equal(c.reshape(m, n), ref)


def test_split_layouts_agree_bitwise():
    rng = np.random.default_rng(77)
    m, n, k = 11, 13, 90
    A = _operand(rng, m, k, 1.0)
    B = _operand(rng, k, n, 1.0)
    outs = []
    for la in (0, 1):
        for lb in (0, 1):
            a = A.ravel().copy() if la == 0 else A.T.copy().ravel()
            b = B.ravel().copy() if lb == 0 else B.T.copy().ravel()
            c = np.zeros(m * n)
            orc.split_gemm(la, lb, m, n, k, 1.0, a, k if la == 0 else m, b, n if lb == 0 else k, 0.0, c, n, 6)
            outs.append(c)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_chunk_rule_keeps_int32_exact():
    for S in range(3, 10):
        ck = orc.split_chunk_k(S, 1 << 20)
        assert ck % 32 == 0
        assert S * ck * 127 * 127 < 2 ** 31
        assert S * (ck + 32) * 127 * 127 >= 2 ** 31
    assert orc.split_chunk_k(8, 16384) == 16384
    assert orc.split_chunk_k(9, 16384) < 16384


def test_nonfinite_rows_give_nan():
    A = np.ones((3, 4))
    A[1, 2] = np.inf
    B = np.ones((4, 2))
    c = np.zeros(6)
    orc.split_gemm(0, 0, 3, 2, 4, 1.0, A.ravel().copy(), 4, B.ravel().copy(), 2, 0.0, c, 2, 5)
    c = c.reshape(3, 2)
    assert np.all(np.isnan(c[1])) and np.all(c[[0, 2]] == 4.0)


def test_bf16_restatement_matches_numpy_emulation():
    """8-bit digits: the C restatement equals an independent numpy emulation."""
    rng = np.random.default_rng(31)
    m, n, k, S = 8, 6, 120, 5
    A = _operand(rng, m, k, 1.0)
    B = _operand(rng, k, n, 1.0)
    ck = orc.split_bf16_chunk_k(S, k)

    def slices(X):
        mx = np.max(np.abs(X), axis=1)
        e = np.array([np.frexp(v)[1] if v > 0 else 0 for v in mx])
        v = X * np.ldexp(1.0, -e)[:, None]
        out = []
        for _ in range(S):
            v = v * 256.0
            q = np.trunc(v)
            v = v - q
            out.append(q)
        return out, e
    As, ea = slices(A)
    Bs, eb = slices(B.T)
    Bs = [x.T for x in Bs]
    acc = [np.zeros((m, n)) for _ in range(S)]
    for t in range(S):
        for u in range(S - t):
            acc[t + u] = acc[t + u] + As[t] @ Bs[u]
    s = np.zeros((m, n))
    for d in range(S - 1, -1, -1):
        s = s + acc[d] * 2.0 ** (-8 * (d + 2))
    ref = s * np.ldexp(1.0, ea)[:, None] * np.ldexp(1.0, eb)[None, :]
    c = np.zeros(m * n)
    orc.split_gemm_bf16(0, 0, m, n, k, 1.0, A.ravel().copy(), k, B.ravel().copy(), n, 0.0, c, n, S, ck)
    assert ck >= k
    assert np.array_equal(c.reshape(m, n), ref)
