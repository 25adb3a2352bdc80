"""Pins of the CRT split-form restatement (oracle/oracle.c, oracle_crt_*).

The CRT form has no reference counterpart (it is the split-precision
component of the DGEMM engine); its oracle is pinned here by what the
mathematics fixes: exact Python integers for residues, reconstruction and
the whole product of the integer images, on small shapes.
"""
from __future__ import annotations

import math
import random
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc


def sym(v: int, m: int) -> int:
    r = v % m
    return r - m if r > (m - 1) // 2 else r


def test_moduli_pairwise_coprime_and_int8_range():
    mods = orc.crt_moduli()
    assert len(mods) == orc.CRT_MAX_MODULI == 20
    assert all(2 < m <= 256 for m in mods)
    for i in range(20):
        for j in range(i + 1, 20):
            assert math.gcd(mods[i], mods[j]) == 1
    for m in mods:  # symmetric residue range fits int8
        assert -128 <= -(m // 2) and (m - 1) // 2 <= 127
    for n in (1, 8, 16, 20):
        assert orc.crt_log2_product(n) == pytest.approx(sum(math.log2(m) for m in mods[:n]), abs=1e-12)


def test_tables_match_big_integer_definition():
    mods = orc.crt_moduli()
    for n in (2, 9, 16, 20):
        w, ml, fi = orc.crt_tables(n)
        M = math.prod(mods[:n])
        assert sum(int(ml[j]) << (40 * j) for j in range(4)) == M
        for t in range(n):
            Mi = M // mods[t]
            W = Mi * pow(Mi, -1, mods[t])
            assert sum(int(w[t, j]) << (40 * j) for j in range(4)) == W
            assert all(0 <= w[t, j] < 2 ** 40 for j in range(4))
            assert W % mods[t] == 1 and all(W % mods[u] == 0 for u in range(n) if u != t)


def test_residues_match_python_integers():
    rng = np.random.default_rng(3)
    x = (rng.random(257) - 0.5) * np.exp(2.0 * rng.standard_normal(257))
    x[:3] = [0.0, -0.0, x[5]]
    e = math.frexp(float(np.max(np.abs(x))))[1]
    for p in (12, 53, 54, 64, 70):
        got = orc.crt_residues(x, e, p, 20)
        for q, v in enumerate(x):
            s = math.ldexp(float(v), p - e)
            img = int(round(s)) if abs(s) < 2 ** 53 else int(s)
            assert abs(img) <= 2 ** p
            for t, m in enumerate(orc.crt_moduli()):
                assert got[t, q] == sym(img, m), (p, q, t)


def test_reconstruction_is_correctly_rounded():
    mods = orc.crt_moduli()
    rnd = random.Random(11)
    for n in (4, 8, 13, 16, 20):
        M = math.prod(mods[:n])
        for _ in range(400):
            v = rnd.getrandbits(rnd.randint(1, int(math.log2(M)) - 2))
            v = -v if rnd.random() < 0.5 else v
            got = orc.crt_reconstruct([sym(v, m) for m in mods[:n]], n)
            assert got == float(v), (n, v, got)
    assert orc.crt_reconstruct([0] * 16, 16) == 0.0


def _norm_sq(x) -> Fraction:
    return sum((Fraction(float(v)) ** 2 for v in x), Fraction(0))


def test_norm_exponent_known_answers():
    assert orc.crt_norm_exponent(np.ones(4)) == 2          # ||x|| = 2; the +1 units make exact powers round up
    assert orc.crt_norm_exponent(np.array([0.75])) == 0    # 0.5 < 0.75 <= 1
    assert orc.crt_norm_exponent(np.array([-3.0, 4.0])) == 3  # ||x|| = 5
    assert orc.crt_norm_exponent(np.zeros(7)) == 0
    assert orc.crt_norm_exponent(np.array([1.0, np.nan])) == orc.SPLIT_NONFINITE
    assert orc.crt_norm_exponent(np.array([np.inf, 0.0])) == orc.SPLIT_NONFINITE
    # tiny and huge scales: no overflow or underflow in the bound
    assert orc.crt_norm_exponent(np.array([3e-300, 4e-300])) == math.frexp(5e-300)[1]
    assert orc.crt_norm_exponent(np.array([3e300, 4e300])) == math.frexp(5e300)[1]


@pytest.mark.parametrize("phi", [0.0, 0.5, 2.0, 8.0])
def test_norm_exponent_bounds_the_norm_tightly(phi):
    """||x|| <= 2^t exactly, and 2^(t-1) lies below the norm plus the bound's
    per-entry slack (2 units of 2^(2e-40) per non-zero entry)."""
    rng = np.random.default_rng(int(phi * 10) + 1)
    for size in (1, 2, 31, 500):
        for _ in range(20):
            x = (rng.random(size) - 0.5) * np.exp(phi * rng.standard_normal(size))
            x[rng.random(size) < 0.1] = 0.0
            t = orc.crt_norm_exponent(x)
            nsq = _norm_sq(x)
            if nsq == 0:
                assert t == 0
                continue
            e = math.frexp(float(np.max(np.abs(x))))[1]
            assert nsq <= Fraction(2) ** (2 * t), (size, t)
            slack = Fraction(2) * Fraction(2) ** (2 * e - 40) * int(np.count_nonzero(x))
            assert Fraction(2) ** (2 * (t - 1)) < nsq + slack, (size, t)
            # integer sums: order and scale do not matter
            assert orc.crt_norm_exponent(rng.permutation(x)) == t
            assert orc.crt_norm_exponent(np.ldexp(x, 7)) == t + 7


def _exact(A, B, pa, pb, alpha=1.0):
    m, k = A.shape
    n = B.shape[1]
    ea = [orc.crt_norm_exponent(A[i]) for i in range(m)]
    eb = [orc.crt_norm_exponent(B[:, j]) for j in range(n)]

    def img(x, e, p):
        s = math.ldexp(float(x), p - e)
        return int(round(s)) if abs(s) < 2 ** 53 else int(s)
    Ai = [[img(A[i, q], ea[i], pa) for q in range(k)] for i in range(m)]
    Bi = [[img(B[q, j], eb[j], pb) for j in range(n)] for q in range(k)]
    # norm scaling: ||a'|| <= 2^pa + sqrt(k)/2 (squared: (2^pa + sqrt(k)/2)^2 <= 4^pa + 2^pa sqrt(k) + k)
    for i in range(m):
        assert sum(v * v for v in Ai[i]) <= 4 ** pa + 2 ** pa * (math.isqrt(k) + 1) + k
    for j in range(n):
        assert sum(Bi[q][j] ** 2 for q in range(k)) <= 4 ** pb + 2 ** pb * (math.isqrt(k) + 1) + k
    C = np.zeros((m, n))
    for i in range(m):
        for j in range(n):
            cp = sum(Ai[i][q] * Bi[q][j] for q in range(k))
            C[i, j] = alpha * math.ldexp(float(cp), ea[i] + eb[j] - pa - pb)
    return C


def _nmod_for(pa, pb, k):
    """Fewest moduli whose product exceeds 4 (2^pa + sqrt(k)/2)(2^pb + sqrt(k)/2):
    by Cauchy-Schwarz every integer product of norm-scaled images is below a quarter of it."""
    h = Fraction(math.isqrt(4 * k) + 1, 4)  # >= sqrt(k)/2
    bound = 4 * (2 ** pa + h) * (2 ** pb + h)
    nmod = next(t for t in range(2, 21) if math.prod(orc.crt_moduli(t)) > bound)
    assert orc.crt_fits(pa, pb, k, nmod)
    return nmod


@pytest.mark.parametrize("m,n,k,pa,pb,phi", [(7, 5, 33, 20, 20, 0.5), (9, 6, 70, 53, 53, 0.5),
                                             (5, 8, 100, 61, 60, 2.0), (4, 4, 300, 64, 64, 1.0),
                                             (3, 2, 1, 30, 10, 0.0)])
def test_crt_gemm_equals_exact_integer_product(m, n, k, pa, pb, phi):
    rng = np.random.default_rng(m * 100 + k)
    A = (rng.random((m, k)) - 0.5) * np.exp(phi * rng.standard_normal((m, k)))
    B = (rng.random((k, n)) - 0.5) * np.exp(phi * rng.standard_normal((k, n)))
    A[0] = 0.0
    nmod = _nmod_for(pa, pb, k)
    C = np.zeros(m * n)
    orc.crt_gemm(0, 0, m, n, k, 1.0, np.ascontiguousarray(A).ravel(), k, np.ascontiguousarray(B).ravel(), n, 0.0, C,
                 n, nmod, pa, pb)
    A0 = A.copy()
    A0[0] = 1.0  # an all-zero row has images 0 whatever its exponent: its outputs are 0
    expect = _exact(A0, B, pa, pb)
    expect[0] = 0.0
    assert np.array_equal(C.reshape(m, n), expect)


def test_crt_gemm_layouts_alpha_beta_and_nonfinite():
    rng = np.random.default_rng(9)
    m, n, k, pa, pb = 6, 7, 40, 40, 40
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    nmod = _nmod_for(pa, pb, k)
    base = np.zeros(m * n)
    orc.crt_gemm(0, 0, m, n, k, 1.0, A.ravel(), k, B.ravel(), n, 0.0, base, n, nmod, pa, pb)
    # column-major op(A) / op(B) storage gives the same numbers
    for la, lb in ((1, 0), (0, 1), (1, 1)):
        a = (A.T if la else A).copy().ravel()
        b = (B.T if lb else B).copy().ravel()
        c = np.zeros(m * n)
        orc.crt_gemm(la, lb, m, n, k, 1.0, a, m if la else k, b, k if lb else n, 0.0, c, n, nmod, pa, pb)
        assert np.array_equal(c, base)
    c0 = rng.uniform(-1, 1, m * n)
    c = c0.copy()
    orc.crt_gemm(0, 0, m, n, k, -1.5, A.ravel(), k, B.ravel(), n, 0.25, c, n, nmod, pa, pb)
    # fma(alpha, v, beta*c): one rounding of the exact alpha*v + fl(beta*c)
    want = [float(Fraction(-1.5) * Fraction(v) + Fraction(0.25 * w)) for v, w in zip(base, c0)]
    assert np.array_equal(c, np.array(want))
    A2 = A.copy()
    A2[2, 5] = np.nan
    B2 = B.copy()
    B2[3, 4] = np.inf
    c = np.zeros(m * n)
    orc.crt_gemm(0, 0, m, n, k, 1.0, A2.ravel(), k, B2.ravel(), n, 0.0, c, n, nmod, pa, pb)
    c = c.reshape(m, n)
    assert np.all(np.isnan(c[2])) and np.all(np.isnan(c[:, 4]))
    mask = np.ones((m, n), bool)
    mask[2] = False
    mask[:, 4] = False
    assert np.array_equal(c[mask], base.reshape(m, n)[mask])
