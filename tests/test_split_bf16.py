"""BF16-slice split form: GPU bit-exactness against the oracle restatement
(8-bit digits) and accuracy against the double-double oracle."""
import numpy as np
import pytest

from jlc_amd import native
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _operand(rng, r, c, phi):
    return (rng.random((r, c)) - 0.5) * np.exp(phi * rng.standard_normal((r, c)))


def _case(rng, la, lb, m, n, k, S, alpha, beta, pad=0, off=0):
    A = _operand(rng, m, k, 1.0)
    B = _operand(rng, k, n, 1.0)
    lda = (k if la == 0 else m) + pad
    ldb = (n if lb == 0 else k) + pad
    ldc = n + pad
    a = np.zeros(off + lda * (m if la == 0 else k))
    b = np.zeros(off + ldb * (k if lb == 0 else n))
    for i in range(m):
        if la == 0:
            a[off + i * lda: off + i * lda + k] = A[i]
        else:
            a[off + i: off + i + lda * k: lda] = A[i]
    for p in range(k):
        if lb == 0:
            b[off + p * ldb: off + p * ldb + n] = B[p]
        else:
            b[off + p: off + p + ldb * n: ldb] = B[p]
    c = rng.uniform(-1, 1, off + ldc * m)
    exp = c.copy()
    orc.split_gemm_bf16(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, exp, ldc, S, a_off=off, b_off=off, c_off=off)
    st = native.dgemm(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                      native.mode_slices(native.MODE_SPLIT_BF16, S), off, off, off)
    assert st.kernel_name == "split_bf16" and st.slices == S
    same = (c == exp) | (np.isnan(c) & np.isnan(exp))
    assert same.all(), (la, lb, m, n, k, S, int((~same).sum()), float(np.nanmax(np.abs(c - exp))))


@pytest.mark.parametrize("S", [3, 4, 5, 6, 7, 8])
def test_bf16_bitwise_equal_to_restatement(S):
    rng = np.random.default_rng(200 + S)
    _case(rng, 0, 0, 64, 64, 48, S, 1.0, 0.0)
    _case(rng, 0, 0, 150, 97, 130, S, -1.5, 0.5)


@pytest.mark.parametrize("la,lb", [(0, 1), (1, 0), (1, 1)])
def test_bf16_layouts(la, lb):
    _case(np.random.default_rng(la * 2 + lb), la, lb, 77, 65, 41, 5, 2.5, -0.5, pad=1, off=3)


@pytest.mark.parametrize("S", [7, 8])
def test_bf16_fp64_flush_chunks(S):
    # S = 8: exact FP32 runs of 2 steps (2^24 / (8 * 255^2 * 16) = 2.02), INT32 flushes every 4128 k
    assert orc.split_bf16_chunk_k(S, 5000) < 5000
    _case(np.random.default_rng(3 + S), 0, 0, 40, 36, 5000, S, 1.0, 0.25)


def test_bf16_accuracy_matches_int8_at_equal_bits():
    """7 BF16 slices carry 56 bits like 8 INT8 slices: errors of the same order."""
    rng = np.random.default_rng(8)
    m = n = k = 512
    A = _operand(rng, m, k, 0.5)
    B = _operand(rng, k, n, 0.5)
    a, b = A.ravel().copy(), B.ravel().copy()
    ii, jj = np.meshgrid(np.arange(0, m, 7), np.arange(0, n, 5), indexing="ij")
    hi, lo, _ = orc.dd_entries(0, 0, k, a, k, b, n, ii.ravel(), jj.ravel())
    exact = hi + lo
    errs = {}
    for name, mode in (("bf16x7", native.mode_slices(native.MODE_SPLIT_BF16, 7)),
                       ("int8x8", native.mode_slices(native.MODE_SPLIT_INT8, 8))):
        c = np.zeros(m * n)
        native.dgemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, mode)
        got = c.reshape(m, n)[ii, jj].ravel()
        errs[name] = float(np.max(np.abs(got - exact) / np.abs(exact)))
    print(errs, flush=True)
    # same digit bits, same truncation bound: errors within an order of magnitude of each other
    assert errs["bf16x7"] < 1e-9 and errs["int8x8"] < 1e-9
    assert max(errs.values()) / min(errs.values()) < 10.0
