"""GPU parity of the CRT split form (JLC_MODE_SPLIT_CRT).

Residues, integer products and the CRT reconstruction are exact integer
work and the final conversion follows a fixed FP64 sequence, so the GPU
result must equal the oracle's restatement (oracle_crt_gemm) bit for bit.
Accuracy is then stated against the double-double oracle.
"""
import numpy as np
import pytest

from jlc_amd import native
from oracle import oracle as orc
from tests.fp64_ref import unsplit_fp64

pytestmark = pytest.mark.gpu


def _operand(rng, r, c, phi):
    return (rng.random((r, c)) - 0.5) * np.exp(phi * rng.standard_normal((r, c)))


def _store(X, layout, ld, off, pad_rows):
    r, c = X.shape
    buf = np.zeros(off + ld * (r if layout == 0 else c))
    if layout == 0:
        for i in range(r):
            buf[off + i * ld: off + i * ld + c] = X[i]
    else:
        for j in range(c):
            buf[off + j * ld: off + j * ld + r] = X[:, j]
    return buf


def _case(rng, la, lb, m, n, k, nmod, alpha, beta, phi=1.0, pad=0, off=0, plant=None):
    A = _operand(rng, m, k, phi)
    B = _operand(rng, k, n, phi)
    if plant is not None:
        plant(A, B)
    lda = (k if la == 0 else m) + pad
    ldb = (n if lb == 0 else k) + pad
    ldc = n + pad
    a = _store(A, la, lda, off, pad)
    b = _store(B, lb, ldb, off, pad)
    c = rng.uniform(-1, 1, off + ldc * m)
    c0 = c.copy()
    st = native.dgemm(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                      native.mode_slices(native.MODE_SPLIT_CRT, nmod), off, off, off)
    assert st.kernel_name == "split_crt", st.kernel_name
    if nmod:
        assert st.slices == nmod and (st.bits_a, st.bits_b) == orc.crt_bits_for(nmod, k)
    exp = c0.copy()
    orc.crt_gemm(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, exp, ldc, st.slices, st.bits_a, st.bits_b,
                 a_off=off, b_off=off, c_off=off)
    same = (c == exp) | (np.isnan(c) & np.isnan(exp))
    assert same.all(), (la, lb, m, n, k, nmod, int((~same).sum()), float(np.nanmax(np.abs(c - exp))))
    return st


@pytest.mark.parametrize("nmod", [4, 5, 6, 8, 10, 11, 12, 15, 16, 20])
def test_bitwise_equal_to_restatement(nmod):
    rng = np.random.default_rng(300 + nmod)
    _case(rng, 0, 0, 256, 256, 64, nmod, 1.0, 0.0)
    _case(rng, 0, 0, 200, 130, 97, nmod, -1.5, 0.5)


@pytest.mark.parametrize("la,lb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_all_layouts_ragged(la, lb):
    rng = np.random.default_rng(20 * la + lb)
    for m, n, k in [(1, 1, 1), (33, 17, 5), (257, 65, 129), (300, 250, 200)]:
        _case(rng, la, lb, m, n, k, 16, 2.5, -0.5, pad=1, off=3)


def test_automatic_moduli_count():
    rng = np.random.default_rng(8)
    st = _case(rng, 0, 0, 300, 280, 1000, 0, 1.0, 0.0, phi=0.5)
    assert 2 <= st.slices <= 20 and st.bits_a >= 40 and st.bits_b >= 40
    # the fewest moduli that hold (2^pa + sqrt(k)/2)(2^pb + sqrt(k)/2) with the M/4 margin
    assert orc.crt_fits(st.bits_a, st.bits_b, 1000, st.slices)
    assert not orc.crt_fits(st.bits_a, st.bits_b, 1000, st.slices - 1)
    # and the moduli's leftover capacity is spent on bits: one more for the narrower operand no longer fits
    more = (st.bits_a + 1, st.bits_b) if st.bits_a <= st.bits_b else (st.bits_a, st.bits_b + 1)
    assert not orc.crt_fits(*more, 1000, st.slices)


def test_modular_fold_beyond_65536_k():
    # INT32 sums are folded modulo m every 65536 k; 70000 crosses one fold.
    rng = np.random.default_rng(70)
    _case(rng, 0, 0, 24, 20, 70000, 16, 1.0, 0.0, phi=0.5)


def _plant_zero_and_wide(A, B):
    A[3] = 0.0                  # all-zero row: exponent 0, norm exponent 0
    A[4, ::7] *= 2.0 ** 300     # a row whose norm sits far above most entries
    A[5, 1] = 2.0 ** -1000      # a lone subnormal-range entry
    B[:, 2] = 0.0               # all-zero column


@pytest.mark.parametrize("la,lb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_long_rows_one_read_statistics(la, lb):
    # Depths in [2048, 16384] take the one-read statistics and norm kernel on
    # k-contiguous operands (row-major op(A), column-major op(B)); 2047 and
    # 16385 the two-pass kernels. Exponents and norm exponents come from a max
    # and integer sums, so every path must give the restatement's bits.
    rng = np.random.default_rng(40 + 2 * la + lb)
    for k in (2047, 2048, 3001, 16384, 16385):
        _case(rng, la, lb, 21, 13, k, 12, 1.0, 0.0, phi=2.0, pad=1, off=1, plant=_plant_zero_and_wide)


def test_long_nonfinite_rows_are_nan():
    rng = np.random.default_rng(41)
    m, n, k = 9, 7, 4096
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    A[2, 4000] = np.inf
    A[6, 17] = np.nan
    for lb in (0, 1):
        b = B.ravel().copy() if lb == 0 else B.T.ravel().copy()
        c = np.zeros(m * n)
        native.dgemm(0, lb, m, n, k, 1.0, A.ravel().copy(), k, b, n if lb == 0 else k, 0.0, c, n,
                     native.mode_slices(native.MODE_SPLIT_CRT, 14))
        c = c.reshape(m, n)
        assert np.all(np.isnan(c[[2, 6]]))
        assert np.all(np.isfinite(np.delete(c, [2, 6], axis=0)))


def test_nonfinite_rows_and_columns_are_nan():
    rng = np.random.default_rng(4)
    m, n, k = 70, 66, 40
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    A[5, 3] = np.inf
    B[7, 9] = np.nan
    c = np.zeros(m * n)
    native.dgemm(0, 0, m, n, k, 1.0, A.ravel().copy(), k, B.ravel().copy(), n, 0.0, c, n,
                 native.mode_slices(native.MODE_SPLIT_CRT, 14))
    c = c.reshape(m, n)
    assert np.all(np.isnan(c[5])) and np.all(np.isnan(c[:, 9]))
    ok = np.ones((m, n), bool)
    ok[5] = False
    ok[:, 9] = False
    assert np.all(np.isfinite(c[ok]))


def test_too_wide_spread_runs_fp64_kernel():
    # Entries spanning ~2^200 within a row need more bits than 20 moduli hold.
    rng = np.random.default_rng(5)
    m, n, k = 64, 64, 64
    A = rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-100, 100, (m, k)))
    B = rng.uniform(-1, 1, (k, n))
    c = np.zeros(m * n)
    st = native.dgemm(0, 0, m, n, k, 1.0, A.ravel().copy(), k, B.ravel().copy(), n, 0.0, c, n,
                      native.MODE_SPLIT_CRT)
    assert st.kernel_name == "mfma_f64"
    c2 = np.zeros(m * n)
    native.dgemm(0, 0, m, n, k, 1.0, A.ravel().copy(), k, B.ravel().copy(), n, 0.0, c2, n, native.MODE_FP64)
    assert np.array_equal(c, c2)


def test_invalid_moduli_count_rejected():
    a = np.ones(4)
    c = np.zeros(4)
    with pytest.raises(native.JlcError):
        native.dgemm(0, 0, 2, 2, 2, 1.0, a, 2, a, 2, 0.0, c, 2, native.mode_slices(native.MODE_SPLIT_CRT, 21))
    with pytest.raises(native.JlcError):
        native.dgemm(0, 0, 2, 2, 2, 1.0, a, 2, a, 2, 0.0, c, 2, native.mode_slices(native.MODE_SPLIT_CRT, 1))


def test_accuracy_matches_fp64_at_auto_moduli():
    """Against double-double sums: the automatic choice keeps the max relative
    and componentwise errors within those of the FP64 MFMA kernel (factor 2).
    The FP64 reference is the unsplit kernel, the one the 16384^3 comparison
    runs: at 512 x 512 x 2048 the tuned split-K policy would cut K into 128-deep
    slices whose separate partial sums make the FP64 result more accurate than
    that kernel's."""
    rng = np.random.default_rng(12)
    m = n = 512
    k = 2048
    A = _operand(rng, m, k, 0.5)
    B = _operand(rng, k, n, 0.5)
    a, b = A.ravel().copy(), B.ravel().copy()
    ii = rng.integers(0, m, 400)
    jj = rng.integers(0, n, 400)
    hi, lo, ab = orc.dd_entries(0, 0, k, a, k, b, n, ii, jj)
    errs = {}
    for name, mode in (("fp64", native.MODE_FP64), ("crt", native.MODE_SPLIT_CRT)):
        c = np.zeros(m * n)
        with unsplit_fp64():
            native.dgemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, mode)
        got = c.reshape(m, n)[ii, jj]
        err = np.abs(got - hi - lo)
        errs[name] = (float(np.max(err / np.abs(hi + lo))), float(np.max(err / ab)))
    assert errs["crt"][0] <= 2 * errs["fp64"][0] + 1e-16, errs
    assert errs["crt"][1] <= 2 * errs["fp64"][1] + 2 ** -53, errs


def test_four_wave_product_kernel():
    """The 4-wavefront (128 x 128 per wavefront) product kernel, selected by
    the crt_waves tuning key, gives the same bits, across a modular fold too."""
    old = native.tune_get("crt_waves")
    native.tune_set("crt_waves", 4)
    try:
        rng = np.random.default_rng(44)
        _case(rng, 0, 0, 300, 270, 130, 16, 1.0, 0.0)
        _case(rng, 1, 1, 257, 65, 129, 9, -1.5, 0.5, pad=1, off=3)
        _case(rng, 0, 0, 24, 20, 70000, 16, 1.0, 0.0, phi=0.5)
    finally:
        native.tune_set("crt_waves", old)


def test_depth_beyond_norm_sum_limit():
    """k >= 2^22 is beyond the integer norm sums: automatic mode hands the call
    to the FP64 kernel, an explicit moduli count is rejected."""
    m = n = 2
    k = 1 << 22
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, m * k)
    b = rng.uniform(-1, 1, k * n)
    c = np.zeros(m * n)
    st = native.dgemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, native.MODE_SPLIT_CRT)
    assert st.kernel_name == "mfma_f64"
    with pytest.raises(native.JlcError):
        native.dgemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, native.mode_slices(native.MODE_SPLIT_CRT, 16))


def test_extreme_magnitudes_norm_scaling():
    """Rows and columns near the ends of the FP64 range, mixed within a row, subnormal
    entries, all-zero rows and a lone non-zero: the norm exponents (integer sums of
    scaled squares) and the whole product stay bit-identical to the restatement."""
    rng = np.random.default_rng(77)
    m, n, k = 70, 66, 45
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    A[0] *= 2.0 ** 900
    A[1] *= 2.0 ** -900
    A[2, :] = 0.0
    A[3, :] = 0.0
    A[3, 7] = -3.5
    A[4, ::2] *= 2.0 ** 600      # one row spanning ~2^1200: its small half underflows in the scaled squares
    A[5, :5] = 5e-324            # subnormals
    A[5, 5:] = 0.0
    B[:, 0] *= 2.0 ** -1000
    B[:, 1] *= 2.0 ** 1000
    B[:, 2] = 0.0
    a, b = A.ravel().copy(), B.ravel().copy()
    for nmod in (8, 16, 20):
        c = np.zeros(m * n)
        st = native.dgemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, c, n, native.mode_slices(native.MODE_SPLIT_CRT, nmod))
        assert st.kernel_name == "split_crt"
        exp = np.zeros(m * n)
        orc.crt_gemm(0, 0, m, n, k, 1.0, a, k, b, n, 0.0, exp, n, st.slices, st.bits_a, st.bits_b)
        same = (c == exp) | (np.isnan(c) & np.isnan(exp))
        assert same.all(), (nmod, int((~same).sum()))
        cm = c.reshape(m, n)
        assert np.all(cm[2] == 0.0) and np.all(cm[:, 2] == 0.0)


@pytest.mark.parametrize("mode", [native.MODE_SPLIT_CRT, native.MODE_SPLIT_INT8, native.MODE_SPLIT_BF16])
def test_beta_zero_never_reads_c(mode):
    """BLAS semantics kept by the split forms (GemmNumericEdgeTest's beta = 0 case): with beta = 0 the
    old C is not read, so NaN or Inf in it cannot reach the result."""
    rng = np.random.default_rng(12)
    m, n, k = 65, 47, 33
    A = rng.uniform(-1, 1, m * k)
    B = rng.uniform(-1, 1, k * n)
    c = np.full(m * n, np.nan)
    c[::7] = np.inf
    native.dgemm(0, 0, m, n, k, 1.25, A, k, B, n, 0.0, c, n, mode)
    assert np.all(np.isfinite(c))
    ref = 1.25 * (A.reshape(m, k) @ B.reshape(k, n))
    assert np.max(np.abs(c.reshape(m, n) - ref)) <= 1e-12 * np.max(np.abs(ref))
