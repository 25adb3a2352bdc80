"""GPU parity of the split-precision INT8 form.

The split form is deterministic integer arithmetic plus a fixed FP64
recombination order, so the GPU result must equal the oracle's restatement
(oracle_split_gemm) bit for bit; its accuracy is then stated against the
double-double oracle.
"""
import numpy as np
import pytest

from jlc_amd import native
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _operand(rng, r, c, phi):
    return (rng.random((r, c)) - 0.5) * np.exp(phi * rng.standard_normal((r, c)))


def _case(rng, la, lb, m, n, k, S, alpha, beta, phi=1.0, pad=0, off=0):
    A = _operand(rng, m, k, phi)
    B = _operand(rng, k, n, phi)
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
    orc.split_gemm(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, exp, ldc, S, a_off=off, b_off=off, c_off=off)
    st = native.dgemm(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                      native.mode_slices(native.MODE_SPLIT_INT8, S), off, off, off)
    assert st.kernel_name == "split_int8" and st.slices == S
    same = (c == exp) | (np.isnan(c) & np.isnan(exp))
    assert same.all(), (la, lb, m, n, k, S, int((~same).sum()), float(np.nanmax(np.abs(c - exp))))


@pytest.mark.parametrize("S", [3, 5, 7, 8, 9])
def test_bitwise_equal_to_restatement(S):
    rng = np.random.default_rng(100 + S)
    _case(rng, 0, 0, 128, 64, 64, S, 1.0, 0.0)
    _case(rng, 0, 0, 200, 130, 97, S, -1.5, 0.5)


@pytest.mark.parametrize("la,lb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_all_layouts_ragged(la, lb):
    rng = np.random.default_rng(10 * la + lb)
    for m, n, k in [(1, 1, 1), (33, 17, 5), (129, 65, 33), (300, 250, 200)]:
        _case(rng, la, lb, m, n, k, 6, 2.5, -0.5, pad=1, off=3)


def test_k_chunking_for_nine_slices():
    # S = 9 needs a diagonal flush every 14784 k to keep INT32 exact.
    rng = np.random.default_rng(9)
    assert orc.split_chunk_k(9, 15000) < 15000
    _case(rng, 0, 0, 40, 36, 15000, 9, 1.0, 0.25)


def test_nonfinite_rows_and_columns_are_nan():
    rng = np.random.default_rng(4)
    m, n, k = 70, 66, 40
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    A[5, 3] = np.inf
    B[7, 9] = np.nan
    c = np.zeros(m * n)
    native.dgemm(0, 0, m, n, k, 1.0, A.ravel().copy(), k, B.ravel().copy(), n, 0.0, c, n,
                 native.mode_slices(native.MODE_SPLIT_INT8, 7))
    c = c.reshape(m, n)
    assert np.all(np.isnan(c[5])) and np.all(np.isnan(c[:, 9]))
    mask = np.ones((m, n), bool)
    mask[5] = False
    mask[:, 9] = False
    assert np.all(np.isfinite(c[mask]))


def test_auto_slices_and_accuracy_vs_fp64():
    """phi = 0.5 inputs: the chooser picks enough slices to match FP64's error."""
    rng = np.random.default_rng(2024)
    m = n = k = 2048
    A = _operand(rng, m, k, 0.5)
    B = _operand(rng, k, n, 0.5)
    a, b = A.ravel().copy(), B.ravel().copy()
    da, db = native.DeviceArray.from_numpy(a), native.DeviceArray.from_numpy(b)
    dc = native.DeviceArray(m * n)
    st = native.dgemm_dev(0, 0, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, dc.ptr, n, native.MODE_SPLIT_INT8)
    print(f"auto slices: {st.slices}", flush=True)
    assert 8 <= st.slices <= 9
    c_split = dc.to_numpy()
    native.dgemm_dev(0, 0, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, dc.ptr, n, native.MODE_FP64)
    c_fp64 = dc.to_numpy()
    srng = np.random.default_rng(5)
    ii = srng.integers(0, m, 2048)
    jj = srng.integers(0, n, 2048)
    hi, lo, ab = orc.dd_entries(0, 0, k, a, k, b, n, ii, jj)
    exact = hi + lo
    idx = ii * n + jj
    rel_split = float(np.max(np.abs(c_split[idx] - exact) / np.abs(exact)))
    rel_fp64 = float(np.max(np.abs(c_fp64[idx] - exact) / np.abs(exact)))
    print(f"max rel err: split {rel_split:.3e}  fp64 {rel_fp64:.3e}", flush=True)
    assert rel_split <= 2.0 * rel_fp64


def test_invalid_slice_counts_rejected():
    a = np.ones(4)
    for s in (1, 2, 10):
        with pytest.raises(native.JlcError) as ei:
            native.dgemm(0, 0, 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a.copy(), 2,
                         native.mode_slices(native.MODE_SPLIT_INT8, s))
        assert ei.value.status == native.JLC_ERR_INVALID_ARG


@pytest.mark.parametrize("form,name", [(native.MODE_SPLIT_INT8, "split_int8"), (native.MODE_SPLIT_BF16, "split_bf16")])
def test_automatic_slices_beyond_maximum_run_fp64_kernel(form, name):
    """Rows spanning ~2^60 need more slices than either form has: the automatic
    mode hands the call to the FP64 kernel, whose result it then equals."""
    rng = np.random.default_rng(31)
    m, n, k = 96, 80, 72
    A = rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-30, 30, (m, k)))
    B = rng.uniform(-1, 1, (k, n)) * np.exp2(rng.integers(-30, 30, (k, n)))
    C0 = rng.uniform(-1, 1, m * n)
    c = C0.copy()
    st = native.dgemm(0, 0, m, n, k, 1.5, A.ravel().copy(), k, B.ravel().copy(), n, 0.25, c, n, form)
    assert st.kernel_name == "mfma_f64"
    c2 = C0.copy()
    native.dgemm(0, 0, m, n, k, 1.5, A.ravel().copy(), k, B.ravel().copy(), n, 0.25, c2, n, native.MODE_FP64)
    assert np.array_equal(c, c2)
    # a narrow spread stays on the slice form
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    st = native.dgemm(0, 0, m, n, k, 1.0, A.ravel().copy(), k, B.ravel().copy(), n, 0.0, np.zeros(m * n), n, form)
    assert st.kernel_name == name
