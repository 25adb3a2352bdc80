"""GPU parity of the HIP DGEMM path against the oracle, through the host mirror
of the reference's API and through the C ABI directly.

The first block restates the reference's GEMM parity tests case by case
(src/test/java/net/faulj/compute/GemmParityTest.java, GemmStridedParityTest.java,
GemmOffHeapParityTest.java, GemmNumericEdgeTest.java, GemmLargeParityTest.java,
GemmCudaParityTest.java) with HIP forced for every size. Tolerances are the
reference's CPU grade: relTol 1e-12 and cpuAbsTolerance (GemmReference.java:175-177).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from jlc_amd import native
from jlc_amd.compute import HIP_POLICY, BLAS3Kernels, Matrix, OffHeapMatrix, OptimizedBLAS3
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4), (12, 12, 12), (64, 64, 64), (128, 64, 16), (16, 64, 128), (96, 96, 1), (128, 3, 128)]
GRID = [0.0, 1.0, -0.5, 2.5]


def seeded_matrix(rows, cols, seed):
    return Matrix.wrap(orc.seeded_array(rows * cols, seed), rows, cols)


def test_device_present():
    rc, count, name = native.probe()
    print(f"device: {name}", flush=True)
    assert rc == native.JLC_OK, native.last_error()
    assert count >= 1


def test_gemm_parity_against_independent_reference():
    """GemmParityTest.gemmParityAgainstIndependentReference (GemmParityTest.java:38-72)."""
    seed = 20260603
    for m, k, n in SHAPES:
        for alpha in GRID:
            for beta in GRID:
                a = seeded_matrix(m, k, seed + 11)
                b = seeded_matrix(k, n, seed + 29)
                c0 = seeded_matrix(m, n, seed + 47)
                expected = orc.gemm(a.getRawData(), b.getRawData(), c0.getRawData(), m, k, n, alpha, beta)
                blas_out = Matrix.wrap(c0.getRawData().copy(), m, n)
                BLAS3Kernels.gemm(a, b, blas_out, alpha, beta, HIP_POLICY)
                orc.assert_parity("BLAS3Kernels.gemm", m, k, n, alpha, beta, seed, expected, blas_out.getRawData(),
                                  1e-12, orc.cpu_abs_tolerance(expected, k))
                opt_out = Matrix.wrap(c0.getRawData().copy(), m, n)
                OptimizedBLAS3.gemm(a, b, opt_out, alpha, beta, HIP_POLICY)
                orc.assert_parity("OptimizedBLAS3.gemm", m, k, n, alpha, beta, seed, expected, opt_out.getRawData(),
                                  1e-12, orc.cpu_abs_tolerance(expected, k))
        seed += 101


def test_tiny_4x4_matches_reference():
    """GemmParityTest.optimizedBlas3WideSimdTiny4x4MatchesReference (GemmParityTest.java:74-89)."""
    a, b, c0 = seeded_matrix(4, 4, 3201), seeded_matrix(4, 4, 3202), seeded_matrix(4, 4, 3203)
    expected = orc.gemm(a.getRawData(), b.getRawData(), c0.getRawData(), 4, 4, 4, 1.0, 0.0)
    c = Matrix.wrap(c0.getRawData().copy(), 4, 4)
    OptimizedBLAS3.gemm(a, b, c, 1.0, 0.0, HIP_POLICY)
    orc.assert_parity("OptimizedBLAS3.gemm 4x4x4", 4, 4, 4, 1.0, 0.0, 3200, expected, c.getRawData(), 1e-12,
                      orc.cpu_abs_tolerance(expected, 4))


def test_matrix_returning_facade():
    """GemmParityTest.gemmMatrixReturningFacadeMatchesReference (GemmParityTest.java:91-101)."""
    a, b = seeded_matrix(16, 9, 3001), seeded_matrix(9, 7, 3002)
    result = BLAS3Kernels.gemm(a, b, HIP_POLICY)
    expected = orc.gemm(a.getRawData(), b.getRawData(), np.zeros(16 * 7), 16, 9, 7, 1.0, 0.0)
    orc.assert_parity("BLAS3Kernels.gemm(Matrix,Matrix,policy)", 16, 9, 7, 1.0, 0.0, 3000, expected,
                      result.getRawData(), 1e-12, orc.cpu_abs_tolerance(expected, 9))


def test_zero_and_degenerate_dimensions():
    """GemmParityTest.gemmSupportsZeroAndDegenerateDimensions (GemmParityTest.java:117-146)."""
    c0 = seeded_matrix(5, 4, 901)
    expected = c0.getRawData() * -0.5
    for api in (BLAS3Kernels.gemm, OptimizedBLAS3.gemm):
        out = Matrix.wrap(c0.getRawData().copy(), 5, 4)
        api(Matrix(5, 0), Matrix(0, 4), out, 2.5, -0.5, HIP_POLICY)
        orc.assert_parity("zero-k", 5, 0, 4, 2.5, -0.5, 901, expected, out.getRawData(), 1e-12, 1e-12)
        assert np.array_equal(out.getRawData(), expected)  # k == 0 => C = beta*C exactly


def test_strided_row_major():
    """GemmStridedParityTest.stridedRowMajorParity (GemmStridedParityTest.java:6-29)."""
    m, k, n, aOff, bOff, cOff = 9, 7, 11, 13, 13, 13
    lda, ldb, ldc, seed = k + 5, n + 7, n + 3, 4101
    a = orc.seeded_array(aOff + lda * m + 17, seed + 1)
    b = orc.seeded_array(bOff + ldb * k + 17, seed + 2)
    c = orc.seeded_array(cOff + ldc * m + 17, seed + 3)
    expected = orc.gemm_strided(a, aOff, lda, b, bOff, ldb, c, cOff, ldc, m, k, n, 2.5, -0.5)
    guard = c.copy()
    BLAS3Kernels.gemmStrided(a, aOff, lda, b, bOff, ldb, c, cOff, ldc, m, k, n, 2.5, -0.5, 19)
    orc.assert_parity("BLAS3Kernels.gemmStrided", m, k, n, 2.5, -0.5, seed, expected,
                      orc.extract_row_major(c, cOff, ldc, m, n), 1e-12, orc.cpu_abs_tolerance(expected, k))
    _assert_outside_untouched(c, guard, cOff, ldc, m, n)


def test_strided_transposed_a():
    """GemmStridedParityTest.stridedTransposedAParity (GemmStridedParityTest.java:31-54)."""
    m, k, n, off = 8, 5, 6, 13
    lda, ldb, ldc, seed = k + 7, n + 5, n + 3, 4201
    a = orc.seeded_array(off + lda * m + 9, seed + 1)
    b = orc.seeded_array(off + ldb * m + 9, seed + 2)
    c = orc.seeded_array(off + ldc * k + 9, seed + 3)
    expected = orc.gemm_strided_trans_a(a, off, lda, b, off, ldb, c, off, ldc, m, k, n, 1.0, 2.5)
    guard = c.copy()
    BLAS3Kernels.gemmStridedTransA(a, off, lda, b, off, ldb, c, off, ldc, m, k, n, 1.0, 2.5, 17)
    orc.assert_parity("BLAS3Kernels.gemmStridedTransA", k, m, n, 1.0, 2.5, seed, expected,
                      orc.extract_row_major(c, off, ldc, k, n), 1e-12, orc.cpu_abs_tolerance(expected, m))
    _assert_outside_untouched(c, guard, off, ldc, k, n)


def test_strided_column_major_variants():
    """GemmStridedParityTest.stridedColumnMajorVariantsParity (GemmStridedParityTest.java:56-107)."""
    m, k, n = 10, 6, 7
    rmA, rmB, cmA, cmB, ldc, off, seed = k + 4, n + 7, m + 5, k + 3, n + 5, 13, 4301
    a_rm = orc.seeded_array(m * k, seed + 1)
    b_rm = orc.seeded_array(k * n, seed + 2)
    cA = orc.seeded_array(off + ldc * m + 9, seed + 3)
    cB = orc.seeded_array(off + ldc * m + 9, seed + 4)
    a_cm = orc.to_column_major(a_rm, m, k, cmA)
    b_cm = orc.to_column_major(b_rm, k, n, cmB)
    b_rs = np.zeros(off + rmB * k + 9)
    a_rs = np.zeros(off + rmA * m + 9)
    for i in range(m):
        a_rs[off + i * rmA: off + i * rmA + k] = a_rm[i * k:(i + 1) * k]
    for i in range(k):
        b_rs[off + i * rmB: off + i * rmB + n] = b_rm[i * n:(i + 1) * n]
    expA = orc.gemm_strided_col_major_a(a_cm, 0, cmA, b_rs, off, rmB, cA, off, ldc, m, k, n, -0.5, 1.0)
    BLAS3Kernels.gemmStridedColMajorA(a_cm, 0, cmA, b_rs, off, rmB, cA, off, ldc, m, k, n, -0.5, 1.0, 15)
    orc.assert_parity("BLAS3Kernels.gemmStridedColMajorA", m, k, n, -0.5, 1.0, seed, expA,
                      orc.extract_row_major(cA, off, ldc, m, n), 1e-12, orc.cpu_abs_tolerance(expA, k))
    expB = orc.gemm_strided_col_major_b(a_rs, off, rmA, b_cm, 0, cmB, cB, off, ldc, m, k, n, 2.5, 0.0)
    BLAS3Kernels.gemmStridedColMajorB(a_rs, off, rmA, b_cm, 0, cmB, cB, off, ldc, m, k, n, 2.5, 0.0, 15)
    orc.assert_parity("BLAS3Kernels.gemmStridedColMajorB", m, k, n, 2.5, 0.0, seed, expB,
                      orc.extract_row_major(cB, off, ldc, m, n), 1e-12, orc.cpu_abs_tolerance(expB, k))


def test_optimized_strided():
    """GemmStridedParityTest.optimizedStridedParity (GemmStridedParityTest.java:109-130)."""
    m, k, n, off = 13, 9, 8, 13
    lda, ldb, ldc, seed = k + 7, n + 5, n + 1, 4401
    a = orc.seeded_array(off + lda * m + 5, seed + 1)
    b = orc.seeded_array(off + ldb * k + 5, seed + 2)
    c = orc.seeded_array(off + ldc * m + 5, seed + 3)
    expected = orc.gemm_strided(a, off, lda, b, off, ldb, c, off, ldc, m, k, n, 1.0, -0.5)
    OptimizedBLAS3.gemmStrided(a, off, lda, b, off, ldb, c, off, ldc, m, k, n, 1.0, -0.5, HIP_POLICY)
    orc.assert_parity("OptimizedBLAS3.gemmStrided", m, k, n, 1.0, -0.5, seed, expected,
                      orc.extract_row_major(c, off, ldc, m, n), 1e-12, orc.cpu_abs_tolerance(expected, k))


def _fill_offheap(mat, seed):
    data = orc.seeded_array(mat.getRowCount() * mat.getColumnCount(), seed)
    mat.getRawData()[:] = data
    mat.syncToOffHeap()


def test_offheap_facade_and_output_sync():
    """GemmOffHeapParityTest (GemmOffHeapParityTest.java:20-78)."""
    a, b = OffHeapMatrix(24, 16), OffHeapMatrix(16, 10)
    _fill_offheap(a, 5101)
    _fill_offheap(b, 5102)
    product = BLAS3Kernels.gemm(a, b, HIP_POLICY)
    assert isinstance(product, OffHeapMatrix)
    expected = orc.gemm(a.getRawData(), b.getRawData(), np.zeros(240), 24, 16, 10, 1.0, 0.0)
    orc.assert_parity("BLAS3Kernels.gemm off-heap", 24, 16, 10, 1.0, 0.0, 5100, expected, product.getRawData(),
                      1e-12, orc.cpu_abs_tolerance(expected, 16))
    a, b, c = OffHeapMatrix(18, 12), OffHeapMatrix(12, 14), OffHeapMatrix(18, 14)
    _fill_offheap(a, 5201)
    _fill_offheap(b, 5202)
    _fill_offheap(c, 5203)
    expected = orc.gemm(a.getRawData(), b.getRawData(), c.getRawData(), 18, 12, 14, -0.5, 2.5)
    OptimizedBLAS3.gemm(a, b, c, -0.5, 2.5, HIP_POLICY)
    orc.assert_parity("OptimizedBLAS3.gemm off-heap", 18, 12, 14, -0.5, 2.5, 5200, expected, c.getRawData(), 1e-12,
                      orc.cpu_abs_tolerance(expected, 12))
    for r, col in ((0, 0), (9, 7), (17, 13)):
        assert abs(c.getOffHeap(r, col) - expected[r * 14 + col]) <= 1e-12


def test_numeric_edge_known_answers():
    """GemmNumericEdgeTest (GemmNumericEdgeTest.java:32-95) known answers, BLAS semantics.

    HIP_POLICY forces the GPU path even for these 2 x 2 and 1 x 1 calls, where beta = 0 clears a
    NaN-seeded C. The reference's OptimizedBLAS3 TINY CPU path keeps that NaN
    (GemmNumericEdgeTest.java:44-47); the expectation here is safe only because the shipped offload
    gate (hip_min_flops 14,155,776) never sends TINY sizes (under 512 flops) to the GPU, so under the
    default policy such calls stay on the CPU path with the reference's semantics."""
    cases = json.loads((Path(__file__).parent / "golden" / "gemm_kat.json").read_text())["cases"]
    for cs in cases:
        conv = lambda v: float(v)
        a = np.array([conv(v) for v in cs["a"]])
        b = np.array([conv(v) for v in cs["b"]])
        c = np.array([conv(v) for v in cs["c"]])
        exp = np.array([conv(v) for v in cs["expected"]])
        out = Matrix.wrap(c, cs["m"], cs["n"])
        OptimizedBLAS3.gemm(Matrix.wrap(a, cs["m"], cs["k"]), Matrix.wrap(b, cs["k"], cs["n"]), out, cs["alpha"],
                            cs["beta"], HIP_POLICY)
        got = out.getRawData()
        assert np.array_equal(np.isnan(got), np.isnan(exp)), cs["name"]
        assert np.array_equal(got[~np.isnan(exp)], exp[~np.isnan(exp)]), cs["name"]


@pytest.mark.parametrize("m,k,n,seed", [(512, 512, 512, 8001), (768, 512, 256, 8101)])
def test_large_parity_cases(m, k, n, seed):
    """GemmLargeParityTest.largeParityCases (GemmLargeParityTest.java:149-175)."""
    a, b, c0 = seeded_matrix(m, k, seed + 1), seeded_matrix(k, n, seed + 2), seeded_matrix(m, n, seed + 3)
    expected = orc.gemm(a.getRawData(), b.getRawData(), c0.getRawData(), m, k, n, 1.0, -0.5)
    for api in (BLAS3Kernels.gemm, OptimizedBLAS3.gemm):
        out = Matrix.wrap(c0.getRawData().copy(), m, n)
        api(a, b, out, 1.0, -0.5, HIP_POLICY)
        orc.assert_parity("large", m, k, n, 1.0, -0.5, seed, expected, out.getRawData(), 1e-12,
                          orc.cpu_abs_tolerance(expected, k))


def test_gpu_grade_case_meets_cpu_grade():
    """GemmCudaParityTest (GemmCudaParityTest.java:184-201): 64^3, alpha=1, beta=-0.5.
    The reference accepts relTol 1e-10 for its GPU; the HIP path is held to the CPU grade."""
    a, b, c0 = seeded_matrix(64, 64, 7001), seeded_matrix(64, 64, 7002), seeded_matrix(64, 64, 7003)
    expected = orc.gemm(a.getRawData(), b.getRawData(), c0.getRawData(), 64, 64, 64, 1.0, -0.5)
    out = Matrix.wrap(c0.getRawData().copy(), 64, 64)
    OptimizedBLAS3.gemm(a, b, out, 1.0, -0.5, HIP_POLICY)
    orc.assert_parity("HipGemm.gemm", 64, 64, 64, 1.0, -0.5, 7000, expected, out.getRawData(), 1e-12,
                      orc.cpu_abs_tolerance(expected, 64))


def _assert_outside_untouched(c, guard, off, ldc, rows, cols):
    mask = np.ones(c.size, bool)
    for i in range(rows):
        mask[off + i * ldc: off + i * ldc + cols] = False
    assert np.array_equal(c[mask], guard[mask])


# --------------------------------------------------------------------------
# Beyond the reference's own cases: every layout, ragged edges, padding.

def _layout_case(rng, la, lb, m, n, k, pad_a, pad_b, pad_c, alpha, beta, a_off=0, b_off=0, c_off=0):
    lda = (k if la == 0 else m) + pad_a
    ldb = (n if lb == 0 else k) + pad_b
    ldc = n + pad_c
    a = rng.uniform(-1, 1, a_off + lda * (m if la == 0 else k))
    b = rng.uniform(-1, 1, b_off + ldb * (k if lb == 0 else n))
    c = rng.uniform(-1, 1, c_off + ldc * m)
    expc = c.copy()
    orc.gemm_general(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, expc, ldc, a_off, b_off, c_off)
    native.dgemm(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, native.MODE_FP64, a_off, b_off, c_off)
    exp_blk = orc.extract_row_major(expc, c_off, ldc, m, n)
    got_blk = orc.extract_row_major(c, c_off, ldc, m, n)
    orc.assert_parity(f"jlc_dgemm la={la} lb={lb}", m, k, n, alpha, beta, 0, exp_blk, got_blk, 1e-12,
                      orc.cpu_abs_tolerance(exp_blk, k))
    assert np.array_equal(np.delete(c, _block_index(c_off, ldc, m, n)), np.delete(expc, _block_index(c_off, ldc, m, n)))


def _block_index(off, ld, rows, cols):
    return (off + np.arange(rows)[:, None] * ld + np.arange(cols)[None, :]).ravel()


@pytest.mark.parametrize("la,lb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_all_layouts_ragged_edges(la, lb):
    rng = np.random.default_rng(100 + 10 * la + lb)
    # K a multiple of 16 with ragged M / N takes the clamped edge loads over every K tile; other K
    # zero-fill only the last K tile (one K tile, two, or many before it).
    shapes = [(1, 1, 1), (1, 7, 1), (2, 3, 5), (17, 33, 15), (127, 129, 16), (128, 128, 17), (129, 127, 255),
              (200, 1, 300), (257, 65, 130), (300, 250, 37), (130, 260, 64), (1, 300, 32), (333, 7, 48),
              (65, 129, 160), (129, 257, 33), (256, 130, 1000)]
    for m, n, k in shapes:
        alpha = float(rng.choice([1.0, -0.5, 2.5]))
        beta = float(rng.choice([0.0, 1.0, -0.5]))
        _layout_case(rng, la, lb, m, n, k, int(rng.integers(0, 4)), int(rng.integers(0, 4)),
                     int(rng.integers(0, 4)), alpha, beta, int(rng.integers(0, 3)), int(rng.integers(0, 3)),
                     int(rng.integers(0, 3)))


@pytest.mark.parametrize("la,lb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_device_pointer_entry_aligned_and_misaligned(la, lb):
    rng = np.random.default_rng(7 + la * 2 + lb)
    for m, n, k, pad, off in [(192, 160, 96, 0, 0), (130, 70, 45, 1, 1), (64, 256, 33, 3, 5)]:
        lda = (k if la == 0 else m) + pad
        ldb = (n if lb == 0 else k) + pad
        ldc = n + pad
        a = rng.uniform(-1, 1, off + lda * (m if la == 0 else k))
        b = rng.uniform(-1, 1, off + ldb * (k if lb == 0 else n))
        c = rng.uniform(-1, 1, off + ldc * m)
        exp = c.copy()
        orc.gemm_general(la, lb, m, n, k, 1.25, a, lda, b, ldb, 0.75, exp, ldc, off, off, off)
        da, db, dc = native.DeviceArray.from_numpy(a), native.DeviceArray.from_numpy(b), native.DeviceArray.from_numpy(c)
        st = native.dgemm_dev(la, lb, m, n, k, 1.25, da.at(off), lda, db.at(off), ldb, 0.75, dc.at(off), ldc)
        assert st.kernel_name == "mfma_f64"
        got = dc.to_numpy()
        e = orc.extract_row_major(exp, off, ldc, m, n)
        g = orc.extract_row_major(got, off, ldc, m, n)
        orc.assert_parity("jlc_dgemm_dev", m, k, n, 1.25, 0.75, 0, e, g, 1e-12, orc.cpu_abs_tolerance(e, k))


def test_exact_integer_products_bitwise_at_2048():
    """Small-integer inputs make every partial sum exact, so any summation order
    gives the same bits: the HIP result must equal numpy's bit for bit."""
    rng = np.random.default_rng(2048)
    m = n = k = 2048
    A = rng.integers(-4, 5, (m, k)).astype(np.float64)
    B = rng.integers(-4, 5, (k, n)).astype(np.float64)
    C0 = rng.integers(-4, 5, (m, n)).astype(np.float64)
    exp = 2.0 * (A @ B) - 1.0 * C0
    for la, lb in [(0, 0), (1, 1)]:
        a = A.reshape(-1).copy() if la == 0 else A.T.copy().reshape(-1)
        b = B.reshape(-1).copy() if lb == 0 else B.T.copy().reshape(-1)
        c = C0.reshape(-1).copy()
        native.dgemm(la, lb, m, n, k, 2.0, a, k if la == 0 else m, b, n if lb == 0 else k, -1.0, c, n)
        assert np.array_equal(c, exp.reshape(-1)), (la, lb)
