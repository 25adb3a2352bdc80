"""GPU parity of the MATVEC kernel (SpecializedKernels.matvec,
src/main/java/net/faulj/compute/SpecializedKernels.java:27-71) against the oracle."""
import numpy as np
import pytest

from jlc_amd import native
from jlc_amd.compute import HIP_POLICY, Matrix, OptimizedBLAS3
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GRID = [0.0, 1.0, -0.5, 2.5]


def test_matvec_shape_of_parity_grid():
    """The {96,96,1} MATVEC shape of GemmParityTest (GemmParityTest.java:31) x alpha, beta grid."""
    m, k = 96, 96
    seed = 20260603 + 5 * 101
    a = orc.seeded_array(m * k, seed + 11)
    x = orc.seeded_array(k, seed + 29)
    y0 = orc.seeded_array(m, seed + 47)
    for alpha in GRID:
        for beta in GRID:
            exp = orc.gemm(a, x, y0, m, k, 1, alpha, beta)
            y = y0.copy()
            st = native.dgemv(m, k, alpha, a, k, x, 1, beta, y, 1)
            assert st.kernel_name == "gemv"
            orc.assert_parity("jlc_dgemv", m, k, 1, alpha, beta, seed, exp, y, 1e-12, orc.cpu_abs_tolerance(exp, k))
            # and as an n == 1 GEMM through the mirror
            out = Matrix.wrap(y0.copy(), m, 1)
            OptimizedBLAS3.gemm(Matrix.wrap(a, m, k), Matrix.wrap(x, k, 1), out, alpha, beta, HIP_POLICY)
            orc.assert_parity("OptimizedBLAS3.gemm n=1", m, k, 1, alpha, beta, seed, exp, out.getRawData(), 1e-12,
                              orc.cpu_abs_tolerance(exp, k))


def test_matvec_one_by_one_edges():
    """1x1 problems are MATVEC (n == 1 is checked first, GemmDispatch.java:80-82):
    GemmNumericEdgeTest.java:77-95 known answers."""
    y = np.array([4.0])
    native.dgemv(1, 1, 0.0, np.array([np.nan]), 1, np.array([np.nan]), 1, -0.5, y, 1)
    assert y[0] == -2.0
    y = np.array([7.0])
    native.dgemv(1, 1, 1.0, np.array([0.0]), 1, np.array([np.inf]), 1, 1.0, y, 1)
    assert np.isnan(y[0])
    y = np.array([np.nan])
    native.dgemv(1, 3, 1.0, np.array([1.0, 2.0, 3.0]), 3, np.array([1.0, 1.0, 1.0]), 1, 0.0, y, 1)
    assert y[0] == 6.0


def test_matvec_strided_and_ragged():
    rng = np.random.default_rng(42)
    for m, k, lda, incx, incy in [(1, 1, 1, 1, 1), (5, 3, 4, 2, 3), (37, 129, 131, 1, 1), (300, 257, 257, 3, 2),
                                  (1000, 1, 1, 1, 1), (7, 4097, 4100, 1, 1)]:
        a = rng.uniform(-1, 1, m * lda)
        x = rng.uniform(-1, 1, k * incx)
        y = rng.uniform(-1, 1, m * incy)
        xs = x[::incx][:k].copy()
        exp = y[::incy][:m].copy()
        orc.matvec(m, k, -1.5, a, lda, xs, 0.5, exp)
        native.dgemv(m, k, -1.5, a, lda, x, incx, 0.5, y, incy)
        got = y[::incy][:m]
        orc.assert_parity("jlc_dgemv strided", m, k, 1, -1.5, 0.5, 0, exp, got, 1e-12, orc.cpu_abs_tolerance(exp, k))


def test_matvec_exact_integers_large_bitwise():
    """At a large shape, small-integer inputs make the dot products exact."""
    rng = np.random.default_rng(8)
    m, k = 4096, 8192
    A = rng.integers(-3, 4, (m, k)).astype(np.float64)
    x = rng.integers(-3, 4, k).astype(np.float64)
    da, dx = native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(x)
    dy = native.DeviceArray.from_numpy(np.full(m, np.nan))
    st = native.dgemv_dev(m, k, 1.0, da.ptr, k, dx.ptr, 1, 0.0, dy.ptr, 1)
    assert st.kernel_name == "gemv"
    assert np.array_equal(dy.to_numpy(), A @ x)
