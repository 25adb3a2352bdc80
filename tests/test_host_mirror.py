"""Host-mirror logic of the reference's GEMM API that runs without a GPU:
dispatch order and thresholds, argument validation, degenerate shapes."""
import numpy as np
import pytest

from jlc_amd.compute import (BLAS3Kernels, DispatchPolicy, GemmDispatch, HipGemm, IllegalArgumentException, Matrix,
                             OffHeapMatrix, OptimizedBLAS3, UnsupportedKernel)

K = GemmDispatch.Kernel
CPU = DispatchPolicy.builder().enableHip(False).parallelism(1).build()


def test_select_kernel_reference_order_without_hip():
    # GemmDispatch.selectKernel (GemmDispatch.java:76-106)
    assert GemmDispatch.selectKernel(96, 1, 96, False, 1) is K.MATVEC
    assert GemmDispatch.selectKernel(1, 1, 1, False, 1) is K.MATVEC  # n == 1 before TINY
    assert GemmDispatch.selectKernel(4, 4, 4, False, 1) is K.TINY
    assert GemmDispatch.selectKernel(6, 6, 7, False, 1) is K.TINY  # 504 flops
    assert GemmDispatch.selectKernel(8, 8, 8, False, 1) is K.SMALL  # 1024 flops > 512
    assert GemmDispatch.selectKernel(9, 8, 3, False, 1) is K.SMALL
    assert GemmDispatch.selectKernel(12, 12, 12, False, 1) is K.SMALL
    assert GemmDispatch.selectKernel(64, 64, 64, False, 1) is K.MICROKERNEL
    assert GemmDispatch.selectKernel(512, 512, 512, False, 8) is K.PARALLEL_MICRO
    assert GemmDispatch.selectKernel(128, 128, 128, False, 8) is K.MICROKERNEL  # 4.2M < 5M


def test_select_kernel_hip_gate():
    pol = DispatchPolicy.builder().hipMinFlops(200_000_000).build()  # the CUDA_THRESHOLD value
    assert GemmDispatch.selectKernel(512, 512, 512, True, 8, pol) is K.HIP  # 2.68e8 >= 2e8
    assert GemmDispatch.selectKernel(256, 256, 256, True, 8, pol) is K.PARALLEL_MICRO  # 3.4e7 < 2e8
    big = DispatchPolicy.builder().hipMinFlops(100_000_000).build()
    assert GemmDispatch.selectKernel(512, 512, 512, True, 8, big) is K.HIP
    assert GemmDispatch.selectKernel(512, 512, 512, False, 8, big) is K.PARALLEL_MICRO
    forced = DispatchPolicy.builder().hipMinFlops(0).build()
    assert GemmDispatch.selectKernel(4, 4, 4, True, 1, forced) is K.HIP
    off = DispatchPolicy.builder().enableHip(False).hipMinFlops(0).build()
    assert GemmDispatch.selectKernel(4, 4, 4, True, 1, off) is K.TINY
    dim = DispatchPolicy.builder().hipMinFlops(0).hipMinDim(256).build()
    assert GemmDispatch.selectKernel(128, 4096, 4096, True, 1, dim) is not K.HIP


def test_hip_gate_matches_should_offload_to_cuda():
    """shouldOffloadToCuda (DispatchPolicy.java:226-235): minDim applies to every shape, n == 1
    included, and the element gate counts m*k + k*n + m*n."""
    dim = DispatchPolicy.builder().hipMinFlops(0).hipMinDim(64).build()
    assert not dim.shouldOffloadToHip(128, 1, 128)
    assert GemmDispatch.selectKernel(128, 1, 128, True, 1, dim) is K.MATVEC
    assert dim.shouldOffloadToHip(64, 64, 64)
    elems = DispatchPolicy.builder().hipMinFlops(0).hipMinElements(1_000_000).build()
    # long k, small C: 8*100000 + 100000*8 + 8*8 = 1,600,064 elements pass the gate
    assert elems.shouldOffloadToHip(8, 8, 100_000)
    assert GemmDispatch.selectKernel(8, 8, 100_000, True, 1, elems) is K.HIP
    # 3 * 512^2 = 786,432 elements stay below it
    assert not elems.shouldOffloadToHip(512, 512, 512)
    assert elems.shouldOffloadToHip(1000, 1000, 1)  # 1000*1 + 1*1000 + 1000*1000 = 1,002,000


def test_rejects_null_and_dimension_mismatch():
    # GemmParityTest.gemmRejectsNullAndDimensionMismatch (GemmParityTest.java:104-115)
    a, b, c = Matrix(2, 3), Matrix(4, 2), Matrix(2, 2)
    for fn in (lambda: BLAS3Kernels.gemm(None, b, c, 1.0, 0.0, CPU),
               lambda: OptimizedBLAS3.gemm(a, None, c, 1.0, 0.0, CPU),
               lambda: BLAS3Kernels.gemm(a, b, c, 1.0, 0.0, CPU),
               lambda: OptimizedBLAS3.gemm(a, b, c, 1.0, 0.0, CPU),
               lambda: BLAS3Kernels.gemm(a, Matrix(3, 2), Matrix(1, 2), 1.0, 0.0, CPU),
               lambda: OptimizedBLAS3.gemm(a, Matrix(3, 2), Matrix(1, 2), 1.0, 0.0, CPU),
               lambda: HipGemm.gemm(a, b, None, 1.0, 0.0)):
        with pytest.raises(IllegalArgumentException) as ei:
            fn()
        assert str(ei.value)


def test_zero_dimensions_need_no_device():
    a, b, c = Matrix(0, 3), Matrix(3, 2), Matrix(0, 2)
    BLAS3Kernels.gemm(a, b, c, 1.0, 0.0, CPU)
    OptimizedBLAS3.gemm(a, b, c, 1.0, 0.0, CPU)
    assert c.getRawData().size == 0
    empty = BLAS3Kernels.gemm(Matrix(2, 0), Matrix(0, 0), CPU)
    assert (empty.getRowCount(), empty.getColumnCount()) == (2, 0)


def test_cpu_regime_is_not_served_here():
    a, b, c = Matrix(4, 4), Matrix(4, 4), Matrix(4, 4)
    with pytest.raises(UnsupportedKernel):
        OptimizedBLAS3.gemm(a, b, c, 1.0, 0.0, CPU)


def test_offheap_sync_orders():
    m = OffHeapMatrix(2, 3, OffHeapMatrix.COLUMN_MAJOR)
    m.getRawData()[:] = np.arange(6.0)
    m.syncToOffHeap()
    assert m.getOffHeap(1, 2) == 5.0 and m.segment()[1] == 3.0
    m.getRawData()[:] = 0
    m.syncFromOffHeap()
    assert np.array_equal(m.getRawData(), np.arange(6.0))


def test_multiply_facade_semantics_without_a_device():
    """BlockedMultiply.multiply (BlockedMultiply.java:55-105): null and inner-dimension errors with a
    message, empty products as a new zero matrix (off-heap when either input is), n == 1 to the Java
    host's matvec, small sizes to its naive / SIMD loops."""
    from jlc_amd.compute import BlockedMultiply
    with pytest.raises(IllegalArgumentException) as ei:
        BlockedMultiply.multiply(None, Matrix(2, 2))
    assert str(ei.value)
    with pytest.raises(IllegalArgumentException) as ei:
        Matrix(2, 3).multiply(Matrix(4, 2))
    assert "3 != 4" in str(ei.value)
    z = BlockedMultiply.multiply(Matrix(3, 0), Matrix(0, 5), CPU)
    assert type(z) is Matrix and (z.getRowCount(), z.getColumnCount()) == (3, 5) and not z.getRawData().any()
    z = BlockedMultiply.multiply(OffHeapMatrix(0, 4), Matrix(4, 2), CPU)
    assert isinstance(z, OffHeapMatrix) and (z.getRowCount(), z.getColumnCount()) == (0, 2)
    z = BlockedMultiply.multiply(Matrix(2, 3), OffHeapMatrix(3, 0), CPU)
    assert isinstance(z, OffHeapMatrix)
    with pytest.raises(UnsupportedKernel):
        BlockedMultiply.multiply(Matrix(300, 300), Matrix(300, 1), CPU)
    with pytest.raises(UnsupportedKernel):
        Matrix(8, 8).multiply(Matrix(8, 8))


def test_select_for_multiply_order():
    """selectBaseAlgorithm (DispatchPolicy.java:254-279): sizes up to naiveThreshold never offload,
    then the HIP gate (in the CUDA slot), then STRASSEN / BLAS3 / SIMD / BLOCKED / NAIVE."""
    from jlc_amd.compute import Algorithm
    on = DispatchPolicy.builder().hipMinFlops(0).hipDetector(lambda: True).build()
    off = DispatchPolicy.builder().hipMinFlops(0).hipDetector(lambda: False).build()
    assert on.selectForMultiply(64, 64, 64) is Algorithm.SIMD      # maxDim <= 64: CPU even with HIP
    assert on.selectForMultiply(65, 65, 65) is Algorithm.HIP
    assert off.selectForMultiply(65, 65, 65) is Algorithm.SIMD     # 65 < simdThreshold 256
    assert off.selectForMultiply(300, 300, 300) is Algorithm.BLAS3
    assert on.selectCpuAlgorithm(300, 300, 300) is Algorithm.BLAS3
    gated = DispatchPolicy.builder().hipMinFlops(10 ** 9).hipDetector(lambda: True).build()
    assert gated.selectForMultiply(300, 300, 300) is Algorithm.BLAS3  # 5.4e7 flops < 1e9
    assert gated.selectForMultiply(1000, 1000, 1000) is Algorithm.HIP
    nosimd = DispatchPolicy(hip_min_flops=0, hip_detector=lambda: False, simd_available=False, enable_blas3=False)
    assert nosimd.selectForMultiply(100, 100, 100) is Algorithm.NAIVE
    assert nosimd.selectForMultiply(300, 100, 100) is Algorithm.BLOCKED
    assert nosimd.selectForMultiply(10, 10, 10) is Algorithm.NAIVE


def test_shipped_offload_gate_keeps_tiny_calls_off_the_gpu():
    """The library's BLAS beta = 0 semantics clear a NaN-seeded C, where the reference's TINY CPU path
    keeps it (GemmNumericEdgeTest.java:44-47); the two never meet while the shipped offload gate sits
    above the TINY regime's 512 flops (GemmDispatch.java:30-33)."""
    from pathlib import Path
    table = Path(__file__).resolve().parent.parent / "jlc_amd" / "tuning" / "mi355x.tune"
    gate = None
    for line in table.read_text().splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0] == "hip_min_flops":
            gate = float(parts[1])
    assert gate is not None and gate > 512
    assert GemmDispatch.selectKernel(2, 2, 2, False, 1) is K.TINY and 2 * 2 * 2 * 2 < gate
