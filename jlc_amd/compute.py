"""Host-side mirror of the reference's GEMM surface, routed to libjlc_hip.

Names, argument order and error behaviour follow the Java classes in
src/main/java/net/faulj/compute/ so that the parity tests read like the
reference's own tests:

* Matrix / OffHeapMatrix       net/faulj/matrix/Matrix.java:18-24, :149, :385;
                               net/faulj/matrix/OffHeapMatrix.java:15-314
* DispatchPolicy               DispatchPolicy.java:57-679 (builder, HIP gating fields
                               mirroring cudaMinDim / cudaMinElements / cudaMinFlops)
* GemmDispatch                 GemmDispatch.java:40-47 (Kernel), :76-106 (selectKernel)
                               with the new HIP member in the CUDA slot
* HipGemm                      the HIP twin of CudaGemm.gemm (CudaGemm.java:53-135):
                               IllegalArgumentException for null/shape errors, False for
                               device failures, never raises for a device problem
* OptimizedBLAS3.gemm /        OptimizedBLAS3.java:61-130, :135-179
  gemmStrided
* BLAS3Kernels.gemm* /         BLAS3Kernels.java:35-139, :244-303
  gemmStrided*
* BlockedHessenbergQR.         eigen/qr/BlockedHessenbergQR.java:277-345 (compact-WY updates)
  applyBlockLeft / Right
* TallSkinnyQR /                decomposition/qr/TallSkinnyQR.java:22-90, CommunicationAvoidingQR.java:18-20
  CommunicationAvoidingQR
* Matrix.multiply /             Matrix.java:1120-1122; BlockedMultiply.java:43-105, with
  BlockedMultiply.multiply      DispatchPolicy.selectForMultiply (:179-193, :254-279) and
                                CudaGemm.multiply (CudaGemm.java:32-41) -> HipGemm.multiply

This build contains no CPU GEMM. Where the reference would run one of its own
CPU regimes (TINY / SMALL / MICROKERNEL / PARALLEL_MICRO), the mirror raises
UnsupportedKernel: in the Java integration those branches stay the
reference's own code. A policy with hip_min_flops(0) routes every size to HIP.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field, replace

import numpy as np

from . import native
from .native import COL, MODE_FP64, ROW


class IllegalArgumentException(ValueError):
    """Mirror of java.lang.IllegalArgumentException (non-empty message)."""


class UnsupportedKernel(RuntimeError):
    """Dispatch chose a CPU regime that only the Java host implements."""


# ----------------------------------------------------------------------------
# Operands

class Matrix:
    """Row-major dense real matrix backed by a float64 vector (Matrix.java:18-24)."""

    def __init__(self, rows: int, cols: int, data: np.ndarray | None = None):
        if rows < 0 or cols < 0:
            raise IllegalArgumentException("Matrix dimensions must be non-negative")
        self._rows = int(rows)
        self._cols = int(cols)
        if data is None:
            data = np.zeros(self._rows * self._cols, dtype=np.float64)
        data = np.asarray(data, dtype=np.float64)
        if data.ndim != 1 or data.size != self._rows * self._cols:
            raise IllegalArgumentException("Data length must equal rows*cols")
        self._data = data

    @staticmethod
    def wrap(data, rows: int, cols: int) -> "Matrix":
        """Matrix.wrap (Matrix.java:149): shares the array, no copy."""
        return Matrix(rows, cols, np.asarray(data, dtype=np.float64).reshape(-1))

    def getRowCount(self) -> int:
        return self._rows

    def getColumnCount(self) -> int:
        return self._cols

    def getRawData(self) -> np.ndarray:
        return self._data

    def getRawImagData(self):
        return None

    def get(self, i: int, j: int) -> float:
        return float(self._data[i * self._cols + j])

    def set(self, i: int, j: int, v: float) -> None:
        self._data[i * self._cols + j] = v

    def to_numpy(self) -> np.ndarray:
        return self._data.reshape(self._rows, self._cols)

    @staticmethod
    def Identity(n: int) -> "Matrix":
        """Matrix.Identity (Matrix.java:1007-1016): IllegalArgumentException for n <= 0."""
        if n <= 0:
            raise IllegalArgumentException("Size n must be positive")
        return Matrix(n, n, np.eye(n).reshape(-1))

    def copy(self) -> "Matrix":
        return Matrix(self._rows, self._cols, self._data.copy())

    @staticmethod
    def zero(rows: int, cols: int) -> "Matrix":
        """Matrix.zero (Matrix.java:1202-1204)."""
        return Matrix(rows, cols)

    def multiply(self, other: "Matrix") -> "Matrix":
        """Matrix.multiply (Matrix.java:1120-1122): BlockedMultiply.multiply under the default policy."""
        return BlockedMultiply.multiply(self, other)


class OffHeapMatrix(Matrix):
    """FFM-segment-backed operand with a heap mirror (OffHeapMatrix.java:15-314).

    The heap mirror is authoritative for kernel inputs; syncToOffHeap copies
    it into the (row- or column-ordered) segment, as the reference does
    after every GEMM that writes an off-heap C (OptimizedBLAS3.java:127-129).
    """

    ROW_MAJOR = "ROW_MAJOR"
    COLUMN_MAJOR = "COLUMN_MAJOR"

    def __init__(self, rows: int, cols: int, order: str = "ROW_MAJOR"):
        super().__init__(rows, cols)
        self.order = order
        self._segment = np.zeros(rows * cols, dtype=np.float64)

    def segment(self) -> np.ndarray:
        return self._segment

    def _seg_index(self, i: int, j: int) -> int:
        return i * self._cols + j if self.order == self.ROW_MAJOR else j * self._rows + i

    def syncToOffHeap(self) -> None:
        m = self.to_numpy()
        if self.order == self.ROW_MAJOR:
            self._segment[:] = m.reshape(-1)
        else:
            self._segment[:] = m.T.reshape(-1)

    def syncFromOffHeap(self) -> None:
        if self.order == self.ROW_MAJOR:
            self._data[:] = self._segment
        else:
            self._data[:] = self._segment.reshape(self._cols, self._rows).T.reshape(-1)

    def getOffHeap(self, i: int, j: int) -> float:
        return float(self._segment[self._seg_index(i, j)])


# ----------------------------------------------------------------------------
# Policy and dispatch

class Algorithm(enum.Enum):
    """DispatchPolicy.Algorithm (DispatchPolicy.java:58-67) with HIP beside CUDA."""
    NAIVE = "NAIVE"
    BLOCKED = "BLOCKED"
    BLAS3 = "BLAS3"
    SIMD = "SIMD"
    CUDA = "CUDA"
    PARALLEL = "PARALLEL"
    STRASSEN = "STRASSEN"
    SPECIALIZED = "SPECIALIZED"
    HIP = "HIP"


@dataclass(frozen=True)
class DispatchPolicy:
    """Subset of DispatchPolicy (DispatchPolicy.java:57-679) that the GEMM route reads.

    HIP gating mirrors the CUDA fields (enableCuda, cudaMinDim, cudaMinElements,
    cudaMinFlops; builder defaults DispatchPolicy.java:439-441). The defaults
    for HIP come from the tuned table (jlc_amd.tuning) when one is loaded.
    """

    enable_hip: bool = True
    hip_min_dim: int = 0
    hip_min_elements: int = 0
    hip_min_flops: int = 200_000_000
    parallelism: int = 1
    mode: int = MODE_FP64
    # Multiply-route thresholds, the reference's builder defaults (DispatchPolicy.java:424-437).
    naive_threshold: int = 64
    blocked_threshold: int = 256
    strassen_threshold: int = 1024
    enable_blas3: bool = True
    blas3_threshold: int = 256
    enable_simd: bool = True
    simd_threshold: int = 256
    enable_strassen: bool = False
    simd_available: bool = True
    # cudaDetector analogue (DispatchPolicy.java:440): None asks HipSupport.
    hip_detector: object = None

    @staticmethod
    def builder() -> "_PolicyBuilder":
        return _PolicyBuilder()

    @staticmethod
    def defaultPolicy() -> "DispatchPolicy":
        """Default policy: the HIP gate comes from the loaded tuning table."""
        return DispatchPolicy(hip_min_flops=int(native.tune_get("hip_min_flops")))

    def getParallelism(self) -> int:
        return self.parallelism

    def shouldOffloadToHip(self, m: int, n: int, k: int) -> bool:
        """Analogue of shouldOffloadToCuda (DispatchPolicy.java:219-236): the smallest
        dimension against hip_min_dim, the elements of all three operands (m*k + k*n + m*n)
        against hip_min_elements, then 2*m*n*k against hip_min_flops."""
        if not self.enable_hip:
            return False
        if min(m, n, k) < self.hip_min_dim:
            return False
        if m * k + k * n + m * n < self.hip_min_elements:
            return False
        return 2 * m * n * k >= self.hip_min_flops

    def _hip_available(self) -> bool:
        return bool(self.hip_detector()) if self.hip_detector is not None else HipSupport.isHipAvailable()

    def _select_base(self, m: int, n: int, k: int, allow_hip: bool) -> Algorithm:
        """selectBaseAlgorithm (DispatchPolicy.java:254-279) with the HIP gate in the CUDA slot."""
        max_dim = max(m, n, k)
        simd = self.enable_simd and max_dim < self.simd_threshold and self.simd_available
        if max_dim <= self.naive_threshold:
            return Algorithm.SIMD if simd else Algorithm.NAIVE
        if allow_hip and self.enable_hip and self._hip_available() and self.shouldOffloadToHip(m, n, k):
            return Algorithm.HIP
        if self.enable_strassen and m == n == k and max_dim >= self.strassen_threshold:
            return Algorithm.STRASSEN
        if self.enable_blas3 and max_dim >= self.blas3_threshold:
            return Algorithm.BLAS3
        if simd:
            return Algorithm.SIMD
        if max_dim >= self.blocked_threshold:
            return Algorithm.BLOCKED
        return Algorithm.NAIVE

    def selectForMultiply(self, m: int, n: int, k: int) -> Algorithm:
        """selectForMultiply (DispatchPolicy.java:179-181)."""
        return self._select_base(m, n, k, True)

    def selectCpuAlgorithm(self, m: int, n: int, k: int) -> Algorithm:
        """selectCpuAlgorithm (DispatchPolicy.java:191-193)."""
        return self._select_base(m, n, k, False)


class _PolicyBuilder:
    def __init__(self):
        self._kw = {}

    def enableHip(self, v: bool):
        self._kw["enable_hip"] = bool(v)
        return self

    def hipMinDim(self, v: int):
        self._kw["hip_min_dim"] = int(v)
        return self

    def hipMinElements(self, v: int):
        self._kw["hip_min_elements"] = int(v)
        return self

    def hipMinFlops(self, v: int):
        self._kw["hip_min_flops"] = int(v)
        return self

    def parallelism(self, v: int):
        self._kw["parallelism"] = int(v)
        return self

    def mode(self, v: int):
        self._kw["mode"] = int(v)
        return self

    def naiveThreshold(self, v: int):
        self._kw["naive_threshold"] = int(v)
        return self

    def hipDetector(self, fn):
        self._kw["hip_detector"] = fn
        return self

    def build(self) -> DispatchPolicy:
        return DispatchPolicy(**self._kw)


class GemmDispatch:
    """GemmDispatch (GemmDispatch.java:21-239) with a HIP kernel choice."""

    class Kernel(enum.Enum):
        TINY = "TINY"
        SMALL = "SMALL"
        MATVEC = "MATVEC"
        MICROKERNEL = "MICROKERNEL"
        CUDA = "CUDA"
        PARALLEL_MICRO = "PARALLEL_MICRO"
        HIP = "HIP"

    TINY_THRESHOLD = 8 * 8 * 8
    SMALL_THRESHOLD = 50_000
    PARALLEL_THRESHOLD = 5_000_000

    @staticmethod
    def selectKernel(m: int, n: int, k: int, hipAvailable: bool, threads: int,
                     policy: DispatchPolicy | None = None) -> "GemmDispatch.Kernel":
        """selectKernel (GemmDispatch.java:76-106) with HIP tried first when the
        policy's gate (tuned thresholds) accepts the shape, then the reference's
        order: MATVEC, TINY, SMALL, PARALLEL_MICRO, MICROKERNEL."""
        K = GemmDispatch.Kernel
        pol = policy or DispatchPolicy.defaultPolicy()
        if hipAvailable and pol.shouldOffloadToHip(m, n, k):
            return K.HIP
        flops = 2 * m * n * k
        if n == 1:
            return K.MATVEC
        if flops <= GemmDispatch.TINY_THRESHOLD and m <= 8 and n <= 8 and k <= 8:
            return K.TINY
        if flops <= GemmDispatch.SMALL_THRESHOLD:
            return K.SMALL
        if threads > 1 and flops >= GemmDispatch.PARALLEL_THRESHOLD:
            return K.PARALLEL_MICRO
        return K.MICROKERNEL


class HipSupport:
    """Detection analogue of CudaSupport (CudaSupport.java:34-123)."""

    _cached: bool | None = None

    @classmethod
    def isHipAvailable(cls) -> bool:
        if cls._cached is None:
            rc, count, _ = native.probe()
            cls._cached = rc == native.JLC_OK and count > 0
        return cls._cached

    @classmethod
    def refresh(cls) -> None:
        cls._cached = None


class HipGemm:
    """The HIP plug behind the unchanged API: HIP twin of CudaGemm.gemm (CudaGemm.java:53-135)."""

    @staticmethod
    def gemm(a: Matrix, b: Matrix, c: Matrix, alpha: float, beta: float, mode: int = MODE_FP64) -> bool:
        if a is None or b is None or c is None:
            raise IllegalArgumentException("Matrices must not be null")
        m, k, n = a.getRowCount(), a.getColumnCount(), b.getColumnCount()
        if k != b.getRowCount():
            raise IllegalArgumentException("Inner dimensions must agree")
        if m != c.getRowCount() or n != c.getColumnCount():
            raise IllegalArgumentException("Output dimensions must match")
        try:
            native.dgemm(ROW, ROW, m, n, k, alpha, a.getRawData(), max(1, k), b.getRawData(), max(1, n), beta,
                         c.getRawData(), max(1, n), mode)
        except native.JlcError:
            return False
        return True

    @staticmethod
    def multiply(a: Matrix, b: Matrix, mode: int = MODE_FP64) -> "Matrix | None":
        """CudaGemm.multiply (CudaGemm.java:32-41): a new heap Matrix C = A B, or None when the device
        call fails (the caller then runs its CPU algorithm)."""
        c = Matrix.zero(a.getRowCount(), b.getColumnCount())
        return c if HipGemm.gemm(a, b, c, 1.0, 0.0, mode) else None


def _require_hip(kernel: "GemmDispatch.Kernel") -> None:
    if kernel is not GemmDispatch.Kernel.HIP:
        raise UnsupportedKernel(
            f"dispatch selected the CPU regime {kernel.value}; CPU regimes are the Java host's own code")


def _validate(a: Matrix, b: Matrix, c: Matrix) -> tuple[int, int, int]:
    if a is None or b is None or c is None:
        raise IllegalArgumentException("Matrices must not be null")
    m, k, n = a.getRowCount(), a.getColumnCount(), b.getColumnCount()
    if k != b.getRowCount():
        raise IllegalArgumentException("Inner dimensions must agree")
    if m != c.getRowCount() or n != c.getColumnCount():
        raise IllegalArgumentException("Output dimensions must match")
    return m, k, n


def _hip_or_raise(a: Matrix, b: Matrix, c: Matrix, alpha: float, beta: float, mode: int) -> None:
    if not HipGemm.gemm(a, b, c, alpha, beta, mode):
        raise native.JlcError(native.JLC_ERR_HIP, "HIP GEMM failed: " + native.last_error())


# ----------------------------------------------------------------------------
# Public GEMM surfaces

class OptimizedBLAS3:
    @staticmethod
    def gemm(a: Matrix, b: Matrix, c: Matrix, alpha: float, beta: float, policy: DispatchPolicy | None) -> None:
        """OptimizedBLAS3.gemm (OptimizedBLAS3.java:61-130)."""
        m, k, n = _validate(a, b, c)
        if m == 0 or n == 0:
            return
        pol = policy or DispatchPolicy.defaultPolicy()
        kernel = GemmDispatch.selectKernel(m, n, k, HipSupport.isHipAvailable(), pol.getParallelism(), pol)
        _require_hip(kernel)
        _hip_or_raise(a, b, c, alpha, beta, pol.mode)
        if isinstance(c, OffHeapMatrix):
            c.syncToOffHeap()

    @staticmethod
    def gemmStrided(a, aOffset, lda, b, bOffset, ldb, c, cOffset, ldc, m, k, n, alpha, beta,
                    policy: DispatchPolicy | None = None) -> None:
        """OptimizedBLAS3.gemmStrided (OptimizedBLAS3.java:135-179); note the m, k, n order.
        Implements the documented strided semantics (GemmReference.java:46-61)."""
        if m <= 0 or n <= 0:
            return
        pol = policy or DispatchPolicy.defaultPolicy()
        _require_hip(GemmDispatch.selectKernel(m, n, k, HipSupport.isHipAvailable(), 1, pol))
        native.dgemm(ROW, ROW, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, pol.mode, aOffset, bOffset, cOffset)


class BLAS3Kernels:
    @staticmethod
    def gemm(*args):
        """gemm(A,B) / gemm(A,B,policy) return a new C (BLAS3Kernels.java:35-63; an
        OffHeapMatrix if either input is off-heap, :58-60); gemm(A,B,C,alpha,beta,policy)
        updates C (BLAS3Kernels.java:75-139)."""
        if len(args) in (2, 3):
            a, b = args[0], args[1]
            policy = args[2] if len(args) == 3 else None
            if a is None or b is None:
                raise IllegalArgumentException("Matrices must not be null")
            if a.getColumnCount() != b.getRowCount():
                raise IllegalArgumentException("Inner dimensions must agree")
            m, n = a.getRowCount(), b.getColumnCount()
            off = isinstance(a, OffHeapMatrix) or isinstance(b, OffHeapMatrix)
            c = OffHeapMatrix(m, n) if off else Matrix(m, n)
            BLAS3Kernels.gemm(a, b, c, 1.0, 0.0, policy)
            return c
        if len(args) == 6:
            a, b, c, alpha, beta, policy = args
            m, k, n = _validate(a, b, c)
            if m == 0 or n == 0:
                return None
            pol = policy or DispatchPolicy.defaultPolicy()
            _require_hip(GemmDispatch.selectKernel(m, n, k, HipSupport.isHipAvailable(), pol.getParallelism(), pol))
            _hip_or_raise(a, b, c, alpha, beta, pol.mode)
            if isinstance(c, OffHeapMatrix):
                c.syncToOffHeap()
            return None
        raise TypeError("gemm expects (A,B), (A,B,policy) or (A,B,C,alpha,beta,policy)")

    @staticmethod
    def gemmStrided(ad, aOff, lda, bd, bOff, ldb, cd, cOff, ldc, m, k, n, alpha, beta, blockSize=0,
                    mode: int = MODE_FP64) -> None:
        """C[m x n] = alpha*A[m x k]*B[k x n] + beta*C, all row-major (BLAS3Kernels.java:244-251)."""
        if m <= 0 or n <= 0:
            return
        native.dgemm(ROW, ROW, m, n, k, alpha, ad, lda, bd, ldb, beta, cd, ldc, mode, aOff, bOff, cOff)

    @staticmethod
    def gemmStridedTransA(ad, aOff, lda, bd, bOff, ldb, cd, cOff, ldc, m, k, n, alpha, beta, blockSize=0,
                          mode: int = MODE_FP64) -> None:
        """C[k x n] = alpha*A^T*B + beta*C with A stored m x k row-major (BLAS3Kernels.java:253-279):
        m is the inner dimension, k the output row count."""
        if k <= 0 or n <= 0:
            return
        native.dgemm(COL, ROW, k, n, m, alpha, ad, lda, bd, ldb, beta, cd, ldc, mode, aOff, bOff, cOff)

    @staticmethod
    def gemmStridedColMajorA(ad, aOff, lda, bd, bOff, ldb, cd, cOff, ldc, m, k, n, alpha, beta, blockSize=0,
                             mode: int = MODE_FP64) -> None:
        """A column-major: element (i,p) at aOff + p*lda + i (BLAS3Kernels.java:284-291)."""
        if m <= 0 or n <= 0:
            return
        native.dgemm(COL, ROW, m, n, k, alpha, ad, lda, bd, ldb, beta, cd, ldc, mode, aOff, bOff, cOff)

    @staticmethod
    def gemmStridedColMajorB(ad, aOff, lda, bd, bOff, ldb, cd, cOff, ldc, m, k, n, alpha, beta, blockSize=0,
                             mode: int = MODE_FP64) -> None:
        """B column-major: element (p,j) at bOff + j*ldb + p (BLAS3Kernels.java:296-303)."""
        if m <= 0 or n <= 0:
            return
        native.dgemm(ROW, COL, m, n, k, alpha, ad, lda, bd, ldb, beta, cd, ldc, mode, aOff, bOff, cOff)


class BlockedMultiply:
    """BlockedMultiply.multiply (BlockedMultiply.java:43-105), the product behind Matrix.multiply."""

    @staticmethod
    def multiply(a: Matrix, b: Matrix, policy: DispatchPolicy | None = None) -> Matrix:
        if a is None or b is None:
            raise IllegalArgumentException("Matrices must not be null")
        m, k = a.getRowCount(), a.getColumnCount()
        k2, n = b.getRowCount(), b.getColumnCount()
        if k != k2:
            raise IllegalArgumentException(f"Inner dimensions must agree for multiplication: {k} != {k2}")
        if m == 0 or n == 0 or k == 0:
            off = isinstance(a, OffHeapMatrix) or isinstance(b, OffHeapMatrix)
            return OffHeapMatrix(m, n) if off else Matrix.zero(m, n)
        if n == 1:
            raise UnsupportedKernel("multiply: n == 1 runs multiplyVector, the Java host's CPU matvec "
                                    "(BlockedMultiply.java:71-73)")
        active = policy or DispatchPolicy.defaultPolicy()
        algorithm = active.selectForMultiply(m, n, k)
        if algorithm is Algorithm.HIP:
            c = HipGemm.multiply(a, b, active.mode)
            if c is not None:
                return c
            algorithm = active.selectCpuAlgorithm(m, n, k)
            raise UnsupportedKernel(f"multiply: the HIP call failed ({native.last_error()}); the reference "
                                    f"then runs its CPU algorithm {algorithm.value}")
        raise UnsupportedKernel(f"multiply: dispatch selected the CPU algorithm {algorithm.value}; CPU "
                                f"algorithms are the Java host's own code")


class BlockedHessenbergQR:
    """The compact-WY block updates of BlockedHessenbergQR (src/main/java/net/faulj/eigen/qr/
    BlockedHessenbergQR.java:277-345), same arguments and in-place semantics, each three-GEMM sequence
    one jlc_wy_apply call on the row-major n x n array `a` (only V, T and the updated block cross PCIe)."""

    @staticmethod
    def applyBlockLeft(a, n, k, colStart, V, T, m, panelWidth, mode: int = MODE_FP64) -> None:
        """applyBlockLeft (:277-306): a[k+1:, colStart:] <- (I - V T^T V^T) a[k+1:, colStart:]."""
        if colStart >= n:
            return
        native.wy_apply(native.LEFT, 1, m, n - colStart, panelWidth, V, panelWidth, T, panelWidth, a, n, mode,
                        c_off=(k + 1) * n + colStart)

    @staticmethod
    def applyBlockRight(a, n, k, colStart, V, T, m, panelWidth, mode: int = MODE_FP64) -> None:
        """applyBlockRight (:308-345): a[:, colStart:] -= (a[:, k+1:] V T) V[colStart-k-1:, :]^T."""
        fullStart = k + 1
        if fullStart >= n or colStart >= n:
            return
        nCols = n - colStart
        colOffset = colStart - fullStart
        if nCols <= 0 or colOffset < 0:
            return
        if colOffset == 0:  # the form the driver calls (:132-133): one WY update of a[:, k+1:]
            native.wy_apply(native.RIGHT, 0, n, nCols, panelWidth, V, panelWidth, T, panelWidth, a, n, mode,
                            c_off=colStart)
            return
        # W from all m trailing columns, the update on the last nCols of them
        W = np.zeros(n * panelWidth)
        native.dgemm(ROW, ROW, n, panelWidth, m, 1.0, a, n, V, panelWidth, 0.0, W, panelWidth, mode, a_off=fullStart)
        W2 = np.zeros(n * panelWidth)
        native.dgemm(ROW, ROW, n, panelWidth, panelWidth, 1.0, W, panelWidth, T, panelWidth, 0.0, W2, panelWidth)
        native.dgemm(ROW, COL, n, nCols, panelWidth, -1.0, W2, panelWidth, V, panelWidth, 1.0, a, n, mode,
                     b_off=colOffset * panelWidth, c_off=colStart)


class QRResult:
    """QRResult (src/main/java/net/faulj/decomposition/result/QRResult.java:102-135): A, Q, R."""

    def __init__(self, A: Matrix, Q: Matrix, R: Matrix):
        self._A, self._Q, self._R = A, Q, R

    def getA(self) -> Matrix:
        return self._A

    def getQ(self) -> Matrix:
        return self._Q

    def getR(self) -> Matrix:
        return self._R


class TallSkinnyQR:
    """TallSkinnyQR.decompose (src/main/java/net/faulj/decomposition/qr/TallSkinnyQR.java:22-90) on the
    GPU through jlc_dtsqr: the reference's two-level tree (row blocks of max(n, 64) rows, a QR of the
    stacked R factors), thin Q (m x min(m, n)) and R (min(m, n) x n)."""

    DEFAULT_BLOCK_SIZE = 64

    @staticmethod
    def decompose(A: Matrix) -> QRResult:
        if A is None:
            raise IllegalArgumentException("Matrix must not be null")
        m, n = A.getRowCount(), A.getColumnCount()
        if m == 0 or n == 0:
            return QRResult(A, Matrix.Identity(m), A.copy())  # Identity(0) raises, as in the reference
        k = min(m, n)
        Q, R = Matrix(m, k), Matrix(k, n)
        native.dtsqr(m, n, np.ascontiguousarray(A.getRawData()), Q.getRawData(), R.getRawData())
        return QRResult(A, Q, R)


class CommunicationAvoidingQR:
    """CommunicationAvoidingQR.decompose delegates to TallSkinnyQR (CommunicationAvoidingQR.java:18-20)."""

    @staticmethod
    def decompose(A: Matrix) -> QRResult:
        return TallSkinnyQR.decompose(A)


HIP_POLICY = DispatchPolicy.builder().enableHip(True).hipMinFlops(0).parallelism(1).build()
