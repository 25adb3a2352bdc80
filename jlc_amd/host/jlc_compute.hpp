// C++ host mirror of the reference's GEMM surface over libjlc_hip's C ABI.
//
// Class and method names, argument order and error behaviour follow
// src/main/java/net/faulj/compute/ (OptimizedBLAS3, BLAS3Kernels,
// GemmDispatch, DispatchPolicy, CudaGemm -> HipGemm, CudaSupport -> HipSupport)
// and src/main/java/net/faulj/matrix/ (Matrix, OffHeapMatrix), so a C++ or
// JNI/FFM host can drop this layer in where the Java classes call JCublas.
// The reference's CPU regimes (TINY/SMALL/MICROKERNEL/PARALLEL_MICRO) are not
// part of this library: when dispatch selects one, UnsupportedKernel is thrown.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace jlc {
namespace compute {

struct IllegalArgumentException : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

struct UnsupportedKernel : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct DeviceError : std::runtime_error {
    int status;
    DeviceError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// Row-major dense real matrix (Matrix.java:18-24).
class Matrix {
   public:
    Matrix(int rows, int cols);
    Matrix(int rows, int cols, std::vector<double> data);
    static Matrix wrap(std::vector<double> data, int rows, int cols);
    virtual ~Matrix() = default;
    int getRowCount() const { return rows_; }
    int getColumnCount() const { return cols_; }
    std::vector<double>& getRawData() { return data_; }
    const std::vector<double>& getRawData() const { return data_; }
    double get(int i, int j) const { return data_[(size_t)i * cols_ + j]; }
    void set(int i, int j, double v) { data_[(size_t)i * cols_ + j] = v; }
    // Matrix.zero (Matrix.java:1202-1204).
    static Matrix zero(int rows, int cols) { return Matrix(rows, cols); }
    // Matrix.multiply (Matrix.java:1120-1122): BlockedMultiply.multiply under the default policy.
    std::unique_ptr<Matrix> multiply(const Matrix& other) const;

   protected:
    int rows_, cols_;
    std::vector<double> data_;
};

// FFM-segment-backed operand with a heap mirror (OffHeapMatrix.java:15-314).
class OffHeapMatrix : public Matrix {
   public:
    enum class Order { ROW_MAJOR, COLUMN_MAJOR };
    OffHeapMatrix(int rows, int cols, Order order = Order::ROW_MAJOR);
    void syncToOffHeap();
    void syncFromOffHeap();
    double getOffHeap(int i, int j) const;
    const std::vector<double>& segment() const { return segment_; }

   private:
    Order order_;
    std::vector<double> segment_;
};

// GEMM-route subset of DispatchPolicy (DispatchPolicy.java:57-679) with HIP gates.
class DispatchPolicy {
   public:
    // DispatchPolicy.Algorithm (DispatchPolicy.java:58-67) with HIP beside CUDA.
    enum class Algorithm { NAIVE, BLOCKED, BLAS3, SIMD, CUDA, PARALLEL, STRASSEN, SPECIALIZED, HIP };
    static const char* name(Algorithm a);
    class Builder;
    static Builder builder();
    // Default policy: the HIP gate comes from the library's tuning table.
    static DispatchPolicy defaultPolicy();
    int getParallelism() const { return parallelism_; }
    int mode() const { return mode_; }
    bool shouldOffloadToHip(int64_t m, int64_t n, int64_t k) const;
    // selectForMultiply / selectCpuAlgorithm (DispatchPolicy.java:179-193): selectBaseAlgorithm
    // (:254-279) with the HIP gate in the CUDA slot.
    Algorithm selectForMultiply(int m, int n, int k) const { return selectBase(m, n, k, true); }
    Algorithm selectCpuAlgorithm(int m, int n, int k) const { return selectBase(m, n, k, false); }

   private:
    Algorithm selectBase(int m, int n, int k, bool allowHip) const;
    bool enable_hip_ = true;
    int64_t hip_min_dim_ = 0;
    int64_t hip_min_elements_ = 0;
    int64_t hip_min_flops_ = 200000000;
    int parallelism_ = 1;
    int mode_ = 0;
    // Multiply-route thresholds: the reference builder's defaults (DispatchPolicy.java:424-437).
    int naive_threshold_ = 64, blocked_threshold_ = 256, strassen_threshold_ = 1024;
    int blas3_threshold_ = 256, simd_threshold_ = 256;
    bool enable_blas3_ = true, enable_simd_ = true, enable_strassen_ = false, simd_available_ = true;
    int hip_detected_ = -1;  // cudaDetector analogue: -1 asks HipSupport, 0 / 1 injected
};

class DispatchPolicy::Builder {
   public:
    Builder& enableHip(bool v) { p_.enable_hip_ = v; return *this; }
    Builder& hipMinDim(int64_t v) { p_.hip_min_dim_ = v; return *this; }
    Builder& hipMinElements(int64_t v) { p_.hip_min_elements_ = v; return *this; }
    Builder& hipMinFlops(int64_t v) { p_.hip_min_flops_ = v; return *this; }
    Builder& parallelism(int v) { p_.parallelism_ = v; return *this; }
    // Arithmetic form of the HIP kernel: JLC_MODE_FP64 / SPLIT_INT8 / SPLIT_BF16.
    Builder& mode(int v) { p_.mode_ = v; return *this; }
    Builder& naiveThreshold(int v) { p_.naive_threshold_ = v; return *this; }
    Builder& enableBlas3(bool v) { p_.enable_blas3_ = v; return *this; }
    Builder& simdAvailable(bool v) { p_.simd_available_ = v; return *this; }
    Builder& hipDetector(bool available) { p_.hip_detected_ = available ? 1 : 0; return *this; }
    DispatchPolicy build() const { return p_; }

   private:
    DispatchPolicy p_;
};

inline DispatchPolicy::Builder DispatchPolicy::builder() { return Builder(); }

class GemmDispatch {
   public:
    enum class Kernel { TINY, SMALL, MATVEC, MICROKERNEL, CUDA, PARALLEL_MICRO, HIP };
    static constexpr int64_t TINY_THRESHOLD = 8 * 8 * 8;
    static constexpr int64_t SMALL_THRESHOLD = 50000;
    static constexpr int64_t PARALLEL_THRESHOLD = 5000000;
    // selectKernel (GemmDispatch.java:76-106) with the HIP gate tried first.
    static Kernel selectKernel(int m, int n, int k, bool hipAvailable, int threads,
                               const DispatchPolicy& policy = DispatchPolicy::defaultPolicy());
    static const char* name(Kernel k);
};

// Detection (CudaSupport.java:34-123 analogue).
class HipSupport {
   public:
    static bool isHipAvailable();
    static void refresh();
};

// The HIP plug (CudaGemm.gemm contract, CudaGemm.java:53-135): throws
// IllegalArgumentException for null / shape errors, returns false for device failures.
class HipGemm {
   public:
    static bool gemm(const Matrix* a, const Matrix* b, Matrix* c, double alpha, double beta, int mode = 0);
    // CudaGemm.multiply (CudaGemm.java:32-41): a new heap Matrix A B, or null when the device call fails.
    static std::unique_ptr<Matrix> multiply(const Matrix* a, const Matrix* b, int mode = 0);
};

// BlockedMultiply.multiply (BlockedMultiply.java:43-105), the product behind Matrix.multiply: null /
// inner-dimension errors throw IllegalArgumentException; an empty product is a new zero matrix (an
// OffHeapMatrix when either input is one); above the HIP gate a new Matrix from HipGemm.multiply.
// n == 1 (multiplyVector) and the CPU algorithms throw UnsupportedKernel.
class BlockedMultiply {
   public:
    static std::unique_ptr<Matrix> multiply(const Matrix* a, const Matrix* b,
                                            const DispatchPolicy* policy = nullptr);
};

class OptimizedBLAS3 {
   public:
    // OptimizedBLAS3.gemm (OptimizedBLAS3.java:61-130).
    static void gemm(const Matrix* a, const Matrix* b, Matrix* c, double alpha, double beta,
                     const DispatchPolicy* policy);
    // OptimizedBLAS3.gemmStrided (OptimizedBLAS3.java:135-179); note the m, k, n order.
    static void gemmStrided(const double* a, int aOffset, int lda, const double* b, int bOffset, int ldb, double* c,
                            int cOffset, int ldc, int m, int k, int n, double alpha, double beta,
                            const DispatchPolicy* policy = nullptr);
};

class BLAS3Kernels {
   public:
    static Matrix gemm(const Matrix* a, const Matrix* b, const DispatchPolicy* policy = nullptr);
    static void gemm(const Matrix* a, const Matrix* b, Matrix* c, double alpha, double beta,
                     const DispatchPolicy* policy);
    static void gemmStrided(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb, double* cd,
                            int cOff, int ldc, int m, int k, int n, double alpha, double beta, int blockSize = 0);
    static void gemmStridedTransA(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb, double* cd,
                                  int cOff, int ldc, int m, int k, int n, double alpha, double beta,
                                  int blockSize = 0);
    static void gemmStridedColMajorA(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb,
                                     double* cd, int cOff, int ldc, int m, int k, int n, double alpha, double beta,
                                     int blockSize = 0);
    static void gemmStridedColMajorB(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb,
                                     double* cd, int cOff, int ldc, int m, int k, int n, double alpha, double beta,
                                     int blockSize = 0);
};

}  // namespace compute
}  // namespace jlc
