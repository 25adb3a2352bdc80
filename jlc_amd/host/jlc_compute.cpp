// C++ host mirror: implementation over the C ABI (include/jlc_hip.h).
#include "jlc_compute.hpp"

#include <algorithm>

#include <atomic>

#include "../../include/jlc_hip.h"

namespace jlc {
namespace compute {

namespace {

void check(int rc, const char* what) {
    if (rc != JLC_OK) throw DeviceError(rc, std::string(what) + ": " + jlc_last_error());
}

void validate(const Matrix* a, const Matrix* b, const Matrix* c, int& m, int& k, int& n) {
    if (!a || !b || !c) throw IllegalArgumentException("Matrices must not be null");
    m = a->getRowCount();
    k = a->getColumnCount();
    n = b->getColumnCount();
    if (k != b->getRowCount()) throw IllegalArgumentException("Inner dimensions must agree");
    if (m != c->getRowCount() || n != c->getColumnCount())
        throw IllegalArgumentException("Output dimensions must match");
}

void require_hip(GemmDispatch::Kernel k) {
    if (k != GemmDispatch::Kernel::HIP)
        throw UnsupportedKernel(std::string("dispatch selected the CPU regime ") + GemmDispatch::name(k) +
                                "; CPU regimes are the Java host's own code");
}

std::atomic<int> g_hip_cached{-1};

}  // namespace

Matrix::Matrix(int rows, int cols) : Matrix(rows, cols, std::vector<double>((size_t)rows * cols, 0.0)) {}

Matrix::Matrix(int rows, int cols, std::vector<double> data) : rows_(rows), cols_(cols), data_(std::move(data)) {
    if (rows < 0 || cols < 0) throw IllegalArgumentException("Matrix dimensions must be non-negative");
    if (data_.size() != (size_t)rows * cols) throw IllegalArgumentException("Data length must equal rows*cols");
}

Matrix Matrix::wrap(std::vector<double> data, int rows, int cols) { return Matrix(rows, cols, std::move(data)); }

OffHeapMatrix::OffHeapMatrix(int rows, int cols, Order order)
    : Matrix(rows, cols), order_(order), segment_((size_t)rows * cols, 0.0) {}

void OffHeapMatrix::syncToOffHeap() {
    for (int i = 0; i < rows_; i++)
        for (int j = 0; j < cols_; j++)
            segment_[order_ == Order::ROW_MAJOR ? (size_t)i * cols_ + j : (size_t)j * rows_ + i] = get(i, j);
}

void OffHeapMatrix::syncFromOffHeap() {
    for (int i = 0; i < rows_; i++)
        for (int j = 0; j < cols_; j++) set(i, j, getOffHeap(i, j));
}

double OffHeapMatrix::getOffHeap(int i, int j) const {
    return segment_[order_ == Order::ROW_MAJOR ? (size_t)i * cols_ + j : (size_t)j * rows_ + i];
}

DispatchPolicy DispatchPolicy::defaultPolicy() {
    DispatchPolicy p;
    double v = 0.0;
    if (jlc_tune_get("hip_min_flops", &v) == JLC_OK) p.hip_min_flops_ = (int64_t)v;
    return p;
}

bool DispatchPolicy::shouldOffloadToHip(int64_t m, int64_t n, int64_t k) const {
    // shouldOffloadToCuda analogue (DispatchPolicy.java:219-236): the smallest dimension against
    // the minimum, the element count of all three operands against the minimum, then the flops.
    if (!enable_hip_) return false;
    if (std::min(m, std::min(n, k)) < hip_min_dim_) return false;
    if (m * k + k * n + m * n < hip_min_elements_) return false;
    return 2 * m * n * k >= hip_min_flops_;
}

const char* DispatchPolicy::name(Algorithm a) {
    switch (a) {
        case Algorithm::NAIVE: return "NAIVE";
        case Algorithm::BLOCKED: return "BLOCKED";
        case Algorithm::BLAS3: return "BLAS3";
        case Algorithm::SIMD: return "SIMD";
        case Algorithm::CUDA: return "CUDA";
        case Algorithm::PARALLEL: return "PARALLEL";
        case Algorithm::STRASSEN: return "STRASSEN";
        case Algorithm::SPECIALIZED: return "SPECIALIZED";
        default: return "HIP";
    }
}

DispatchPolicy::Algorithm DispatchPolicy::selectBase(int m, int n, int k, bool allowHip) const {
    const int maxDim = std::max(m, std::max(n, k));
    const bool simd = enable_simd_ && maxDim < simd_threshold_ && simd_available_;
    if (maxDim <= naive_threshold_) return simd ? Algorithm::SIMD : Algorithm::NAIVE;
    const bool hip = hip_detected_ < 0 ? HipSupport::isHipAvailable() : hip_detected_ == 1;
    if (allowHip && hip && shouldOffloadToHip(m, n, k)) return Algorithm::HIP;
    if (enable_strassen_ && m == n && n == k && maxDim >= strassen_threshold_) return Algorithm::STRASSEN;
    if (enable_blas3_ && maxDim >= blas3_threshold_) return Algorithm::BLAS3;
    if (simd) return Algorithm::SIMD;
    if (maxDim >= blocked_threshold_) return Algorithm::BLOCKED;
    return Algorithm::NAIVE;
}

GemmDispatch::Kernel GemmDispatch::selectKernel(int m, int n, int k, bool hipAvailable, int threads,
                                                const DispatchPolicy& policy) {
    if (hipAvailable && policy.shouldOffloadToHip(m, n, k)) return Kernel::HIP;
    const int64_t flops = 2LL * m * n * k;
    if (n == 1) return Kernel::MATVEC;
    if (flops <= TINY_THRESHOLD && m <= 8 && n <= 8 && k <= 8) return Kernel::TINY;
    if (flops <= SMALL_THRESHOLD) return Kernel::SMALL;
    if (threads > 1 && flops >= PARALLEL_THRESHOLD) return Kernel::PARALLEL_MICRO;
    return Kernel::MICROKERNEL;
}

const char* GemmDispatch::name(Kernel k) {
    switch (k) {
        case Kernel::TINY: return "TINY";
        case Kernel::SMALL: return "SMALL";
        case Kernel::MATVEC: return "MATVEC";
        case Kernel::MICROKERNEL: return "MICROKERNEL";
        case Kernel::CUDA: return "CUDA";
        case Kernel::PARALLEL_MICRO: return "PARALLEL_MICRO";
        default: return "HIP";
    }
}

bool HipSupport::isHipAvailable() {
    int cached = g_hip_cached.load();
    if (cached < 0) {
        int count = 0;
        cached = (jlc_hip_probe(&count, nullptr, 0) == JLC_OK && count > 0) ? 1 : 0;
        g_hip_cached.store(cached);
    }
    return cached == 1;
}

void HipSupport::refresh() { g_hip_cached.store(-1); }

bool HipGemm::gemm(const Matrix* a, const Matrix* b, Matrix* c, double alpha, double beta, int mode) {
    int m, k, n;
    validate(a, b, c, m, k, n);
    if (m == 0 || n == 0) return true;
    const int rc = jlc_dgemm(JLC_ROW, JLC_ROW, m, n, k, alpha, a->getRawData().data(), k > 0 ? k : 1,
                             b->getRawData().data(), n, beta, c->getRawData().data(), n, mode, nullptr);
    return rc == JLC_OK;
}

std::unique_ptr<Matrix> HipGemm::multiply(const Matrix* a, const Matrix* b, int mode) {
    auto c = std::make_unique<Matrix>(a->getRowCount(), b->getColumnCount());
    if (!gemm(a, b, c.get(), 1.0, 0.0, mode)) return nullptr;
    return c;
}

std::unique_ptr<Matrix> BlockedMultiply::multiply(const Matrix* a, const Matrix* b, const DispatchPolicy* policy) {
    if (!a || !b) throw IllegalArgumentException("Matrices must not be null");
    const int m = a->getRowCount(), k = a->getColumnCount(), k2 = b->getRowCount(), n = b->getColumnCount();
    if (k != k2)
        throw IllegalArgumentException("Inner dimensions must agree for multiplication: " + std::to_string(k) +
                                       " != " + std::to_string(k2));
    if (m == 0 || n == 0 || k == 0) {
        if (dynamic_cast<const OffHeapMatrix*>(a) || dynamic_cast<const OffHeapMatrix*>(b))
            return std::make_unique<OffHeapMatrix>(m, n);
        return std::make_unique<Matrix>(m, n);
    }
    if (n == 1)
        throw UnsupportedKernel("multiply: n == 1 runs multiplyVector, the Java host's CPU matvec");
    const DispatchPolicy active = policy ? *policy : DispatchPolicy::defaultPolicy();
    DispatchPolicy::Algorithm algorithm = active.selectForMultiply(m, n, k);
    if (algorithm == DispatchPolicy::Algorithm::HIP) {
        if (auto c = HipGemm::multiply(a, b, active.mode())) return c;
        algorithm = active.selectCpuAlgorithm(m, n, k);
        throw UnsupportedKernel(std::string("multiply: the HIP call failed (") + jlc_last_error() +
                                "); the reference then runs its CPU algorithm " + DispatchPolicy::name(algorithm));
    }
    throw UnsupportedKernel(std::string("multiply: dispatch selected the CPU algorithm ") +
                            DispatchPolicy::name(algorithm));
}

std::unique_ptr<Matrix> Matrix::multiply(const Matrix& other) const { return BlockedMultiply::multiply(this, &other); }

static void hip_or_throw(const Matrix* a, const Matrix* b, Matrix* c, double alpha, double beta, int mode) {
    if (!HipGemm::gemm(a, b, c, alpha, beta, mode)) throw DeviceError(JLC_ERR_HIP, jlc_last_error());
}

void OptimizedBLAS3::gemm(const Matrix* a, const Matrix* b, Matrix* c, double alpha, double beta,
                          const DispatchPolicy* policy) {
    int m, k, n;
    validate(a, b, c, m, k, n);
    if (m == 0 || n == 0) return;
    const DispatchPolicy pol = policy ? *policy : DispatchPolicy::defaultPolicy();
    require_hip(GemmDispatch::selectKernel(m, n, k, HipSupport::isHipAvailable(), pol.getParallelism(), pol));
    hip_or_throw(a, b, c, alpha, beta, pol.mode());
    if (auto* off = dynamic_cast<OffHeapMatrix*>(c)) off->syncToOffHeap();
}

void OptimizedBLAS3::gemmStrided(const double* a, int aOffset, int lda, const double* b, int bOffset, int ldb,
                                 double* c, int cOffset, int ldc, int m, int k, int n, double alpha, double beta,
                                 const DispatchPolicy* policy) {
    if (m <= 0 || n <= 0) return;
    const DispatchPolicy pol = policy ? *policy : DispatchPolicy::defaultPolicy();
    require_hip(GemmDispatch::selectKernel(m, n, k, HipSupport::isHipAvailable(), 1, pol));
    check(jlc_dgemm(JLC_ROW, JLC_ROW, m, n, k, alpha, a + aOffset, lda, b + bOffset, ldb, beta, c + cOffset, ldc,
                    pol.mode(), nullptr),
          "OptimizedBLAS3.gemmStrided");
}

Matrix BLAS3Kernels::gemm(const Matrix* a, const Matrix* b, const DispatchPolicy* policy) {
    if (!a || !b) throw IllegalArgumentException("Matrices must not be null");
    if (a->getColumnCount() != b->getRowCount()) throw IllegalArgumentException("Inner dimensions must agree");
    Matrix c(a->getRowCount(), b->getColumnCount());
    gemm(a, b, &c, 1.0, 0.0, policy);
    return c;
}

void BLAS3Kernels::gemm(const Matrix* a, const Matrix* b, Matrix* c, double alpha, double beta,
                        const DispatchPolicy* policy) {
    int m, k, n;
    validate(a, b, c, m, k, n);
    if (m == 0 || n == 0) return;
    const DispatchPolicy pol = policy ? *policy : DispatchPolicy::defaultPolicy();
    require_hip(GemmDispatch::selectKernel(m, n, k, HipSupport::isHipAvailable(), pol.getParallelism(), pol));
    hip_or_throw(a, b, c, alpha, beta, pol.mode());
    if (auto* off = dynamic_cast<OffHeapMatrix*>(c)) off->syncToOffHeap();
}

void BLAS3Kernels::gemmStrided(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb, double* cd,
                               int cOff, int ldc, int m, int k, int n, double alpha, double beta, int) {
    if (m <= 0 || n <= 0) return;
    check(jlc_dgemm(JLC_ROW, JLC_ROW, m, n, k, alpha, ad + aOff, lda, bd + bOff, ldb, beta, cd + cOff, ldc, 0,
                    nullptr),
          "BLAS3Kernels.gemmStrided");
}

void BLAS3Kernels::gemmStridedTransA(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb,
                                     double* cd, int cOff, int ldc, int m, int k, int n, double alpha, double beta,
                                     int) {
    // C (k x n) = alpha * A^T B + beta C with A stored m x k (BLAS3Kernels.java:253-279).
    if (k <= 0 || n <= 0) return;
    check(jlc_dgemm(JLC_COL, JLC_ROW, k, n, m, alpha, ad + aOff, lda, bd + bOff, ldb, beta, cd + cOff, ldc, 0,
                    nullptr),
          "BLAS3Kernels.gemmStridedTransA");
}

void BLAS3Kernels::gemmStridedColMajorA(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb,
                                        double* cd, int cOff, int ldc, int m, int k, int n, double alpha,
                                        double beta, int) {
    if (m <= 0 || n <= 0) return;
    check(jlc_dgemm(JLC_COL, JLC_ROW, m, n, k, alpha, ad + aOff, lda, bd + bOff, ldb, beta, cd + cOff, ldc, 0,
                    nullptr),
          "BLAS3Kernels.gemmStridedColMajorA");
}

void BLAS3Kernels::gemmStridedColMajorB(const double* ad, int aOff, int lda, const double* bd, int bOff, int ldb,
                                        double* cd, int cOff, int ldc, int m, int k, int n, double alpha,
                                        double beta, int) {
    if (m <= 0 || n <= 0) return;
    check(jlc_dgemm(JLC_ROW, JLC_COL, m, n, k, alpha, ad + aOff, lda, bd + bOff, ldb, beta, cd + cOff, ldc, 0,
                    nullptr),
          "BLAS3Kernels.gemmStridedColMajorB");
}

}  // namespace compute
}  // namespace jlc
