// Parity tests for the C++ host mirror (jlc_amd/host/jlc_compute.hpp), written
// after the reference's JUnit suites for the GEMM path:
//   GemmParityTest.java:25-72, :118-156     shape x alpha/beta grid, degenerate dims, messages
//   GemmStridedParityTest.java:7-130         strided row-major, transposed A, column-major A/B
//   GemmOffHeapParityTest.java:21-78         off-heap output sync
//   GemmNumericEdgeTest.java:32-95           known answers
//   DispatchPolicyTest.java:18-138           selection (detector injected as a flag)
// Expected values come from the oracle library (oracle/oracle.h); the product
// library is reached only through the mirror classes.
//
// Usage: gemm_parity_test [--cpu]   (--cpu: tests that need no device)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../jlc_amd/host/jlc_compute.hpp"
#include "../../oracle/oracle.h"

using namespace jlc::compute;

static int g_failed = 0, g_passed = 0;

#define CHECK(cond, ...)                                        \
    do {                                                        \
        if (!(cond)) {                                          \
            std::printf("  FAIL %s:%d: ", __FILE__, __LINE__);  \
            std::printf(__VA_ARGS__);                           \
            std::printf("\n");                                  \
            throw std::runtime_error("check failed");           \
        }                                                       \
    } while (0)

static void run(const char* name, const std::function<void()>& fn) {
    std::printf("[ RUN  ] %s\n", name);
    std::fflush(stdout);
    try {
        fn();
        g_passed++;
        std::printf("[  OK  ] %s\n", name);
    } catch (const std::exception& e) {
        g_failed++;
        std::printf("[ FAIL ] %s: %s\n", name, e.what());
    }
    std::fflush(stdout);
}

static std::vector<double> seeded(int64_t len, int64_t seed) {
    std::vector<double> v((size_t)len);
    if (len) oracle_seeded_array(v.data(), len, seed);
    return v;
}

// assertParity (GemmReference.java:148-173): relative Frobenius error <= relTol and
// max abs error <= absTol (cpuAbsTolerance, GemmReference.java:175-177).
static void assert_parity(const std::vector<double>& expected, const std::vector<double>& actual, int64_t k,
                          const char* label) {
    CHECK(expected.size() == actual.size(), "%s: length", label);
    double max_abs = 0, rel = 0;
    oracle_parity(expected.data(), actual.data(), (int64_t)expected.size(), &max_abs, &rel);
    const double abs_tol = oracle_cpu_abs_tol(expected.data(), (int64_t)expected.size(), k);
    CHECK(rel <= 1e-12 && max_abs <= abs_tol, "%s: rel %.3e max_abs %.3e tol %.3e", label, rel, max_abs, abs_tol);
}

static const DispatchPolicy HIP_POLICY = DispatchPolicy::builder().enableHip(true).hipMinFlops(0).parallelism(1).build();

template <class F>
static bool throws_iae(F f, std::string* msg = nullptr) {
    try {
        f();
    } catch (const IllegalArgumentException& e) {
        if (msg) *msg = e.what();
        return true;
    }
    return false;
}

static void cpu_tests() {
    run("DispatchPolicy.selection_reference_order_without_hip", [] {
        const auto pol = DispatchPolicy::builder().parallelism(4).build();
        using K = GemmDispatch::Kernel;
        CHECK(GemmDispatch::selectKernel(96, 1, 96, false, 4, pol) == K::MATVEC, "matvec");
        CHECK(GemmDispatch::selectKernel(4, 4, 4, false, 4, pol) == K::TINY, "tiny");
        CHECK(GemmDispatch::selectKernel(8, 8, 8, false, 4, pol) == K::SMALL, "8^3 is SMALL (1024 flops)");
        CHECK(GemmDispatch::selectKernel(12, 12, 12, false, 4, pol) == K::SMALL, "small");
        CHECK(GemmDispatch::selectKernel(64, 64, 64, false, 1, pol) == K::MICROKERNEL, "micro at parallelism 1");
        CHECK(GemmDispatch::selectKernel(256, 256, 256, false, 4, pol) == K::PARALLEL_MICRO, "parallel");
    });
    run("DispatchPolicy.hip_gate", [] {
        using K = GemmDispatch::Kernel;
        const auto pol = DispatchPolicy::builder().hipMinFlops(200000000).parallelism(4).build();
        CHECK(GemmDispatch::selectKernel(512, 512, 512, true, 4, pol) == K::HIP, "512^3 offloads");
        CHECK(GemmDispatch::selectKernel(256, 256, 256, true, 4, pol) == K::PARALLEL_MICRO, "256^3 stays");
        CHECK(GemmDispatch::selectKernel(512, 512, 512, false, 4, pol) == K::PARALLEL_MICRO, "no device");
        const auto off = DispatchPolicy::builder().enableHip(false).hipMinFlops(0).build();
        CHECK(GemmDispatch::selectKernel(512, 512, 512, true, 1, off) == K::MICROKERNEL, "disabled");
        const auto dim = DispatchPolicy::builder().hipMinFlops(0).hipMinDim(64).build();
        CHECK(GemmDispatch::selectKernel(32, 128, 128, true, 1, dim) == K::MICROKERNEL, "min dim");
        // shouldOffloadToCuda (DispatchPolicy.java:226-235) applies the minimum dimension to n == 1 too.
        CHECK(GemmDispatch::selectKernel(128, 1, 128, true, 1, dim) == K::MATVEC, "matvec under min dim");
        // The element gate counts all three operands, m*k + k*n + m*n: a long-k product with a small C
        // passes it (8*100000 + 100000*8 + 64 = 1,600,064 elements) ...
        const auto elems = DispatchPolicy::builder().hipMinFlops(0).hipMinElements(1000000).build();
        CHECK(GemmDispatch::selectKernel(8, 8, 100000, true, 1, elems) == K::HIP, "long-k elements");
        // ... and a product whose three operands hold fewer elements does not (3 * 512^2 = 786,432).
        CHECK(GemmDispatch::selectKernel(512, 512, 512, true, 4, elems) == K::PARALLEL_MICRO, "elements gate");
        CHECK(GemmDispatch::selectKernel(4, 4, 4, true, 1, HIP_POLICY) == K::HIP, "gate 0 routes all");
    });
    run("GemmParityTest.rejects_null_and_mismatch_with_message", [] {
        Matrix a(3, 4), b(5, 2), c(3, 2), b2(4, 2), c2(2, 2);
        std::string msg;
        CHECK(throws_iae([&] { BLAS3Kernels::gemm(nullptr, &b2, &c, 1, 0, &HIP_POLICY); }, &msg), "null");
        CHECK(!msg.empty(), "message");
        CHECK(throws_iae([&] { BLAS3Kernels::gemm(&a, &b, &c, 1, 0, &HIP_POLICY); }, &msg), "inner");
        CHECK(!msg.empty(), "message");
        CHECK(throws_iae([&] { OptimizedBLAS3::gemm(&a, &b2, &c2, 1, 0, &HIP_POLICY); }, &msg), "output");
        CHECK(!msg.empty(), "message");
        CHECK(throws_iae([&] { HipGemm::gemm(&a, &b, &c, 1, 0); }, &msg), "HipGemm inner");
        CHECK(throws_iae([&] { BLAS3Kernels::gemm(&a, &b); }), "facade inner");
        CHECK(throws_iae([] { Matrix bad(2, 2, std::vector<double>(3)); }), "data length");
    });
    run("GemmParityTest.empty_outputs_need_no_device", [] {
        Matrix a(0, 5), b(5, 3), c(0, 3);
        BLAS3Kernels::gemm(&a, &b, &c, 1.0, 0.0, &HIP_POLICY);
        OptimizedBLAS3::gemm(&a, &b, &c, 1.0, 0.0, &HIP_POLICY);
        Matrix a2(4, 5), b2(5, 0), c2(4, 0);
        BLAS3Kernels::gemm(&a2, &b2, &c2, 2.0, 1.0, &HIP_POLICY);
        CHECK(BLAS3Kernels::gemm(&a, &b, &HIP_POLICY).getRowCount() == 0, "shape");
    });
    // BlockedMultiply.multiply facade (BlockedMultiply.java:55-105) and selectBaseAlgorithm
    // (DispatchPolicy.java:254-279), without a device.
    run("BlockedMultiply.facade_and_selection", [] {
        using A = DispatchPolicy::Algorithm;
        std::string msg;
        Matrix a(2, 3), b(4, 2);
        CHECK(throws_iae([&] { BlockedMultiply::multiply(nullptr, &b); }, &msg), "null");
        CHECK(throws_iae([&] { a.multiply(b); }, &msg), "inner");
        CHECK(msg.find("3 != 4") != std::string::npos, "message names the dimensions: %s", msg.c_str());
        const auto cpu = DispatchPolicy::builder().enableHip(false).build();
        Matrix e1(3, 0), e2(0, 5);
        auto z = BlockedMultiply::multiply(&e1, &e2, &cpu);
        CHECK(z->getRowCount() == 3 && z->getColumnCount() == 5 && !dynamic_cast<OffHeapMatrix*>(z.get()),
              "empty heap product");
        for (double v : z->getRawData()) CHECK(v == 0.0, "zero entries");
        OffHeapMatrix o1(0, 4);
        Matrix m2(4, 2);
        CHECK(dynamic_cast<OffHeapMatrix*>(BlockedMultiply::multiply(&o1, &m2, &cpu).get()) != nullptr,
              "empty product with an off-heap input is off-heap");
        Matrix v1(300, 300), v2(300, 1);
        bool unsupported = false;
        try {
            BlockedMultiply::multiply(&v1, &v2, &cpu);
        } catch (const UnsupportedKernel&) {
            unsupported = true;
        }
        CHECK(unsupported, "n == 1 is the host's matvec");
        const auto on = DispatchPolicy::builder().hipMinFlops(0).hipDetector(true).build();
        const auto off = DispatchPolicy::builder().hipMinFlops(0).hipDetector(false).build();
        CHECK(on.selectForMultiply(64, 64, 64) == A::SIMD, "naive threshold before the HIP gate");
        CHECK(on.selectForMultiply(65, 65, 65) == A::HIP, "HIP above the naive threshold");
        CHECK(off.selectForMultiply(65, 65, 65) == A::SIMD, "SIMD below 256");
        CHECK(off.selectForMultiply(300, 300, 300) == A::BLAS3, "BLAS3 from 256");
        CHECK(on.selectCpuAlgorithm(300, 300, 300) == A::BLAS3, "CPU choice ignores HIP");
        const auto plain = DispatchPolicy::builder().hipDetector(false).simdAvailable(false).enableBlas3(false).build();
        CHECK(plain.selectForMultiply(100, 100, 100) == A::NAIVE, "naive below 256");
        CHECK(plain.selectForMultiply(300, 100, 100) == A::BLOCKED, "blocked from 256");
    });
    run("OffHeapMatrix.sync_orders", [] {
        for (auto order : {OffHeapMatrix::Order::ROW_MAJOR, OffHeapMatrix::Order::COLUMN_MAJOR}) {
            OffHeapMatrix m(3, 5, order);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 5; j++) m.set(i, j, 10 * i + j);
            m.syncToOffHeap();
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 5; j++) CHECK(m.getOffHeap(i, j) == 10 * i + j, "offheap (%d,%d)", i, j);
            CHECK(m.segment()[1] == (order == OffHeapMatrix::Order::ROW_MAJOR ? 1.0 : 10.0), "storage order");
        }
    });
}

static void gpu_tests() {
    run("HipSupport.device_present", [] { CHECK(HipSupport::isHipAvailable(), "no HIP device"); });

    // GemmParityTest.java:25-35, :39-72.
    run("GemmParityTest.shape_alpha_beta_grid_both_facades", [] {
        const int shapes[][3] = {{4, 4, 4}, {12, 12, 12}, {64, 64, 64}, {128, 64, 16},
                                 {16, 64, 128}, {128, 3, 128}, {96, 96, 1}};
        const double ab[] = {0.0, 1.0, -0.5, 2.5};
        int64_t seed = 1000;
        for (auto& s : shapes) {
            const int m = s[0], k = s[1], n = s[2];
            for (double alpha : ab)
                for (double beta : ab) {
                    auto av = seeded((int64_t)m * k, seed++), bv = seeded((int64_t)k * n, seed++);
                    auto cv = seeded((int64_t)m * n, seed++);
                    std::vector<double> expected = cv;
                    oracle_gemm(av.data(), bv.data(), expected.data(), m, k, n, alpha, beta);
                    Matrix a(m, k, av), b(k, n, bv), c1(m, n, cv), c2(m, n, cv);
                    BLAS3Kernels::gemm(&a, &b, &c1, alpha, beta, &HIP_POLICY);
                    OptimizedBLAS3::gemm(&a, &b, &c2, alpha, beta, &HIP_POLICY);
                    char label[96];
                    std::snprintf(label, sizeof label, "%dx%dx%d a=%g b=%g", m, k, n, alpha, beta);
                    assert_parity(expected, c1.getRawData(), k, label);
                    assert_parity(expected, c2.getRawData(), k, label);
                }
        }
    });

    // GemmParityTest.java:126-141: k = 0 gives C = beta*C exactly.
    run("GemmParityTest.k_zero_is_beta_scale", [] {
        Matrix a(3, 0), b(0, 4);
        auto cv = seeded(12, 77);
        Matrix c(3, 4, cv);
        BLAS3Kernels::gemm(&a, &b, &c, 1.5, -0.5, &HIP_POLICY);
        for (int i = 0; i < 12; i++) CHECK(c.getRawData()[i] == -0.5 * cv[i], "entry %d", i);
    });

    // BLAS3Kernels.gemm(A,B) facade (BLAS3Kernels.java:35-63).
    run("BLAS3Kernels.facade_returns_product", [] {
        auto av = seeded(40 * 24, 5), bv = seeded(24 * 33, 6);
        Matrix a(40, 24, av), b(24, 33, bv);
        Matrix c = BLAS3Kernels::gemm(&a, &b, &HIP_POLICY);
        std::vector<double> expected(40 * 33, 0.0);
        oracle_gemm(av.data(), bv.data(), expected.data(), 40, 24, 33, 1.0, 0.0);
        assert_parity(expected, c.getRawData(), 24, "facade");
    });

    // GemmStridedParityTest.java:7-29 (offsets, lda = k + 5, ...).
    run("GemmStridedParityTest.row_major_strided", [] {
        const int m = 37, k = 29, n = 41, aOff = 13, bOff = 7, cOff = 11, lda = k + 5, ldb = n + 3, ldc = n + 2;
        auto av = seeded(aOff + (int64_t)m * lda, 21), bv = seeded(bOff + (int64_t)k * ldb, 22);
        auto cv = seeded(cOff + (int64_t)m * ldc, 23);
        auto expected = cv, actual = cv;
        oracle_gemm_strided(av.data(), aOff, lda, bv.data(), bOff, ldb, expected.data(), cOff, ldc, m, k, n, 1.25,
                            -0.75);
        BLAS3Kernels::gemmStrided(av.data(), aOff, lda, bv.data(), bOff, ldb, actual.data(), cOff, ldc, m, k, n,
                                  1.25, -0.75);
        assert_parity(expected, actual, k, "strided");
    });

    // GemmStridedParityTest.java:32-54: C (k x n) = alpha A^T B + beta C, A stored m x k.
    run("GemmStridedParityTest.transposed_a", [] {
        const int m = 45, k = 23, n = 31, aOff = 5, bOff = 9, cOff = 3, lda = k + 4, ldb = n + 6, ldc = n + 1;
        auto av = seeded(aOff + (int64_t)m * lda, 31), bv = seeded(bOff + (int64_t)m * ldb, 32);
        auto cv = seeded(cOff + (int64_t)k * ldc, 33);
        auto expected = cv, actual = cv;
        oracle_gemm_strided_trans_a(av.data(), aOff, lda, bv.data(), bOff, ldb, expected.data(), cOff, ldc, m, k, n,
                                    0.5, 2.0);
        BLAS3Kernels::gemmStridedTransA(av.data(), aOff, lda, bv.data(), bOff, ldb, actual.data(), cOff, ldc, m, k,
                                        n, 0.5, 2.0);
        assert_parity(expected, actual, m, "transA");
    });

    // GemmStridedParityTest.java:57-107.
    run("GemmStridedParityTest.column_major_a_and_b", [] {
        const int m = 33, k = 27, n = 35, aOff = 4, bOff = 6, cOff = 2;
        {
            const int lda = m + 3, ldb = n + 2, ldc = n + 5;
            auto av = seeded(aOff + (int64_t)k * lda, 41), bv = seeded(bOff + (int64_t)k * ldb, 42);
            auto cv = seeded(cOff + (int64_t)m * ldc, 43);
            auto expected = cv, actual = cv;
            oracle_gemm_strided_col_major_a(av.data(), aOff, lda, bv.data(), bOff, ldb, expected.data(), cOff, ldc,
                                            m, k, n, -1.5, 0.25);
            BLAS3Kernels::gemmStridedColMajorA(av.data(), aOff, lda, bv.data(), bOff, ldb, actual.data(), cOff, ldc,
                                               m, k, n, -1.5, 0.25);
            assert_parity(expected, actual, k, "colMajorA");
        }
        {
            const int lda = k + 2, ldb = k + 7, ldc = n + 1;
            auto av = seeded(aOff + (int64_t)m * lda, 51), bv = seeded(bOff + (int64_t)n * ldb, 52);
            auto cv = seeded(cOff + (int64_t)m * ldc, 53);
            auto expected = cv, actual = cv;
            oracle_gemm_strided_col_major_b(av.data(), aOff, lda, bv.data(), bOff, ldb, expected.data(), cOff, ldc,
                                            m, k, n, 2.0, 1.0);
            BLAS3Kernels::gemmStridedColMajorB(av.data(), aOff, lda, bv.data(), bOff, ldb, actual.data(), cOff, ldc,
                                               m, k, n, 2.0, 1.0);
            assert_parity(expected, actual, k, "colMajorB");
        }
    });

    // GemmStridedParityTest.java:110-130 (13 x 9 x 8) and a size that would reach
    // the reference's MICROKERNEL branch; the documented semantics
    // (GemmReference.java:46-61) are the expected values for both.
    run("GemmStridedParityTest.optimized_strided", [] {
        const int dims[][3] = {{13, 9, 8}, {160, 96, 130}};
        int64_t seed = 61;
        for (auto& d : dims) {
            const int m = d[0], k = d[1], n = d[2], aOff = 3, bOff = 5, cOff = 7, lda = k + 2, ldb = n + 3,
                      ldc = n + 4;
            auto av = seeded(aOff + (int64_t)m * lda, seed++), bv = seeded(bOff + (int64_t)k * ldb, seed++);
            auto cv = seeded(cOff + (int64_t)m * ldc, seed++);
            auto expected = cv, actual = cv;
            oracle_gemm_strided(av.data(), aOff, lda, bv.data(), bOff, ldb, expected.data(), cOff, ldc, m, k, n,
                                1.0, 0.5);
            OptimizedBLAS3::gemmStrided(av.data(), aOff, lda, bv.data(), bOff, ldb, actual.data(), cOff, ldc, m, k,
                                        n, 1.0, 0.5, &HIP_POLICY);
            assert_parity(expected, actual, k, "optimizedStrided");
        }
    });

    // GemmOffHeapParityTest.java:38-78.
    run("GemmOffHeapParityTest.output_sync", [] {
        const int m = 48, k = 40, n = 36;
        auto av = seeded(m * k, 71), bv = seeded(k * n, 72);
        Matrix a(m, k, av), b(k, n, bv);
        OffHeapMatrix c(m, n, OffHeapMatrix::Order::COLUMN_MAJOR);
        OptimizedBLAS3::gemm(&a, &b, &c, 1.0, 0.0, &HIP_POLICY);
        std::vector<double> expected((size_t)m * n, 0.0);
        oracle_gemm(av.data(), bv.data(), expected.data(), m, k, n, 1.0, 0.0);
        assert_parity(expected, c.getRawData(), k, "offheap heap mirror");
        const int probes[][2] = {{0, 0}, {m - 1, n - 1}, {7, 29}, {31, 3}};
        for (auto& p : probes)
            CHECK(c.getOffHeap(p[0], p[1]) == c.get(p[0], p[1]), "segment (%d,%d)", p[0], p[1]);
    });

    // GemmNumericEdgeTest.java:32-95 known answers.
    run("GemmNumericEdgeTest.known_answers", [] {
        Matrix a(2, 2, {1, 2, 3, 4}), b(2, 2, {5, 6, 7, 8}), c(2, 2);
        OptimizedBLAS3::gemm(&a, &b, &c, 1.0, 0.0, &HIP_POLICY);
        const double want[] = {19, 22, 43, 50};
        for (int i = 0; i < 4; i++) CHECK(c.getRawData()[i] == want[i], "2x2 entry %d", i);

        Matrix c2(2, 2, {2, -6, 10, -14});
        BLAS3Kernels::gemm(&a, &b, &c2, 0.0, 2.5, &HIP_POLICY);
        const double want2[] = {5, -15, 25, -35};
        for (int i = 0; i < 4; i++) CHECK(c2.getRawData()[i] == want2[i], "alpha=0 entry %d", i);

        Matrix an(1, 1, {NAN}), bn(1, 1, {NAN}), c3(1, 1, {4.0});
        OptimizedBLAS3::gemm(&an, &bn, &c3, 0.0, -0.5, &HIP_POLICY);
        CHECK(c3.get(0, 0) == -2.0, "alpha=0 drops NaN product: %g", c3.get(0, 0));

        Matrix az(1, 1, {0.0}), bi(1, 1, {INFINITY}), c4(1, 1, {7.0});
        OptimizedBLAS3::gemm(&az, &bi, &c4, 1.0, 1.0, &HIP_POLICY);
        CHECK(std::isnan(c4.get(0, 0)), "0*inf is NaN: %g", c4.get(0, 0));
    });

    // GemmLargeParityTest.java:19-45.
    run("GemmLargeParityTest.512_and_768x512x256", [] {
        const int dims[][3] = {{512, 512, 512}, {768, 512, 256}};
        int64_t seed = 8001;
        for (auto& d : dims) {
            const int m = d[0], k = d[1], n = d[2];
            auto av = seeded((int64_t)m * k, seed++), bv = seeded((int64_t)k * n, seed++);
            Matrix a(m, k, av), b(k, n, bv), c(m, n);
            OptimizedBLAS3::gemm(&a, &b, &c, 1.0, 0.0, &HIP_POLICY);
            std::vector<double> expected((size_t)m * n, 0.0);
            oracle_gemm(av.data(), bv.data(), expected.data(), m, k, n, 1.0, 0.0);
            assert_parity(expected, c.getRawData(), k, "large");
        }
    });

    // Matrix.multiply -> BlockedMultiply.multiply -> HipGemm.multiply (Matrix.java:1120-1122,
    // BlockedMultiply.java:79-85, CudaGemm.java:32-41): a new heap Matrix, off-heap inputs read
    // through their heap mirrors.
    run("BlockedMultiply.multiply_through_hip", [] {
        const int m = 300, k = 250, n = 200;
        auto av = seeded((int64_t)m * k, 9101), bv = seeded((int64_t)k * n, 9102);
        Matrix a(m, k, av), b(k, n, bv);
        std::vector<double> expected((size_t)m * n, 0.0);
        oracle_gemm(av.data(), bv.data(), expected.data(), m, k, n, 1.0, 0.0);
        auto c = a.multiply(b);
        CHECK(!dynamic_cast<OffHeapMatrix*>(c.get()), "heap result");
        assert_parity(expected, c->getRawData(), k, "Matrix.multiply");
        OffHeapMatrix oa(m, k);
        oa.getRawData() = av;
        auto c2 = BlockedMultiply::multiply(&oa, &b, &HIP_POLICY);
        CHECK(c2->getRowCount() == m && c2->getColumnCount() == n, "shape");
        assert_parity(expected, c2->getRawData(), k, "BlockedMultiply.multiply off-heap input");
    });

    // Default policy: the gate read from the tuning table sends 1024^3 to HIP.
    run("DispatchPolicy.default_policy_offloads_large", [] {
        CHECK(GemmDispatch::selectKernel(1024, 1024, 1024, HipSupport::isHipAvailable(), 4,
                                         DispatchPolicy::defaultPolicy()) == GemmDispatch::Kernel::HIP,
              "1024^3 under the default policy");
    });
}

int main(int argc, char** argv) {
    const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    cpu_tests();
    if (!cpu_only) gpu_tests();
    std::printf("%d passed, %d failed\n", g_passed, g_failed);
    return g_failed == 0 ? 0 : 1;
}
