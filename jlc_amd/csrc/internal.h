// Internal declarations shared by the libjlc_hip translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jlc {

// Tile shape of a launch, reported back through jlc_stats.
struct TileInfo {
    int tile_m = 0, tile_n = 0, tile_k = 0;
    int launches = 0;
};

// C = beta*C on an m x n row-major block (beta == 0 writes zeros, C unread).
hipError_t launch_scale(int64_t m, int64_t n, double beta, double* c, int64_t ldc, hipStream_t stream,
                        TileInfo* info);

// FP64 MFMA GEMM, device pointers, layouts as in jlc_hip.h. Requires m, n, k > 0 and alpha != 0.
hipError_t launch_dgemm_f64(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha,
                            const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                            int64_t ldc, hipStream_t stream, TileInfo* info);

// y = alpha*A*x + beta*y, A row-major. Requires m > 0; k == 0 or alpha == 0 handled inside.
hipError_t launch_dgemv(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x,
                        int64_t incx, double beta, double* y, int64_t incy, hipStream_t stream, TileInfo* info);

// Device-to-device copy kernel (bytes may be any size).
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

}  // namespace jlc
