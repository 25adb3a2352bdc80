// Bandwidth-bound FP64 matvec for gfx950: y = alpha*A*x + beta*y.
//
// Device body of the MATVEC regime (SpecializedKernels.matvec,
// src/main/java/net/faulj/compute/SpecializedKernels.java:27-71): beta
// scales y first (beta == 0 writes zero without reading y), alpha == 0 or
// k == 0 leaves beta*y, otherwise y += alpha * dot(A row, x) with alpha applied
// after the dot product, as the reference does.
//
// A streams through once, so the kernel is HBM-bound (8 bytes of A per 2
// flops). Each wavefront owns ROWS consecutive rows and walks K with 16-byte
// loads, two iterations in flight, so one x fragment serves ROWS rows; the
// partial sums are reduced across the 64 lanes with cross-lane shuffles.
#include "internal.h"

namespace jlc {
namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int ROWS = 4;     // rows per wavefront
constexpr int WAVES = 4;    // wavefronts per workgroup

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <bool VEC>
__global__ __launch_bounds__(64 * WAVES) void dgemv_rows(int M, int K, double alpha, const double* __restrict__ A,
                                                         int64_t lda, const double* __restrict__ x, int64_t incx,
                                                         double beta, double* __restrict__ y, int64_t incy) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * WAVES + wave) * ROWS;
    if (row0 >= M) return;
    const int nrows = min(ROWS, M - row0);

    double acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) acc[r] = 0.0;

    if (alpha != 0.0) {
        if (VEC) {
            // 2 doubles per lane per step, 128 per wavefront, two steps per iteration.
            const int kvec = K & ~1;
            int kk = lane * 2;
            for (; kk + 128 < kvec; kk += 256) {
                d2 xv0 = *reinterpret_cast<const d2*>(x + kk);
                d2 xv1 = *reinterpret_cast<const d2*>(x + kk + 128);
                d2 av0[ROWS], av1[ROWS];
#pragma unroll
                for (int r = 0; r < ROWS; r++) {
                    if (r < nrows) {
                        const double* ar = A + (int64_t)(row0 + r) * lda;
                        av0[r] = __builtin_nontemporal_load(reinterpret_cast<const d2*>(ar + kk));
                        av1[r] = __builtin_nontemporal_load(reinterpret_cast<const d2*>(ar + kk + 128));
                    }
                }
#pragma unroll
                for (int r = 0; r < ROWS; r++) {
                    if (r < nrows) {
                        acc[r] = fma(av0[r].x, xv0.x, acc[r]);
                        acc[r] = fma(av0[r].y, xv0.y, acc[r]);
                        acc[r] = fma(av1[r].x, xv1.x, acc[r]);
                        acc[r] = fma(av1[r].y, xv1.y, acc[r]);
                    }
                }
            }
            for (; kk < kvec; kk += 128) {
                d2 xv = *reinterpret_cast<const d2*>(x + kk);
#pragma unroll
                for (int r = 0; r < ROWS; r++) {
                    if (r < nrows) {
                        d2 av = *reinterpret_cast<const d2*>(A + (int64_t)(row0 + r) * lda + kk);
                        acc[r] = fma(av.x, xv.x, acc[r]);
                        acc[r] = fma(av.y, xv.y, acc[r]);
                    }
                }
            }
            if ((K & 1) && lane == 0) {
                double xv = x[K - 1];
#pragma unroll
                for (int r = 0; r < ROWS; r++)
                    if (r < nrows) acc[r] = fma(A[(int64_t)(row0 + r) * lda + K - 1], xv, acc[r]);
            }
        } else {
            for (int kk = lane; kk < K; kk += 64) {
                double xv = x[(int64_t)kk * incx];
#pragma unroll
                for (int r = 0; r < ROWS; r++)
                    if (r < nrows) acc[r] = fma(A[(int64_t)(row0 + r) * lda + kk], xv, acc[r]);
            }
        }
    }

#pragma unroll
    for (int r = 0; r < ROWS; r++) acc[r] = wave_sum(acc[r]);

    if (lane < nrows) {
        double dot = acc[0];
#pragma unroll
        for (int r = 1; r < ROWS; r++)
            if (lane == r) dot = acc[r];
        double* yp = y + (int64_t)(row0 + lane) * incy;
        double base = (beta == 0.0) ? 0.0 : (beta == 1.0 ? *yp : *yp * beta);
        *yp = (alpha == 0.0 || K == 0) ? base : base + alpha * dot;
    }
}

}  // namespace

hipError_t launch_dgemv(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x,
                        int64_t incx, double beta, double* y, int64_t incy, hipStream_t stream, TileInfo* info) {
    if (m <= 0) return hipSuccess;
    const bool vec = incx == 1 && (reinterpret_cast<uintptr_t>(a) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(x) % 16 == 0) && (lda % 2 == 0);
    const int rows_per_block = ROWS * WAVES;
    const int64_t blocks = (m + rows_per_block - 1) / rows_per_block;
    if (info) {
        info->tile_m = rows_per_block;
        info->tile_n = 1;
        info->tile_k = 256;
        info->launches++;
    }
    mark_main(info, true, stream);
    if (vec)
        hipLaunchKernelGGL(dgemv_rows<true>, dim3((unsigned)blocks), dim3(64 * WAVES), 0, stream, (int)m, (int)k,
                           alpha, a, lda, x, incx, beta, y, incy);
    else
        hipLaunchKernelGGL(dgemv_rows<false>, dim3((unsigned)blocks), dim3(64 * WAVES), 0, stream, (int)m, (int)k,
                           alpha, a, lda, x, incx, beta, y, incy);
    mark_main(info, false, stream);
    return hipGetLastError();
}

}  // namespace jlc
