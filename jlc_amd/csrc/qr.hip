// Blocked Householder QR on the GPU: FP64 panel kernels, T factor, and the
// batched small-matrix factorisation.
//
// Restates the reference's HouseholderQR (src/main/java/net/faulj/
// decomposition/qr/HouseholderQR.java): column-major storage (the
// reference's AT, :98-127), one reflector per column with the EJML-style
// max-scaled norm and LAPACK sign convention (computeHouseholder, :219-291):
//   beta = -sign(x0) * ||x||, tau = (beta - x0) / beta, v = x / (x0 - beta),
//   v0 = 1 implicit, R(k,k) = beta, with tau = 0 for a zero column and the
//   tau = 2 / negated-pivot case when ||x(1:)|| == 0 and x0 < 0;
// panel apply (factorizePanel / applyHouseholderToColumn, :151-162, :343-385);
// the compact-WY factor T (buildT, :296-320) and the trailing update
// W = V^T C, W = T^T W, C -= V W (updateTrailingMatrix, :167-209) — the
// three GEMMs of that update run on this library's GEMM entry points.
//
// Panel kernels. Rows k0..m-1 of the panel are cut into chunks of PCHUNK rows,
// one workgroup each. Per column k two launches:
//   panel_update_norm: apply reflector k-1 to columns k..k1-1 (w_j reduced
//     from the dot partials, every workgroup in the same order) and produce
//     the per-chunk (max, scaled sum of squares) of column k below the pivot;
//     the owner of row k also saves the pivot x0.
//   panel_reflect_dots: every workgroup combines the norm partials in the same
//     order (so all agree on beta and tau), scales its part of v, and emits
//     per-chunk dot products v . a_j for j = k+1..k1-1; the owner of row k
//     writes R(k,k) = beta and tau[k].
// No workgroup reads data another one writes in the same launch, so no grid
// synchronisation is needed.
#include <math.h>

#include "internal.h"

namespace jlc {
namespace {

constexpr int PCHUNK = 256;   // panel rows per workgroup
constexpr int PTHREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block-wide sum / max over 256 threads (4 wavefronts); result in every thread.
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

__device__ double block_max(double v, double* red) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Norm partial layout: part[b*2 + 0] = chunk max |x|, part[b*2 + 1] = sum (x/max)^2;
// part[2*G] = pivot x0 of the current column.
__global__ __launch_bounds__(PTHREADS) void panel_update_norm(double* __restrict__ at, int64_t lda, int m, int k0,
                                                              int k1, int k, const double* __restrict__ tau,
                                                              const double* __restrict__ dots, int G,
                                                              double* __restrict__ part) {
    __shared__ double red[4];
    __shared__ double w_s[128];
    const int b = blockIdx.x;
    const int row = k0 + b * PCHUNK + threadIdx.x;
    // 1. apply reflector k-1 (if any) to columns k..k1-1 on rows >= k-1
    if (k > k0) {
        const int kp = k - 1;
        const double t = tau[kp];
        if (t != 0.0) {
            for (int j = k + threadIdx.x; j < k1; j += PTHREADS) {
                double s = 0.0;
                for (int q = 0; q < G; q++) s += dots[(int64_t)q * 128 + (j - k0)];
                w_s[j - k0 - (k - k0)] = s;
            }
            __syncthreads();
            if (row >= kp && row < m) {
                const double v = row == kp ? 1.0 : at[(int64_t)kp * lda + row];
                for (int j = k; j < k1; j++) at[(int64_t)j * lda + row] -= t * v * w_s[j - k];
            }
        }
    }
    // 2. norm partial of column k below the pivot
    const bool live = row > k && row < m;
    const double x = live ? at[(int64_t)k * lda + row] : 0.0;
    const double mx = block_max(fabs(x), red);
    double sq = 0.0;
    if (mx > 0.0) {
        const double y = x / mx;
        sq = y * y;
    }
    sq = block_sum(sq, red);
    if (threadIdx.x == 0) {
        part[2 * b] = mx;
        part[2 * b + 1] = sq;
    }
    if (row == k) part[2 * G] = at[(int64_t)k * lda + k];
}

__global__ __launch_bounds__(PTHREADS) void panel_reflect_dots(double* __restrict__ at, int64_t lda, int m, int k0,
                                                               int k1, int k, double* __restrict__ tau, int G,
                                                               const double* __restrict__ part,
                                                               double* __restrict__ dots) {
    __shared__ double red[4];
    __shared__ double v_s[PCHUNK];
    const int b = blockIdx.x;
    const int row = k0 + b * PCHUNK + threadIdx.x;
    // Combine the norm partials in a fixed order: every workgroup gets the same bits.
    const double x0 = part[2 * G];
    double M = fabs(x0);
    for (int q = 0; q < G; q++) M = fmax(M, part[2 * q]);
    double t = 0.0, beta = x0, v0inv = 0.0;
    bool negate = false;
    if (k < m - 1 && M >= 2.2250738585072014e-308) {
        double sigma = 0.0;
        for (int q = 0; q < G; q++) {
            const double r = part[2 * q] / M;
            sigma += part[2 * q + 1] * r * r;
        }
        const double xnorm = sqrt(sigma);
        const double x0n = x0 / M;
        if (xnorm == 0.0) {
            if (x0n < 0.0) {
                t = 2.0;
                beta = -x0;
                negate = true;
            }
        } else {
            const double nbeta = -copysign(hypot(x0n, xnorm), x0n);
            beta = nbeta * M;
            t = (nbeta - x0n) / nbeta;
            v0inv = 1.0 / ((x0n - nbeta) * M);
        }
    }
    (void)negate;
    // scale v on this chunk (rows > k) and stage it
    double v = 0.0;
    if (row > k && row < m) {
        v = at[(int64_t)k * lda + row];
        if (v0inv != 0.0) {
            v *= v0inv;
            at[(int64_t)k * lda + row] = v;
        }
    } else if (row == k) {
        v = 1.0;
    }
    v_s[threadIdx.x] = (row >= k && row < m) ? v : 0.0;
    if (row == k) {
        at[(int64_t)k * lda + k] = beta;
        tau[k] = t;
    }
    __syncthreads();
    // dot partials for columns k+1..k1-1: each wavefront takes every 4th column
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rbase = k0 + b * PCHUNK;
    for (int j = k + 1 + w; j < k1; j += 4) {
        double s = 0.0;
        for (int i = lane; i < PCHUNK; i += 64) {
            const int r = rbase + i;
            if (r >= k && r < m) s += v_s[i] * at[(int64_t)j * lda + r];
        }
        s = wave_sum(s);
        if (lane == 0) dots[(int64_t)b * 128 + (j - k0)] = s;
    }
}

// Apply the last reflector of a panel (column k1-1) has no panel columns left; nothing to do.

// Vc (nb x mk, row-major): Vc[j][i] = v_j at panel row i (unit diagonal, zeros above).
__global__ __launch_bounds__(256) void build_vc(const double* __restrict__ at, int64_t lda, int m, int k0, int nb,
                                                double* __restrict__ vc) {
    const int mk = m - k0;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nb * mk) return;
    const int j = (int)(idx / mk), i = (int)(idx % mk);
    double v;
    if (i < j) v = 0.0;
    else if (i == j) v = 1.0;
    else v = at[(int64_t)(k0 + j) * lda + k0 + i];
    vc[idx] = v;
}

// T (nb x nb row-major, upper triangular) from G = V^T V and tau:
// T(i,i) = tau_i, T(0:i, i) = T(0:i, 0:i) * (-tau_i * G(0:i, i)) (buildT, HouseholderQR.java:296-320).
__global__ __launch_bounds__(256) void build_t(const double* __restrict__ g, int nb, const double* __restrict__ tau,
                                               double* __restrict__ T) {
    extern __shared__ double tmp[];
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) T[idx] = 0.0;
    __syncthreads();
    for (int i = 0; i < nb; i++) {
        const double ti = tau[i];
        for (int r = threadIdx.x; r < i; r += blockDim.x) tmp[r] = -ti * g[r * nb + i];
        __syncthreads();
        for (int r = threadIdx.x; r < i; r += blockDim.x) {
            double s = 0.0;
            for (int c = r; c < i; c++) s += T[r * nb + c] * tmp[c];
            T[r * nb + i] = s;
        }
        if (threadIdx.x == 0) T[i * nb + i] = ti;
        __syncthreads();
    }
}

// Extract the upper triangle R (n x n row-major) from the factored column-major matrix.
__global__ __launch_bounds__(256) void extract_r(const double* __restrict__ at, int64_t lda, int n,
                                                 double* __restrict__ r) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    r[idx] = i <= j ? at[(int64_t)j * lda + i] : 0.0;
}

// Fill a rows x cols row-major block with the identity pattern (1 where i == j + diag).
__global__ __launch_bounds__(256) void fill_eye(double* __restrict__ x, int64_t ld, int rows, int cols, double diag_val) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * cols) return;
    const int i = (int)(idx / cols), j = (int)(idx % cols);
    x[(int64_t)i * ld + j] = i == j ? diag_val : 0.0;
}

// Sum of squares of an m x n row-major block, per-workgroup partials (fixed order reduce on host).
__global__ __launch_bounds__(256) void sumsq_partial(const double* __restrict__ x, int64_t ld, int rows, int cols,
                                                     double* __restrict__ part) {
    __shared__ double red[4];
    double s = 0.0;
    const int64_t total = (int64_t)rows * cols;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const double v = x[(idx / cols) * ld + idx % cols];
        s += v * v;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// Batched small QR: one workgroup factors one column-major m x n matrix held
// in LDS (m*n <= 20480 doubles), unblocked, same reflector convention.
__global__ __launch_bounds__(256) void qr_batched_small(const QrBatchDesc* __restrict__ descs) {
    extern __shared__ double a_s[];
    __shared__ double red[4];
    __shared__ double wcol[256];
    const QrBatchDesc d = descs[blockIdx.x];
    const int m = d.m, n = d.n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int idx = threadIdx.x; idx < m * n; idx += blockDim.x) {
        const int j = idx / m, i = idx % m;
        a_s[idx] = d.a[(int64_t)j * d.lda + i];
    }
    __syncthreads();
    const int kmax = m < n ? m : n;
    for (int k = 0; k < kmax; k++) {
        double* col = a_s + (int64_t)k * m;
        const double x0 = col[k];
        double mx = 0.0;
        for (int i = k + 1 + threadIdx.x; i < m; i += blockDim.x) mx = fmax(mx, fabs(col[i]));
        mx = block_max(mx, red);
        const double M = fmax(mx, fabs(x0));
        double t = 0.0, beta = x0, v0inv = 0.0;
        bool negate = false;
        if (k < m - 1 && M >= 2.2250738585072014e-308) {
            double sq = 0.0;
            for (int i = k + 1 + threadIdx.x; i < m; i += blockDim.x) {
                const double y = col[i] / M;
                sq += y * y;
            }
            sq = block_sum(sq, red);
            const double xnorm = sqrt(sq);
            const double x0n = x0 / M;
            if (xnorm == 0.0) {
                if (x0n < 0.0) {
                    t = 2.0;
                    beta = -x0;
                    negate = true;
                }
            } else {
                const double nbeta = -copysign(hypot(x0n, xnorm), x0n);
                beta = nbeta * M;
                t = (nbeta - x0n) / nbeta;
                v0inv = 1.0 / ((x0n - nbeta) * M);
            }
        }
        (void)negate;
        __syncthreads();
        if (v0inv != 0.0)
            for (int i = k + 1 + threadIdx.x; i < m; i += blockDim.x) col[i] *= v0inv;
        __syncthreads();
        if (threadIdx.x == 0) {
            col[k] = beta;
            d.tau[k] = t;
        }
        if (t != 0.0) {
            // w_j = a(k,j) + sum_{i>k} v_i a(i,j); a(:,j) -= t * v * w_j
            for (int j = k + 1 + w; j < n; j += 4) {
                const double* cj = a_s + (int64_t)j * m;
                double s = 0.0;
                for (int i = k + 1 + lane; i < m; i += 64) s += col[i] * cj[i];
                s = wave_sum(s);
                if (lane == 0) wcol[j - k - 1] = s + cj[k];
            }
            __syncthreads();
            for (int j = k + 1 + w; j < n; j += 4) {
                double* cj = a_s + (int64_t)j * m;
                const double s = t * wcol[j - k - 1];
                if (lane == 0) cj[k] -= s;
                for (int i = k + 1 + lane; i < m; i += 64) cj[i] -= s * col[i];
            }
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < m * n; idx += blockDim.x) {
        const int j = idx / m, i = idx % m;
        d.a[(int64_t)j * d.lda + i] = a_s[idx];
    }
}

}  // namespace

int qr_panel_chunk() { return PCHUNK; }

hipError_t launch_qr_panel(double* at, int64_t lda, int m, int k0, int k1, double* tau, double* part, double* dots,
                           hipStream_t stream) {
    const int G = (m - k0 + PCHUNK - 1) / PCHUNK;
    for (int k = k0; k < k1; k++) {
        hipLaunchKernelGGL(panel_update_norm, dim3(G), dim3(PTHREADS), 0, stream, at, lda, m, k0, k1, k, tau, dots, G,
                           part);
        hipLaunchKernelGGL(panel_reflect_dots, dim3(G), dim3(PTHREADS), 0, stream, at, lda, m, k0, k1, k, tau, G,
                           part, dots);
    }
    return hipGetLastError();
}

hipError_t launch_build_vc(const double* at, int64_t lda, int m, int k0, int nb, double* vc, hipStream_t stream) {
    const int64_t total = (int64_t)nb * (m - k0);
    hipLaunchKernelGGL(build_vc, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, at, lda, m, k0, nb, vc);
    return hipGetLastError();
}

hipError_t launch_build_t(const double* g, int nb, const double* tau, double* T, hipStream_t stream) {
    hipLaunchKernelGGL(build_t, dim3(1), dim3(256), nb * sizeof(double), stream, g, nb, tau, T);
    return hipGetLastError();
}

hipError_t launch_extract_r(const double* at, int64_t lda, int n, double* r, hipStream_t stream) {
    const int64_t total = (int64_t)n * n;
    hipLaunchKernelGGL(extract_r, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, at, lda, n, r);
    return hipGetLastError();
}

hipError_t launch_fill_eye(double* x, int64_t ld, int rows, int cols, double diag, hipStream_t stream) {
    const int64_t total = (int64_t)rows * cols;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_eye, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, x, ld, rows, cols, diag);
    return hipGetLastError();
}

hipError_t launch_sumsq_partial(const double* x, int64_t ld, int rows, int cols, double* part, int nparts,
                                hipStream_t stream) {
    hipLaunchKernelGGL(sumsq_partial, dim3(nparts), dim3(256), 0, stream, x, ld, rows, cols, part);
    return hipGetLastError();
}

hipError_t launch_qr_batched_small(const QrBatchDesc* descs, int count, int max_mn, hipStream_t stream) {
    const size_t lds = (size_t)max_mn * sizeof(double);
    (void)hipFuncSetAttribute((const void*)qr_batched_small, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(qr_batched_small, dim3(count), dim3(256), lds, stream, descs);
    return hipGetLastError();
}

}  // namespace jlc
