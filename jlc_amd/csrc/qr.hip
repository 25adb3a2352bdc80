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
// Panel kernels. Rows k0..m-1 of the panel are cut into chunks of PCHUNK rows
// and the panel columns into groups of CW; a launch has one workgroup per
// (row chunk, column group), so every column step runs on thousands of
// workgroups with eight independent columns per thread. Per column k:
//   panel_update: apply reflector k-1 to columns k..k1-1 (w_j reduced from the
//     dot partials, every workgroup in the same order), store v_{k-1}, and
//     produce the per-chunk (max, scaled sum of squares) of column k below the
//     pivot; the owner of row k also saves the pivot x0.
//   panel_reflect: every workgroup combines the norm partials in the same
//     order (so all agree on beta and tau), forms its part of v into a side
//     buffer, and emits per-chunk dot products v . a_j for j = k+1..k1-1; the
//     owner of row k writes R(k,k) = beta and tau[k].
// No workgroup reads data another one writes in the same launch (v travels
// through the side buffer and reaches the matrix one launch later), so no
// grid synchronisation is needed.
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
// Grid of both panel kernels: x = row chunk b (PCHUNK rows from k0), y = group of CW
// consecutive columns starting at the first column the launch works on.
constexpr int CW = 8;

// Reflector scalars of column k from the norm partials, combined in a fixed
// order so every workgroup gets the same bits (computeHouseholder, :219-291).
struct Reflector {
    double tau, beta, v0inv;
};
// The partials are combined by the whole workgroup (each thread takes every
// 256th chunk, then the fixed block reduction tree), so every workgroup of the
// launch computes the same bits.
__device__ Reflector reflector_from(const double* __restrict__ part, int G, int k, int m, double* red) {
    const double x0 = part[2 * G];
    double mloc = 0.0;
    for (int q = threadIdx.x; q < G; q += PTHREADS) mloc = fmax(mloc, part[2 * q]);
    const double M = fmax(fabs(x0), block_max(mloc, red));
    Reflector r{0.0, x0, 0.0};
    if (k < m - 1 && M >= 2.2250738585072014e-308) {
        double sloc = 0.0;
        for (int q = threadIdx.x; q < G; q += PTHREADS) {
            const double s = part[2 * q] / M;
            sloc += part[2 * q + 1] * s * s;
        }
        const double sigma = block_sum(sloc, red);
        const double xnorm = sqrt(sigma);
        const double x0n = x0 / M;
        if (xnorm == 0.0) {
            if (x0n < 0.0) {
                r.tau = 2.0;
                r.beta = -x0;
            }
        } else {
            const double nbeta = -copysign(hypot(x0n, xnorm), x0n);
            r.beta = nbeta * M;
            r.tau = (nbeta - x0n) / nbeta;
            r.v0inv = 1.0 / ((x0n - nbeta) * M);
        }
    }
    return r;
}

// Column k's step, first half. With k > k0: apply reflector k-1 (v from vtmp,
// tau[k-1], w_j = sum over row chunks of dots in a fixed order) to columns
// k + CW*y .. ; the y == 0 groups also store v_{k-1} into the matrix (nobody
// reads column k-1 in this launch) and, for every k, emit the norm partials of
// column k (its first column) after the update.
__global__ __launch_bounds__(PTHREADS) void panel_update(double* __restrict__ at, int64_t lda, int m, int k0, int k1,
                                                         int k, const double* __restrict__ tau,
                                                         const double* __restrict__ vtmp,
                                                         const double* __restrict__ dots, int G,
                                                         double* __restrict__ part) {
    __shared__ double red[4];
    __shared__ double w_s[CW];
    const int b = blockIdx.x;
    const int row = k0 + b * PCHUNK + threadIdx.x;
    const int j0 = k + CW * blockIdx.y;
    const int jn = min(CW, k1 - j0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (k > k0) {
        const int kp = k - 1;
        const double t = tau[kp];
        // w_j: wave wv reduces columns wv and wv + 4 of the group, lanes over chunks, fixed tree
        for (int c = wv; c < CW; c += 4) {
            double s = 0.0;
            if (c < jn)
                for (int q = lane; q < G; q += 64) s += dots[(int64_t)q * 128 + (j0 + c - k0)];
            s = wave_sum(s);
            if (lane == 0) w_s[c] = s;
        }
        __syncthreads();
        if (t != 0.0 && row >= kp && row < m) {
            const double v = row == kp ? 1.0 : vtmp[row];
            const double tv = t * v;
#pragma unroll
            for (int c = 0; c < CW; c++)
                if (c < jn) at[(int64_t)(j0 + c) * lda + row] -= tv * w_s[c];
        }
        if (blockIdx.y == 0 && row > kp && row < m) at[(int64_t)kp * lda + row] = vtmp[row];
    }
    if (blockIdx.y == 0) {
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
}

// Column k's step, second half: every workgroup derives the same reflector from
// the partials; v = x * v0inv below the pivot (1 at the pivot) is computed from
// the unscaled column, which nobody writes in this launch. The y == 0 groups
// store v into vtmp; the owner of row k writes R(k,k) = beta and tau[k]; every
// group emits the per-chunk dot products v . a_j for its columns j > k.
__global__ __launch_bounds__(PTHREADS) void panel_reflect(double* __restrict__ at, int64_t lda, int m, int k0,
                                                          int k1, int k, double* __restrict__ tau, int G,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ vtmp, double* __restrict__ dots) {
    __shared__ double v_s[PCHUNK];
    __shared__ double red[4];
    const int b = blockIdx.x;
    const int row = k0 + b * PCHUNK + threadIdx.x;
    const Reflector rf = reflector_from(part, G, k, m, red);
    double v = 0.0;
    if (row > k && row < m) {
        v = at[(int64_t)k * lda + row];
        if (rf.v0inv != 0.0) v *= rf.v0inv;
        if (blockIdx.y == 0) vtmp[row] = v;
    } else if (row == k) {
        v = 1.0;
    }
    v_s[threadIdx.x] = (row >= k && row < m) ? v : 0.0;
    if (row == k && blockIdx.y == 0) {
        at[(int64_t)k * lda + k] = rf.beta;
        tau[k] = rf.tau;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j0 = k + 1 + CW * blockIdx.y;
    const int rbase = k0 + b * PCHUNK;
    for (int c = wv; c < CW; c += 4) {
        const int j = j0 + c;
        if (j >= k1) break;
        double s = 0.0;
        for (int i = lane; i < PCHUNK; i += 64) {
            const int r = rbase + i;
            if (r >= k && r < m) s += v_s[i] * at[(int64_t)j * lda + r];
        }
        s = wave_sum(s);
        if (lane == 0) dots[(int64_t)b * 128 + (j - k0)] = s;
    }
}

// After the panel's last column: store its v (rows below the pivot) from vtmp.
__global__ __launch_bounds__(256) void panel_store_v(double* __restrict__ at, int64_t lda, int m, int k,
                                                     const double* __restrict__ vtmp) {
    const int row = k + 1 + blockIdx.x * 256 + threadIdx.x;
    if (row < m) at[(int64_t)k * lda + row] = vtmp[row];
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

// T (nb x nb row-major, upper triangular) from G = V^T V and tau, with T(i,i) = tau_i and
// I - V T V^T = H_1 ... H_nb (buildT, HouseholderQR.java:296-320, builds the same T column by column:
// T(0:i, i) = T(0:i, 0:i) * (-tau_i * G(0:i, i))). Here by recursive doubling: for adjacent blocks
// [s, s+b) and [s+b, s+b+b2) whose own T11, T22 are done, T12 = -T11 * G12 * T22 (Schreiber and
// Van Loan's compact-WY merge), all block pairs of one level at once; log2(nb) levels instead of nb
// dependent column steps. The nb x nb LDS array holds T in its upper triangle (diagonal tau) and G in
// its strict lower triangle; each level computes X = G12 T22 into registers, stores it in T12's place,
// then T12 = -T11 X the same way (at most nb^2/4 entries a level: 16 per thread).
__global__ __launch_bounds__(256) void build_t(const double* __restrict__ g, int nb, const double* __restrict__ tau,
                                               double* __restrict__ T) {
    extern __shared__ double S[];  // nb * nb
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int r = idx / nb, c = idx % nb;
        if (r < c) S[c * nb + r] = g[idx];  // G(r, c) in the lower triangle
        else if (r == c) S[idx] = tau[r];
    }
    __syncthreads();
    for (int b = 1; b < nb; b *= 2) {
        const int total = ((nb + 2 * b - 1) / (2 * b)) * b * b;  // block pairs x b x b entries
        double val[16];
        // X = G12 * T22: X(i, j) = sum_{l <= j} G(s+i, s+b+l) T(s+b+l, s+b+j)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int e = threadIdx.x + 256 * q;
            double x = 0.0;
            if (e < total) {
                const int pr = e / (b * b), rem = e % (b * b), i = rem / b, j = rem % b;
                const int s0 = 2 * b * pr, b2 = min(b, nb - s0 - b);
                if (j < b2)
                    for (int l = 0; l <= j; l++) x += S[(s0 + b + l) * nb + s0 + i] * S[(s0 + b + l) * nb + s0 + b + j];
            }
            val[q] = x;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int e = threadIdx.x + 256 * q;
            if (e < total) {
                const int pr = e / (b * b), rem = e % (b * b), i = rem / b, j = rem % b;
                const int s0 = 2 * b * pr, b2 = min(b, nb - s0 - b);
                if (j < b2) S[(s0 + i) * nb + s0 + b + j] = val[q];
            }
        }
        __syncthreads();
        // T12 = -T11 * X: T12(i, j) = -sum_{l >= i} T(s+i, s+l) X(l, j)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int e = threadIdx.x + 256 * q;
            double x = 0.0;
            if (e < total) {
                const int pr = e / (b * b), rem = e % (b * b), i = rem / b, j = rem % b;
                const int s0 = 2 * b * pr, b2 = min(b, nb - s0 - b);
                if (j < b2)
                    for (int l = i; l < b; l++) x += S[(s0 + i) * nb + s0 + l] * S[(s0 + l) * nb + s0 + b + j];
            }
            val[q] = -x;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int e = threadIdx.x + 256 * q;
            if (e < total) {
                const int pr = e / (b * b), rem = e % (b * b), i = rem / b, j = rem % b;
                const int s0 = 2 * b * pr, b2 = min(b, nb - s0 - b);
                if (j < b2) S[(s0 + i) * nb + s0 + b + j] = val[q];
            }
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int r = idx / nb, c = idx % nb;
        T[idx] = r <= c ? S[idx] : 0.0;
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
            // w_j = a(k,j) + sum_{i>k} v_i a(i,j); a(:,j) -= t * v * w_j.
            // Eight threads per column (32 columns at a time) take interleaved rows
            // and three lane swaps add their partial sums.
            const int sub = threadIdx.x & 7, slot = threadIdx.x >> 3;
            for (int jb = k + 1; jb < n; jb += 32) {
                const int j = jb + slot;
                double s = 0.0;
                if (j < n) {
                    const double* cj = a_s + (int64_t)j * m;
                    for (int i = k + 1 + sub; i < m; i += 8) s += col[i] * cj[i];
                }
                s += __shfl_xor(s, 1, 64);
                s += __shfl_xor(s, 2, 64);
                s += __shfl_xor(s, 4, 64);
                if (j < n && sub == 0) wcol[j - k - 1] = s + a_s[(int64_t)j * m + k];
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
                           double* vtmp, hipStream_t stream) {
    const int G = (m - k0 + PCHUNK - 1) / PCHUNK;
    for (int k = k0; k < k1; k++) {
        const int gu = k > k0 ? (k1 - k + CW - 1) / CW : 1;
        hipLaunchKernelGGL(panel_update, dim3(G, gu), dim3(PTHREADS), 0, stream, at, lda, m, k0, k1, k, tau, vtmp,
                           dots, G, part);
        const int gr = k + 1 < k1 ? (k1 - k - 1 + CW - 1) / CW : 1;
        hipLaunchKernelGGL(panel_reflect, dim3(G, gr), dim3(PTHREADS), 0, stream, at, lda, m, k0, k1, k, tau, G,
                           part, vtmp, dots);
    }
    const int last = k1 - 1;
    if (last + 1 < m)
        hipLaunchKernelGGL(panel_store_v, dim3((m - last - 1 + 255) / 256), dim3(256), 0, stream, at, lda, m, last,
                           vtmp);
    return hipGetLastError();
}

hipError_t launch_build_vc(const double* at, int64_t lda, int m, int k0, int nb, double* vc, hipStream_t stream) {
    const int64_t total = (int64_t)nb * (m - k0);
    hipLaunchKernelGGL(build_vc, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, at, lda, m, k0, nb, vc);
    return hipGetLastError();
}

hipError_t launch_build_t(const double* g, int nb, const double* tau, double* T, hipStream_t stream) {
    // Each doubling level holds at most 16 entries per thread: nb <= 128 keeps every level within 4096.
    if (nb < 1 || nb > 128) return hipErrorInvalidValue;
    const size_t lds = (size_t)nb * nb * sizeof(double);
    (void)hipFuncSetAttribute((const void*)build_t, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(build_t, dim3(1), dim3(256), lds, stream, g, nb, tau, T);
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
