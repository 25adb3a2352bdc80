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
// Panel kernel. Rows k0..m-1 of the panel are cut into chunks of PCHUNK rows
// and the panel columns into groups of CW; a launch has one workgroup per
// (row chunk, column group), so every column step runs on thousands of
// workgroups with eight independent columns per thread, and one launch per
// column (panel_col) forms reflector k from partials the previous launch left,
// applies it, and leaves the partials of column k+1: its norm partials, and the
// dot products of its entries with the later columns, from which reflector
// k+1's own dot products follow without another pass. No workgroup reads data
// another one writes in the same launch, so no grid synchronisation is needed.
#include <math.h>

#include "internal.h"

namespace jlc {
namespace {

constexpr int PCHUNK = 512;   // panel rows per workgroup
constexpr int PTHREADS = 512;
constexpr int PWAVES = PTHREADS / 64;

// 1/d from the hardware estimate refined by two Newton steps (within an ulp or two of the IEEE quotient)
// for normal d: the batched kernels' reflector reciprocals sit on each column's critical path.
__device__ __forceinline__ double recip_nr(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

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

// Block-wide sum / max over NW wavefronts; result in every thread (red holds NW doubles).
template <int NW = 4>
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < NW; i++) s += red[i];
    return s;
}

template <int NW = 4>
__device__ double block_max(double v, double* red) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < NW; i++) s = fmax(s, red[i]);
    return s;
}

// Norm partial layout: part[b*2 + 0] = chunk max |x|, part[b*2 + 1] = sum (x/max)^2;
// part[2*G] = pivot x0 of the column.
// Grid of the panel kernel: x = row chunk b (PCHUNK rows from k0), y = group of CW
// consecutive columns starting after the launch's pivot column.
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
    const double M = fmax(fabs(x0), block_max<PWAVES>(mloc, red));
    Reflector r{0.0, x0, 0.0};
    if (k < m - 1 && M >= 2.2250738585072014e-308) {
        double sloc = 0.0;
        for (int q = threadIdx.x; q < G; q += PTHREADS) {
            const double s = part[2 * q] / M;
            sloc += part[2 * q + 1] * s * s;
        }
        const double sigma = block_sum<PWAVES>(sloc, red);
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

// One launch per panel column (panel_col). Launch k forms reflector k and applies it, and prepares
// column k+1 for the next launch; the first launch of a panel (k = k0 - 1) only prepares column k0.
// Column k+1's data travels in one of two buffer sets (P read, Q written; they alternate):
//   part (2G + 1): per row chunk the max |x| and sum (x / max)^2 of column k+1 below its pivot, then
//        the pivot x0;
//   dots (G x 128): per row chunk, sum over the rows below column k+1's pivot of (x_i / chunk max) a_j[i]
//        for the panel columns j > k+1 (after reflector k);
//   prow (128): a_j at column k+1's pivot row; xb (m): column k+1 below its pivot.
// With v = (1 at the pivot, x * v0inv below), v . a_j = prow[j] + sum over chunks of
// (v0inv * chunk max) * dots[chunk][j]: the reflector's dot products come out of the pass that made
// its column, so a column costs one launch. Every workgroup reduces the partials in the same order
// and gets the same bits. Column k+1 itself is not written back (only its row k, R(k, k+1)): every
// workgroup recomputes its rows after reflector k from the unmodified matrix, and the next launch
// reads them from xb; column k is not read from the matrix at all (its x comes from xb), so launch k
// writes v and R(k, k) straight into it.
struct PanelSet {
    double* part;
    double* dots;
    double* prow;
    double* xb;
};

__global__ __launch_bounds__(PTHREADS) void panel_col(double* __restrict__ at, int64_t lda, int m, int k0, int k1,
                                                      int k, double* __restrict__ tau, int G, PanelSet P,
                                                      PanelSet Q) {
    __shared__ double red[PWAVES];
    __shared__ double w_s[CW + 1];       // w_j of the group's columns, then w of column k+1
    __shared__ double xs_s[PCHUNK];      // column k+1 below its pivot, divided by the chunk max
    __shared__ double col_s[CW][PCHUNK]; // the group's columns after reflector k, below column k+1's pivot
    const int b = blockIdx.x, y = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i = k0 + b * PCHUNK + tid;  // this thread's row
    const bool first = k < k0;
    const int j1 = k + 1;                 // next pivot column
    const int jb = j1 + CW * y;           // first column of this group
    // Loads that need nothing from this launch go out first and land during the partial reductions:
    // the group's columns, column k+1 (both unwritten here except by their owners, after the loads)
    // and column k's copy.
    double acol[CW];
#pragma unroll
    for (int c = 0; c < CW; c++) {
        const int j = jb + c;
        acol[c] = (j > j1 && j < k1 && i < m && i >= k) ? at[(int64_t)j * lda + i] : 0.0;
    }
    const double xold = (j1 < k1 && i < m && i > k) ? at[(int64_t)j1 * lda + i] : 0.0;
    const double rk1 = (!first && y == 0 && i == k && j1 < k1) ? at[(int64_t)j1 * lda + k] : 0.0;
    const double xk = (!first && i > k && i < m) ? P.xb[i] : 0.0;
    const Reflector rf = first ? Reflector{0.0, 0.0, 0.0} : reflector_from(P.part, G, k, m, red);
    // w_j = tau * v . a_j for the group's columns (slots 0..CW-1) and column k+1 (slot CW).
    for (int c = wv; c <= CW; c += PWAVES) {
        const int j = c < CW ? jb + c : j1;
        double s = 0.0;
        if (!first && j < k1) {
            for (int q = lane; q < G; q += 64) s += (rf.v0inv * P.part[2 * q]) * P.dots[(int64_t)q * 128 + (j - k0)];
            s = wave_sum(s);
        }
        if (lane == 0) w_s[c] = (first || j >= k1) ? 0.0 : rf.tau * (P.prow[j - k0] + s);
    }
    __syncthreads();
    // v of column k at this row (0 above the pivot; the first launch applies nothing).
    double v = 0.0;
    if (!first && i >= k && i < m) v = i == k ? 1.0 : xk * rf.v0inv;
    if (!first && y == 0) {
        if (i > k && i < m) at[(int64_t)k * lda + i] = rf.v0inv != 0.0 ? xk * rf.v0inv : xk;
        if (i == k) {
            at[(int64_t)k * lda + k] = rf.beta;
            tau[k] = rf.tau;
        }
    }
    // Column k+1 after reflector k: row k becomes R(k, k+1) (group 0 writes it), rows from k+1 on are
    // recomputed by every group.
    double xn = 0.0;
    if (j1 < k1 && i < m) {
        if (i > k) xn = xold - v * w_s[CW];
        else if (!first && i == k && y == 0) at[(int64_t)j1 * lda + k] = rk1 - w_s[CW];
    }
    // The group's own columns (those after column k+1): update in place, keep rows below k+1 in LDS.
#pragma unroll
    for (int c = 0; c < CW; c++) {
        const int j = jb + c;
        double a = 0.0;
        if (j > j1 && j < k1 && i < m && i >= k) {
            a = acol[c];
            if (!first) {
                a -= v * w_s[c];
                at[(int64_t)j * lda + i] = a;
            }
            if (i == j1) Q.prow[j - k0] = a;
        }
        col_s[c][tid] = i > j1 ? a : 0.0;
    }
    if (j1 >= k1) return;  // the panel's last column: nothing to prepare (uniform per launch)
    const double xl = (i > j1 && i < m) ? xn : 0.0;
    const double mb = block_max<PWAVES>(fabs(xl), red);
    const double xsc = mb > 0.0 ? xl / mb : 0.0;
    xs_s[tid] = xsc;
    if (y == 0) {
        const double sq = block_sum<PWAVES>(xsc * xsc, red);
        if (tid == 0) {
            Q.part[2 * b] = mb;
            Q.part[2 * b + 1] = sq;
        }
        if (i == j1) Q.part[2 * G] = xn;
        if (i > j1 && i < m) Q.xb[i] = xn;
    }
    __syncthreads();
    for (int c = wv; c < CW; c += PWAVES) {
        const int j = jb + c;
        if (j <= j1 || j >= k1) continue;
        double s = 0.0;
        for (int t = lane; t < PCHUNK; t += 64) s += xs_s[t] * col_s[c][t];
        s = wave_sum(s);
        if (lane == 0) Q.dots[(int64_t)b * 128 + (j - k0)] = s;
    }
}

// Vc (nb x mk, row-major): Vc[j][i] = v_j at panel row i (unit diagonal, zeros above).
// Row j < nb of vc (row stride ldv): pad zeros, then column k0 + j of V over rows k0..m-1 (zeros above
// the diagonal, the implicit 1 on it, the stored vector below).
__global__ __launch_bounds__(256) void build_vc(const double* __restrict__ at, int64_t lda, int m, int k0, int nb,
                                                double* __restrict__ vc, int64_t ldv, int pad) {
    const int mk = m - k0, w = pad + mk;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nb * w) return;
    const int j = (int)(idx / w), p = (int)(idx % w), i = p - pad;
    double v;
    if (i < j) v = 0.0;
    else if (i == j) v = 1.0;
    else v = at[(int64_t)(k0 + j) * lda + k0 + i];
    vc[(int64_t)j * ldv + p] = v;
}

// T (nb x nb row-major, upper triangular) from G = V^T V and tau, with T(i,i) = tau_i and
// I - V T V^T = H_1 ... H_nb (buildT, HouseholderQR.java:296-320, builds the same T column by column:
// T(0:i, i) = T(0:i, 0:i) * (-tau_i * G(0:i, i))). Here by recursive doubling: for adjacent blocks
// [s, s+b) and [s+b, s+b+b2) whose own T11, T22 are done, T12 = -T11 * G12 * T22 (Schreiber and
// Van Loan's compact-WY merge), all block pairs of one level at once; log2(nb) levels instead of nb
// dependent column steps. The nb x nb LDS array holds T in its upper triangle (diagonal tau) and G in
// its strict lower triangle; each level computes X = G12 T22 into registers, stores it in T12's place,
// then T12 = -T11 X the same way.
// Work split of a level: the P block pairs side by side make W = b * P (pair, column) slots (W <= 64
// for nb <= 128), and thread t takes slot t % W at rows t / W + k * (BT_THREADS / W), k < BT_ROWS, so the
// column operand (T22(l, j), then X(l, j)) is loaded once per l for all of the thread's rows, and the row
// operands are the same for every lane of a wavefront at b >= 8 (LDS broadcast).
constexpr int BT_THREADS = 512, BT_ROWS = 8;  // rows per thread: b <= 64 and BT_THREADS / W >= 8

__global__ __launch_bounds__(BT_THREADS) void build_t(const double* __restrict__ g, int nb,
                                                      const double* __restrict__ tau, double* __restrict__ T) {
    extern __shared__ double S[];  // nb * nb
    const int tid = threadIdx.x;
    // G = V^T V is symmetric: its own lower triangle goes in place. The loads are unconditional (index
    // clamped into the matrix) so that all of them are in flight at once.
    constexpr int FILL = 128 * 128 / BT_THREADS;
    double gv[FILL];
#pragma unroll
    for (int q = 0; q < FILL; q++) gv[q] = g[min(tid + q * BT_THREADS, nb * nb - 1)];
#pragma unroll
    for (int q = 0; q < FILL; q++) {
        const int idx = tid + q * BT_THREADS;
        if (idx < nb * nb) {
            const int r = idx / nb;
            if (r > idx - r * nb) S[idx] = gv[q];
        }
    }
    if (tid < nb) S[tid * nb + tid] = tau[tid];
    __syncthreads();
    for (int b = 1; b < nb; b *= 2) {
        const int P = (nb + 2 * b - 1) / (2 * b), W = b * P, R = BT_THREADS / W;
        const int c = tid % W, i0 = tid / W, j = c % b;
        const int s0 = 2 * b * (c / b), b2 = min(b, nb - s0 - b);
        const bool act = i0 < R && j < b2;
        double x[BT_ROWS];
        // X = G12 * T22: X(i, j) = sum_{l <= j} G(s0+i, s0+b+l) T(s0+b+l, s0+b+j)
#pragma unroll
        for (int k = 0; k < BT_ROWS; k++) x[k] = 0.0;
        if (act) {
            for (int l = 0; l <= j; l++) {
                const int row = (s0 + b + l) * nb + s0;
                const double tv = S[row + b + j];
#pragma unroll
                for (int k = 0; k < BT_ROWS; k++) x[k] += S[row + min(i0 + k * R, b - 1)] * tv;
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int k = 0; k < BT_ROWS; k++)
                if (i0 + k * R < b) S[(s0 + i0 + k * R) * nb + s0 + b + j] = x[k];
        }
        __syncthreads();
        // T12 = -T11 * X: T12(i, j) = -sum_{l >= i} T(s+i, s+l) X(l, j)
#pragma unroll
        for (int k = 0; k < BT_ROWS; k++) x[k] = 0.0;
        if (act) {
            for (int l = i0; l < b; l++) {
                const double xv = S[(s0 + l) * nb + s0 + b + j];
#pragma unroll
                for (int k = 0; k < BT_ROWS; k++) {
                    const int i = i0 + k * R;
                    const double tv = S[(s0 + min(i, b - 1)) * nb + s0 + l];
                    x[k] = fma(l >= i ? tv : 0.0, xv, x[k]);
                }
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int k = 0; k < BT_ROWS; k++)
                if (i0 + k * R < b) S[(s0 + i0 + k * R) * nb + s0 + b + j] = -x[k];
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nb * nb; idx += BT_THREADS) {
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
// Batched small QR: one wavefront factors one column-major m x n matrix held in LDS (n <= 256), the
// reference's unblocked Householder sequence (HouseholderQR.computeHouseholder, HouseholderQR.java:
// 219-291, applied column by column). The leading dimension in LDS is m | 1 when that fits: an odd
// stride puts the 64 lanes' column-parallel reads (lane = column) in distinct banks. A column step is
// two shuffle reductions (max, then the scaled sum of squares), the column scaled by 1 / v0, the dot
// products w_j = v^T a_j with one lane per trailing column (two lanes, halves of the rows, when 32 or
// fewer columns remain) and the rank-1 update with one lane per row; with one wavefront per workgroup
// the barriers between them cost nothing but the ordering of LDS accesses.
__device__ __forceinline__ int qr_wave_ld(int m, int n) {
    return (int64_t)(m | 1) * n <= QR_ODD_LD_MAX ? (m | 1) : m;
}

__global__ __launch_bounds__(64) void qr_batched_wave(const QrBatchDesc* __restrict__ descs) {
    extern __shared__ double a_s[];
    __shared__ double wbuf[256];
    const QrBatchDesc d = descs[blockIdx.x];
    const int m = d.m, n = d.n, ld = qr_wave_ld(m, n);
    const int lane = threadIdx.x;
    for (int idx = lane; idx < m * n; idx += 64) {
        const int j = idx / m, i = idx % m;
        a_s[j * ld + i] = d.a[(int64_t)j * d.lda + i];
    }
    __syncthreads();
    const int kmax = m < n ? m : n;
    for (int k = 0; k < kmax; k++) {
        double* col = a_s + k * ld;
        const double x0 = col[k];
        double mx = 0.0;
        for (int i = k + 1 + lane; i < m; i += 64) mx = fmax(mx, fabs(col[i]));
        const double M = fmax(wave_max(mx), fabs(x0));
        double t = 0.0, beta = x0, v0inv = 0.0;
        if (k < m - 1 && M >= 2.2250738585072014e-308) {
            // x / M as (x s) * (1 / (M s)), s = 2^-100 when M > 2^1000 (else 1), and sqrt(x0n^2 + sum) for
            // the hypotenuse of the scaled values (at least 1): no division on the column's critical path
            const double sc = M > 0x1p1000 ? 0x1p-100 : 1.0;
            const double inv = recip_nr(M * sc);
            double sq = 0.0;
            for (int i = k + 1 + lane; i < m; i += 64) {
                const double y = (col[i] * sc) * inv;
                sq += y * y;
            }
            sq = wave_sum(sq);
            const double x0n = (x0 * sc) * inv;
            if (sq == 0.0) {
                if (x0n < 0.0) {  // H = I - 2 e_k e_k^T negates the column's pivot
                    t = 2.0;
                    beta = -x0;
                }
            } else {
                const double nbeta = -copysign(sqrt(fma(x0n, x0n, sq)), x0n);
                beta = nbeta * M;
                t = (nbeta - x0n) * recip_nr(nbeta);
                v0inv = recip_nr((x0n - nbeta) * M);
            }
        }
        if (v0inv != 0.0)
            for (int i = k + 1 + lane; i < m; i += 64) col[i] *= v0inv;
        __syncthreads();
        if (lane == 0) d.tau[k] = t;
        if (t != 0.0) {
            // w_j = a(k, j) + sum_{i > k} v_i a(i, j) for the trailing columns j > k.
            const int nc = n - k - 1;
            if (nc <= 32) {
                const int j = k + 1 + (lane & 31), half = lane >> 5;
                double s = 0.0;
                if (j < n) {
                    const double* cj = a_s + j * ld;
                    for (int i = k + 1 + half; i < m; i += 2) s += col[i] * cj[i];
                }
                s += __shfl_xor(s, 32, 64);
                if (j < n && half == 0) wbuf[j - k - 1] = s + a_s[j * ld + k];
            } else {
                for (int j = k + 1 + lane; j < n; j += 64) {
                    const double* cj = a_s + j * ld;
                    double s0 = 0.0, s1 = 0.0;
                    int i = k + 1;
                    for (; i + 1 < m; i += 2) {
                        s0 += col[i] * cj[i];
                        s1 += col[i + 1] * cj[i + 1];
                    }
                    if (i < m) s0 += col[i] * cj[i];
                    wbuf[j - k - 1] = (s0 + s1) + cj[k];
                }
            }
            __syncthreads();
            // a(i, j) -= t v_i w_j (v_k = 1), one lane per row.
            for (int i = k + lane; i < m; i += 64) {
                const double tv = t * (i == k ? 1.0 : col[i]);
                for (int j = k + 1; j < n; j++) a_s[j * ld + i] -= tv * wbuf[j - k - 1];
            }
        }
        __syncthreads();
        if (lane == 0) col[k] = beta;
    }
    __syncthreads();
    for (int idx = lane; idx < m * n; idx += 64) {
        const int j = idx / m, i = idx % m;
        d.a[(int64_t)j * d.lda + i] = a_s[j * ld + i];
    }
}

// lo / hi for 0 <= lo <= hi, hi > 0, as (lo s) * (1 / (hi s)): both scaled by a power of two that keeps
// hi s inside [2^-900, 2^900], so its reciprocal is a normal number (no division: these ratios sit on
// the next pivot's critical path).
__device__ __forceinline__ double ratio_le1(double lo, double hi) {
    const double sc = hi < 0x1p-900 ? 0x1p100 : (hi > 0x1p900 ? 0x1p-100 : 1.0);
    return (lo * sc) * recip_nr(hi * sc);
}

// (max |x|, sum (x / max)^2) of two parts of a column, merged (a max of 0 is an empty part).
__device__ __forceinline__ void merge_norm(double& m1, double& s1, double m2, double s2) {
    if (m2 > m1) {
        const double r = ratio_le1(m1, m2);
        s1 = s2 + s1 * r * r;
        m1 = m2;
    } else if (m2 > 0.0) {
        const double r = ratio_le1(m2, m1);
        s1 += s2 * r * r;
    }
}

// One wavefront's (max |x_i|, sum (x_i / max)^2) over rows i in [lo, m) of a column, in every lane
// (the column is not written while it runs).
__device__ __forceinline__ void wave_norm(const double* col, int lo, int m, int lane, double& mx, double& sq) {
    mx = 0.0;
    sq = 0.0;
    for (int i = lo + lane; i < m; i += 64) merge_norm(mx, sq, fabs(col[i]), 1.0);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double m2 = __shfl_xor(mx, off, 64), s2 = __shfl_xor(sq, off, 64);
        merge_norm(mx, sq, m2, s2);
    }
}

// Batched small QR, NT / 64 wavefronts per matrix (the larger size classes): one workgroup factors one
// column-major m x n matrix held in LDS (m*n <= QR_BATCH_MAX_MN doubles, n <= 256, the same odd leading
// dimension as qr_batched_wave when it fits), the same reflector per column as the reference
// (max-scaled norm, beta = -sign(x0) ||x||, tau = (beta - x0) / beta, v = x / (x0 - beta)), with two
// barriers per column: the dot products w_j = a(k,j) + (1 / v0) sum_{i>k} x_i a(i,j) on the unscaled
// column (eight threads per trailing column, NT / 8 columns at a time), then the rank-1 update (one
// wavefront per column). Wavefront 0 updates column k + 1 and leaves its norm statistics (one pass,
// merged by shuffles) for step k + 1, so no step reduces across wavefronts; column k is scaled to v
// during step k + 1, when nothing reads it.
template <int NT>
__global__ __launch_bounds__(NT) void qr_batched_block(const QrBatchDesc* __restrict__ descs) {
    constexpr int NW = NT / 64, SLOTS = NT / 8;
    extern __shared__ double a_s[];
    __shared__ double stat[2];
    __shared__ double wcol[256];
    const QrBatchDesc d = descs[blockIdx.x];
    const int m = d.m, n = d.n, ld = qr_wave_ld(m, n);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int idx = threadIdx.x; idx < m * n; idx += NT) {
        const int j = idx / m, i = idx % m;
        a_s[j * ld + i] = d.a[(int64_t)j * d.lda + i];
    }
    __syncthreads();
    if (w == 0) {
        double mx, sq;
        wave_norm(a_s, 1, m, lane, mx, sq);
        if (lane == 0) {
            stat[0] = mx;
            stat[1] = sq;
        }
    }
    __syncthreads();
    const int kmax = m < n ? m : n;
    double prev_scale = 0.0;  // 1 / v0 of the previous column, applied to it during this step
    for (int k = 0; k < kmax; k++) {
        double* col = a_s + k * ld;
        const double x0 = col[k];
        const double mx = stat[0], sqm = stat[1];
        const double M = fmax(mx, fabs(x0));
        double t = 0.0, beta = x0, v0inv = 0.0;
        if (k < m - 1 && M >= 2.2250738585072014e-308) {
            // scaled as in qr_batched_wave: no division on the column's critical path
            const double sc = M > 0x1p1000 ? 0x1p-100 : 1.0;
            const double inv = recip_nr(M * sc);
            const double r = (mx * sc) * inv;
            const double xnorm = sqrt(sqm * r * r);
            const double x0n = (x0 * sc) * inv;
            if (xnorm == 0.0) {
                if (x0n < 0.0) {  // H = I - 2 e_k e_k^T negates the column's pivot
                    t = 2.0;
                    beta = -x0;
                }
            } else {
                const double nbeta = -copysign(sqrt(fma(x0n, x0n, xnorm * xnorm)), x0n);
                beta = nbeta * M;
                t = (nbeta - x0n) * recip_nr(nbeta);
                v0inv = recip_nr((x0n - nbeta) * M);
            }
        }
        if (k > 0 && prev_scale != 0.0) {  // column k - 1 becomes v (no one reads it in this step)
            double* pc = a_s + (k - 1) * ld;
            for (int i = k + threadIdx.x; i < m; i += NT) pc[i] *= prev_scale;
        }
        if (threadIdx.x == 0) d.tau[k] = t;
        if (t != 0.0) {
            const int sub = threadIdx.x & 7, slot = threadIdx.x >> 3;
            for (int jb = k + 1; jb < n; jb += SLOTS) {
                const int j = jb + slot;
                double s0 = 0.0, s1 = 0.0;
                if (j < n) {
                    const double* cj = a_s + j * ld;
                    int i = k + 1 + sub;
                    for (; i + 8 < m; i += 16) {
                        s0 += col[i] * cj[i];
                        s1 += col[i + 8] * cj[i + 8];
                    }
                    if (i < m) s0 += col[i] * cj[i];
                }
                double sd = s0 + s1;
                sd += __shfl_xor(sd, 1, 64);
                sd += __shfl_xor(sd, 2, 64);
                sd += __shfl_xor(sd, 4, 64);
                if (j < n && sub == 0) wcol[j - k - 1] = t * (a_s[j * ld + k] + v0inv * sd);
            }
        }
        __syncthreads();
        // a(k,j) -= t w_j and a(i,j) -= t w_j x_i / v0 for i > k; wavefront 0 owns column k + 1 and
        // leaves its statistics over rows k + 2 .. m - 1 for the next step.
        for (int j = k + 1 + w; j < n; j += NW) {
            double* cj = a_s + j * ld;
            const double tw = t != 0.0 ? wcol[j - k - 1] : 0.0, twv = tw * v0inv;
            if (t != 0.0 && lane == 0) cj[k] -= tw;
            if (j == k + 1) {
                // Each lane accumulates the rows it updates itself (rows k + 2 .. m - 1), in registers.
                double nmx = 0.0, nsq = 0.0;
                for (int i = k + 1 + lane; i < m; i += 64) {
                    const double v = t != 0.0 ? cj[i] - twv * col[i] : cj[i];
                    if (t != 0.0) cj[i] = v;
                    if (i >= k + 2) merge_norm(nmx, nsq, fabs(v), 1.0);
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const double m2 = __shfl_xor(nmx, off, 64), s2 = __shfl_xor(nsq, off, 64);
                    merge_norm(nmx, nsq, m2, s2);
                }
                if (lane == 0) {
                    stat[0] = nmx;
                    stat[1] = nsq;
                }
            } else if (t != 0.0) {
                for (int i = k + 1 + lane; i < m; i += 64) cj[i] -= twv * col[i];
            }
        }
        if (threadIdx.x == 0) col[k] = beta;
        prev_scale = v0inv;
        __syncthreads();
    }
    if (kmax > 0 && prev_scale != 0.0) {
        double* pc = a_s + (kmax - 1) * ld;
        for (int i = kmax + threadIdx.x; i < m; i += NT) pc[i] *= prev_scale;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < m * n; idx += NT) {
        const int j = idx / m, i = idx % m;
        d.a[(int64_t)j * d.lda + i] = a_s[j * ld + i];
    }
}

// Batched small QR with the matrix in registers (m <= 8 R, n <= 128): eight threads own each column
// (thread t: column t / 8, rows t % 8 + 8 r for r < R), so a column step moves no matrix data through
// LDS. The reflector of column k is formed by its eight owners (max |x| and the sum of (x / max)^2 by
// tree reductions and shuffles inside the group, x0 by a shuffle from its owner), which write the
// whole v (zero above row k, 1 at row k) and tau into a double-buffered LDS vector; after one barrier
// every thread reads v at its rows unconditionally, forms its column's dot product by three shuffles
// and updates its column in registers. The owners of column k + 1 then form reflector k + 1 before the
// next barrier: one barrier per column.
template <int R>
__device__ __forceinline__ void qr_reg_factor(const QrBatchDesc d) {
    // v of the current reflector, thread-major: row sub + 8r at [sub * (R + 2) + r], so each thread's R rows
    // are contiguous (16-byte LDS reads) and the eight threads of a column start on different banks
    constexpr int VS = R + 2;
    __shared__ __attribute__((aligned(16))) double vbuf[2][8 * VS];
    __shared__ double tbuf[2];
    const int m = d.m, n = d.n, kmax = m < n ? m : n;
    const int c = threadIdx.x >> 3, sub = threadIdx.x & 7;
    const bool own = c < n;
    double a[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = sub + 8 * r;
        a[r] = (own && i < m) ? d.a[(int64_t)c * d.lda + i] : 0.0;
    }
    // Reflector of this thread's column at row k, formed by the column's eight owners together.
    auto reflect = [&](int k, int buf) {
        double mxr[4] = {0.0, 0.0, 0.0, 0.0};
        double own_x0 = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = sub + 8 * r;
            if (i == k) own_x0 = a[r];
            mxr[r & 3] = fmax(mxr[r & 3], i > k ? fabs(a[r]) : 0.0);  // rows >= m hold zeros
        }
        double mx = fmax(fmax(mxr[0], mxr[1]), fmax(mxr[2], mxr[3]));
        mx = fmax(mx, __shfl_xor(mx, 1, 64));
        mx = fmax(mx, __shfl_xor(mx, 2, 64));
        mx = fmax(mx, __shfl_xor(mx, 4, 64));
        const double x0 = __shfl(own_x0, k & 7, 8);  // from the owner of row k, bits as they are
        const double M = fmax(mx, fabs(x0));
        double t = 0.0, beta = x0, v0inv = 0.0;
        if (k < m - 1 && M >= 2.2250738585072014e-308) {
            // x / M as (x s) * (1 / (M s)) with s = 1, or s = 2^-100 when M > 2^1000 so that 1 / (M s) stays a
            // normal number (the power-of-two scaling is exact; an x s that underflows has x / M below
            // 2^-1900). No division: these products sit on the per-column critical path.
            const double sc = M > 0x1p1000 ? 0x1p-100 : 1.0;
            const double inv = recip_nr(M * sc);
            double sqr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int i = sub + 8 * r;
                const double y = i > k ? (a[r] * sc) * inv : 0.0;
                sqr[r & 3] += y * y;
            }
            double sq = (sqr[0] + sqr[1]) + (sqr[2] + sqr[3]);
            sq += __shfl_xor(sq, 1, 64);
            sq += __shfl_xor(sq, 2, 64);
            sq += __shfl_xor(sq, 4, 64);
            const double x0n = (x0 * sc) * inv;
            if (sq == 0.0) {
                if (x0n < 0.0) {  // H = I - 2 e_k e_k^T negates the column's pivot
                    t = 2.0;
                    beta = -x0;
                }
            } else {
                // |x0n| <= 1 and sq <= m after the scaling by M, so sqrt(x0n^2 + sq) cannot overflow or
                // underflow (it is at least 1: M is |x0| or the largest |x_i| below the pivot); |nbeta| and
                // |x0n - nbeta| >= 1 keep both reciprocals' arguments normal.
                const double nbeta = -copysign(sqrt(fma(x0n, x0n, sq)), x0n);
                beta = nbeta * M;
                t = (nbeta - x0n) * recip_nr(nbeta);
                v0inv = recip_nr((x0n - nbeta) * M);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = sub + 8 * r;
            if (i == k) {
                a[r] = beta;
            } else if (i > k && v0inv != 0.0) {
                a[r] *= v0inv;  // rows >= m stay zero
            }
            vbuf[buf][sub * VS + r] = i == k ? 1.0 : (i > k && t != 0.0 ? a[r] : 0.0);
        }
        if (sub == 0) {
            tbuf[buf] = t;
            d.tau[k] = t;
        }
    };
    if (c == 0 && kmax > 0) reflect(0, 0);
    __syncthreads();
    for (int k = 0; k < kmax; k++) {
        const int buf = k & 1;
        const double t = tbuf[buf];
        if (t != 0.0 && own && c > k) {
            double dr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < R; r++) dr[r & 3] += vbuf[buf][sub * VS + r] * a[r];
            double dot = (dr[0] + dr[1]) + (dr[2] + dr[3]);
            dot += __shfl_xor(dot, 1, 64);
            dot += __shfl_xor(dot, 2, 64);
            dot += __shfl_xor(dot, 4, 64);
            const double tw = t * dot;
#pragma unroll
            for (int r = 0; r < R; r++) a[r] -= tw * vbuf[buf][sub * VS + r];
        }
        if (c == k + 1 && k + 1 < kmax) reflect(k + 1, buf ^ 1);
        __syncthreads();
    }
    if (own) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = sub + 8 * r;
            if (i < m) d.a[(int64_t)c * d.lda + i] = a[r];
        }
    }
}

// A batch: one workgroup (8 threads per column) per matrix.
template <int R>
__global__ __launch_bounds__(1024) void qr_batched_reg(const QrBatchDesc* __restrict__ descs) {
    qr_reg_factor<R>(descs[blockIdx.x]);
}

// One panel of the blocked QR (m <= 8R rows, n <= 128 columns) on one workgroup: one launch instead of the
// column kernels' n + 1.
template <int R>
__global__ __launch_bounds__(1024) void qr_panel_reg(QrBatchDesc d) {
    qr_reg_factor<R>(d);
}

// Upper trapezoids of stacked row blocks: dst(s, j), s < rows_total, column-major (ldd), is row
// i = s % per of block t = s / per of src (block t starting at row t * stride, column-major, lds),
// kept on and above the diagonal (i <= j), zero below it. With one block this extracts R.
__global__ __launch_bounds__(256) void copy_upper_blocks(const double* __restrict__ src, int64_t lds,
                                                         int64_t stride, int per, int rows_total, int cols,
                                                         double* __restrict__ dst, int64_t ldd) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)rows_total * cols) return;
    const int j = (int)(idx / rows_total), s = (int)(idx % rows_total);
    const int t = s / per, i = s - t * per;
    dst[(int64_t)j * ldd + s] = i <= j ? src[(int64_t)j * lds + (int64_t)t * stride + i] : 0.0;
}

// TSQR's start of Q: dst(i, j) for row i of block t = i / stride (ii = i % stride) is row t * per + ii of
// src when ii < per and that row exists (rows_src of them), else zero.
__global__ __launch_bounds__(256) void spread_blocks(const double* __restrict__ src, int64_t lds, int per,
                                                     int stride, int rows_total, int cols, int rows_src,
                                                     double* __restrict__ dst, int64_t ldd) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)rows_total * cols) return;
    const int j = (int)(idx / rows_total), i = (int)(idx % rows_total);
    const int t = i / stride, ii = i - t * stride, sr = t * per + ii;
    dst[(int64_t)j * ldd + i] = (ii < per && sr < rows_src) ? src[(int64_t)j * lds + sr] : 0.0;
}

// Batched X <- H_0 H_1 ... H_{k-1} X: one workgroup per descriptor holds its column-major m x nx block
// of X in LDS (m <= 256, nx <= 256) and applies the block's reflectors (v_j below the diagonal of the
// factored block, v_j[j] = 1, scalar tau_j) last to first: w = v_j^T X, X -= tau_j v_j w.
__global__ __launch_bounds__(256) void qr_batched_apply_q(const QrApplyDesc* __restrict__ descs) {
    extern __shared__ double x_s[];
    __shared__ double vcol[256];
    __shared__ double wrow[256];
    const QrApplyDesc d = descs[blockIdx.x];
    const int m = d.m, nx = d.nx;
    for (int idx = threadIdx.x; idx < m * nx; idx += blockDim.x) {
        const int c = idx / m, i = idx % m;
        x_s[idx] = d.x[(int64_t)c * d.ldx + i];
    }
    const int sub = threadIdx.x & 7, slot = threadIdx.x >> 3;
    for (int j = d.k - 1; j >= 0; j--) {
        const double t = d.tau[j];
        if (t == 0.0) continue;  // the same value in every thread of the workgroup
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < m; i += blockDim.x) vcol[i] = d.a[(int64_t)j * d.lda + i];
        if (threadIdx.x == 0) vcol[j] = 1.0;
        __syncthreads();
        // w_c = sum_{i >= j} v_i x(i, c): eight threads per column, 32 columns at a time.
        for (int cb = 0; cb < nx; cb += 32) {
            const int c = cb + slot;
            double sacc = 0.0;
            if (c < nx) {
                const double* xc = x_s + (int64_t)c * m;
                for (int i = j + sub; i < m; i += 8) sacc += vcol[i] * xc[i];
            }
            sacc += __shfl_xor(sacc, 1, 64);
            sacc += __shfl_xor(sacc, 2, 64);
            sacc += __shfl_xor(sacc, 4, 64);
            if (c < nx && sub == 0) wrow[c] = sacc;
        }
        __syncthreads();
        const int mj = m - j;
        for (int idx = threadIdx.x; idx < mj * nx; idx += blockDim.x) {
            const int c = idx / mj, i = j + idx % mj;
            x_s[(int64_t)c * m + i] -= t * vcol[i] * wrow[c];
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < m * nx; idx += blockDim.x) {
        const int c = idx / m, i = idx % m;
        d.x[(int64_t)c * d.ldx + i] = x_s[idx];
    }
}

}  // namespace

int qr_panel_chunk() { return PCHUNK; }

hipError_t launch_copy_upper_blocks(const double* src, int64_t lds, int64_t stride, int per, int rows_total, int cols,
                                    double* dst, int64_t ldd, hipStream_t stream) {
    const int64_t total = (int64_t)rows_total * cols;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_upper_blocks, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, src, lds, stride,
                       per, rows_total, cols, dst, ldd);
    return hipGetLastError();
}

hipError_t launch_spread_blocks(const double* src, int64_t lds, int per, int stride, int rows_total, int cols,
                                int rows_src, double* dst, int64_t ldd, hipStream_t stream) {
    const int64_t total = (int64_t)rows_total * cols;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(spread_blocks, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, src, lds, per,
                       stride, rows_total, cols, rows_src, dst, ldd);
    return hipGetLastError();
}

hipError_t launch_qr_batched_apply_q(const QrApplyDesc* descs, int count, int max_mnx, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    const size_t lds = (size_t)max_mnx * sizeof(double);
    (void)hipFuncSetAttribute((const void*)qr_batched_apply_q, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(qr_batched_apply_q, dim3(count), dim3(256), lds, stream, descs);
    return hipGetLastError();
}

hipError_t launch_qr_panel_fused(double* at, int64_t lda, int m, int k0, int k1, double* tau, double* part,
                                 double* dots, double* xbuf, hipStream_t stream) {
    const int G = (m - k0 + PCHUNK - 1) / PCHUNK;
    const int64_t partn = 2 * (int64_t)G + 1;
    PanelSet S[2];
    for (int p = 0; p < 2; p++) {
        S[p].part = part + p * (partn + 128);
        S[p].prow = S[p].part + partn;
        S[p].dots = dots + p * (int64_t)G * 128;
        S[p].xb = xbuf + p * (int64_t)m;
    }
    for (int k = k0 - 1; k < k1; k++) {
        const int groups = k + 1 < k1 ? (k1 - k - 1 + CW - 1) / CW : 1;
        const int rd = (k - k0 + 2) & 1, wr = rd ^ 1;  // launch k reads set rd, writes set wr
        hipLaunchKernelGGL(panel_col, dim3(G, groups), dim3(PTHREADS), 0, stream, at, lda, m, k0, k1, k, tau, G,
                           S[rd], S[wr]);
    }
    return hipGetLastError();
}

hipError_t launch_build_vc(const double* at, int64_t lda, int m, int k0, int nb, double* vc, hipStream_t stream) {
    return launch_build_vc(at, lda, m, k0, nb, vc, m - k0, 0, stream);
}

hipError_t launch_build_vc(const double* at, int64_t lda, int m, int k0, int nb, double* vc, int64_t ldv, int pad,
                           hipStream_t stream) {
    const int64_t total = (int64_t)nb * (pad + m - k0);
    if (ldv < pad + m - k0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(build_vc, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, at, lda, m, k0, nb, vc,
                       ldv, pad);
    return hipGetLastError();
}

hipError_t launch_build_t(const double* g, int nb, const double* tau, double* T, hipStream_t stream) {
    // nb <= 128 keeps every level's (pair, column) slots within 64, so BT_ROWS rows per thread cover it.
    if (nb < 1 || nb > 128) return hipErrorInvalidValue;
    const size_t lds = (size_t)nb * nb * sizeof(double);
    (void)hipFuncSetAttribute((const void*)build_t, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(build_t, dim3(1), dim3(BT_THREADS), lds, stream, g, nb, tau, T);
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

int qr_batched_lds_entries(int m, int n) {
    return ((int64_t)(m | 1) * n <= QR_ODD_LD_MAX ? (m | 1) : m) * n;
}

template <int NT>
static hipError_t launch_block(const QrBatchDesc* descs, int count, size_t lds, hipStream_t stream) {
    (void)hipFuncSetAttribute((const void*)qr_batched_block<NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(qr_batched_block<NT>, dim3(count), dim3(NT), lds, stream, descs);
    return hipGetLastError();
}

hipError_t launch_qr_batched_reg(const QrBatchDesc* descs, int count, int rows_per_thread, int max_n,
                                 hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    if (max_n < 1 || max_n > 128) return hipErrorInvalidValue;
    const int threads = ((max_n * 8 + 63) / 64) * 64;
    switch (rows_per_thread) {
        case 4: hipLaunchKernelGGL(qr_batched_reg<4>, dim3(count), dim3(threads), 0, stream, descs); break;
        case 8: hipLaunchKernelGGL(qr_batched_reg<8>, dim3(count), dim3(threads), 0, stream, descs); break;
        case 12: hipLaunchKernelGGL(qr_batched_reg<12>, dim3(count), dim3(threads), 0, stream, descs); break;
        case 16: hipLaunchKernelGGL(qr_batched_reg<16>, dim3(count), dim3(threads), 0, stream, descs); break;
        case 20: hipLaunchKernelGGL(qr_batched_reg<20>, dim3(count), dim3(threads), 0, stream, descs); break;
        case 24: hipLaunchKernelGGL(qr_batched_reg<24>, dim3(count), dim3(threads), 0, stream, descs); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_qr_panel_reg(double* a, int64_t lda, int m, int n, double* tau, hipStream_t stream) {
    if (m < 1 || n < 1 || n > 128 || m > 192) return hipErrorInvalidValue;
    const int rows = ((m + 31) / 32) * 4;  // rows per thread: m <= 8 * rows
    const int threads = ((n * 8 + 63) / 64) * 64;
    const QrBatchDesc d{m, n, lda, a, tau};
    switch (rows) {
        case 4: hipLaunchKernelGGL(qr_panel_reg<4>, dim3(1), dim3(threads), 0, stream, d); break;
        case 8: hipLaunchKernelGGL(qr_panel_reg<8>, dim3(1), dim3(threads), 0, stream, d); break;
        case 12: hipLaunchKernelGGL(qr_panel_reg<12>, dim3(1), dim3(threads), 0, stream, d); break;
        case 16: hipLaunchKernelGGL(qr_panel_reg<16>, dim3(1), dim3(threads), 0, stream, d); break;
        case 20: hipLaunchKernelGGL(qr_panel_reg<20>, dim3(1), dim3(threads), 0, stream, d); break;
        case 24: hipLaunchKernelGGL(qr_panel_reg<24>, dim3(1), dim3(threads), 0, stream, d); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_qr_batched_small(const QrBatchDesc* descs, int count, int max_lds, int threads,
                                   hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    const size_t lds = (size_t)max_lds * sizeof(double);
    // Only the 256-thread kernel's static LDS (2080 bytes) leaves room past QR_ODD_LD_MAX entries.
    if (max_lds > QR_ODD_LD_MAX && threads != 256) threads = 256;
    switch (threads) {
        case 64:
            (void)hipFuncSetAttribute((const void*)qr_batched_wave, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds);
            hipLaunchKernelGGL(qr_batched_wave, dim3(count), dim3(64), lds, stream, descs);
            return hipGetLastError();
        case 256: return launch_block<256>(descs, count, lds, stream);
        case 512: return launch_block<512>(descs, count, lds, stream);
        case 1024: return launch_block<1024>(descs, count, lds, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace jlc
