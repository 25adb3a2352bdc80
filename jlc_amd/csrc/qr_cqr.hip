// Tall QR panels by CholeskyQR2 with Householder reconstruction.
//
// The column-at-a-time panel (qr.hip) needs one grid-wide reduction per column, so a 128-wide panel
// of a tall matrix is 129 dependent launches, each re-reading the panel from memory. For a panel
// A_p (m_p x b, m_p >= 2b) this path instead forms, with a handful of launches:
//
//   CholeskyQR2:   G1 = A_p^T A_p, R1 = chol(G1), Q1 = A_p R1^-1,
//                  G2 = Q1^T Q1,   R2 = chol(G2), Q  = Q1 R2^-1,  R = R2 R1
//   reconstruction (Ballard, Demmel, Grigori, Jacquelin, Nguyen, Solomonik, "Reconstructing
//   Householder vectors from Tall-Skinny QR", IPDPS 2014): the LU factorisation without pivoting
//   Q_top - D = Y1 U, with the sign D_ii = -sign of the i-th pivot taken as the elimination reaches
//   it, and Y2 = Q_bottom U^-1. Then H = I - Y T Y^T with Y = [Y1; Y2] unit lower trapezoidal is the
//   product of the b reflectors Householder QR of A_p forms, R_H = D R is its R factor and
//   tau_i = -U_ii D_ii (T_ii of T = -U D Y1^-T).
//
// With R's diagonal positive, Householder QR of A_p and of Q = A_p R^-1 make the same reflectors, and
// the sign rule above is the reference's beta = -sign(x0) * ||x|| (HouseholderQR.computeHouseholder,
// src/main/java/net/faulj/decomposition/qr/HouseholderQR.java:219-291), so in exact arithmetic the
// factors are the reference's. Its one discontinuous case, ||x(1:)|| == 0 with x0 >= 0 (tau = 0,
// R(k,k) = x0 instead of tau = 2, R(k,k) = -x0), shows as a column of Y that is exactly zero below the
// diagonal; cqr_special gives such a column tau = 0 and R's row its other sign.
//
// CholeskyQR2 needs kappa(A_p) below ~u^-1/2: the first Cholesky reports a pivot that is not positive
// and finite, the second one an orthogonality defect max|Q1^T Q1 - I| above 2^-8 (which bounds
// ||Q1^T Q1 - I||_2 by 1/2 for b <= 128, where the second pass is exact to O(u)). Either sets a flag
// the host reads before anything is written to the matrix; the panel then runs the column kernels.
//
// Kernels run one workgroup of 1024 threads on a b x b (b <= 128) matrix held in registers: thread t
// owns column k = t % 128 and rows 16g + q (g = t / 128, q < 16), so a thread's rows are the same
// across a wavefront and contiguous (LDS reads of row-indexed values are 16-byte broadcasts). Each
// elimination or substitution step broadcasts one row and/or column through a double-buffered LDS
// vector: one barrier per step.
#include <math.h>

#include "internal.h"

namespace jlc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int CQ_NT = 512;           // 8 wavefronts
constexpr int CQ_W = CQ_NT / 64;
constexpr int SLD = 129;             // LDS matrix: 128 x SLD doubles, row-major (odd: conflict-free columns)
constexpr int CQ_DYN = (128 * SLD + 8 * 256) * 8;  // dynamic LDS: the matrix and 8 diagonal-block inverses
constexpr int GR_MAXT = 5;   // upper 16 x 16 tiles per wavefront: 36 tiles of a 128 x 128 matrix over 8

// 1/d and 1/sqrt(d) from the hardware estimates refined by Newton steps (within an ulp or two; the
// IEEE division and square root sequences would sit on every elimination step's critical path).
__device__ __forceinline__ double rcp_nr(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

__device__ __forceinline__ double rsq_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int i = 0; i < 2; i++) y = fma(0.5 * y, fma(-d * y, y, 1.0), y);
    return y;
}

__device__ __forceinline__ d4 mfma(d4 acc, double a, double b) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}

// v_mfma_f64_16x16x4 fragments: lane l (i16 = l % 16, g = l / 16) supplies A(i16, k) and B(k, i16) with
// k = g + 4 * step for the four steps of a 16-deep product; accumulator register r of lane l holds
// C(g + 4r, i16).
// acc += sa * A * B, A(i, k) at a[i * lda + k], B(k, j) at b[k * ldb + j]; b_upper reads B's entries
// below its diagonal as zero.
__device__ __forceinline__ d4 tile_mma(d4 acc, const double* a, int lda, const double* b, int ldb, double sa,
                                       bool b_upper = false) {
    const int l = threadIdx.x & 63, i16 = l & 15, g = l >> 4;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const double bv = b[(g + 4 * s) * ldb + i16];
        acc = mfma(acc, sa * a[i16 * lda + g + 4 * s], (b_upper && g + 4 * s > i16) ? 0.0 : bv);
    }
    return acc;
}

// acc += sa * A^T * B, A(k, i) at a[k * lda + i].
__device__ __forceinline__ d4 tile_mma_tn(d4 acc, const double* a, int lda, const double* b, int ldb, double sa) {
    const int l = threadIdx.x & 63, i16 = l & 15, g = l >> 4;
#pragma unroll
    for (int s = 0; s < 4; s++) acc = mfma(acc, sa * a[(g + 4 * s) * lda + i16], b[(g + 4 * s) * ldb + i16]);
    return acc;
}

__device__ __forceinline__ d4 tile_load(const double* c, int ldc) {
    const int l = threadIdx.x & 63, i16 = l & 15, g = l >> 4;
    d4 v;
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = c[(g + 4 * r) * ldc + i16];
    return v;
}

// upper_only: entries below the tile's diagonal are left as they are.
__device__ __forceinline__ void tile_store(double* c, int ldc, d4 v, bool upper_only) {
    const int l = threadIdx.x & 63, i16 = l & 15, g = l >> 4;
#pragma unroll
    for (int r = 0; r < 4; r++)
        if (!upper_only || g + 4 * r <= i16) c[(g + 4 * r) * ldc + i16] = v[r];
}

// In place, X = U^-1 for the upper-triangular nblk*16 square U in S (entries below the diagonal are not
// read, and only the diagonal tiles' upper triangles are written, so a strictly lower part survives).
// dinv[i] = 1 / U(i, i). First the nblk diagonal 16 x 16 blocks are inverted side by side (16 lanes per
// block, each lane one column by back substitution) into dblk; then block rows from the bottom up:
// X(ib, cb) = -U(ib, ib)^-1 sum_{ib < lb <= cb} U(ib, lb) X(lb, cb) on FP64 MFMA, one tile per wavefront.
__device__ void upper_inverse_blocked(double* S, double* dblk, const double* dinv, int nblk) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    if (w < 2) {
        const int blk = 4 * w + (l >> 4), c = l & 15;
        if (blk < nblk) {
            const double* U = S + 16 * blk * SLD + 16 * blk;
            double x[16];
#pragma unroll
            for (int r = 15; r >= 0; r--) {
                double v = r == c ? 1.0 : 0.0;
#pragma unroll
                for (int q = r + 1; q < 16; q++) v = fma(-U[r * SLD + q], x[q], v);
                x[r] = v * dinv[16 * blk + r];
            }
#pragma unroll
            for (int r = 0; r < 16; r++) dblk[blk * 256 + r * 16 + c] = x[r];
        }
    }
    __syncthreads();
    if (w == 0) tile_store(S + 16 * (nblk - 1) * (SLD + 1), SLD, tile_load(dblk + (nblk - 1) * 256, 16), true);
    for (int ib = nblk - 2; ib >= 0; ib--) {
        const int cb = ib + 1 + w;
        d4 res = {0.0, 0.0, 0.0, 0.0};
        if (cb < nblk) {
            // two independent accumulator chains (alternate lb) so consecutive MFMAs do not wait on each other
            d4 a0 = {0.0, 0.0, 0.0, 0.0}, a1 = {0.0, 0.0, 0.0, 0.0};
            for (int lb = ib + 1; lb <= cb; lb += 2) {  // X(cb, cb)'s strictly lower part may hold other data
                a0 = tile_mma(a0, S + 16 * ib * SLD + 16 * lb, SLD, S + 16 * lb * SLD + 16 * cb, SLD, 1.0, lb == cb);
                if (lb + 1 <= cb)
                    a1 = tile_mma(a1, S + 16 * ib * SLD + 16 * (lb + 1), SLD, S + 16 * (lb + 1) * SLD + 16 * cb, SLD,
                                  1.0, lb + 1 == cb);
            }
            const d4 acc = a0 + a1;
            // res = -Dinv(ib) * acc: acc's register s is the B fragment of product step s.
            const int i16 = l & 15, g = l >> 4;
            const double* di = dblk + ib * 256;
#pragma unroll
            for (int s = 0; s < 4; s++) res = mfma(res, -di[i16 * 16 + g + 4 * s], acc[s]);
        }
        __syncthreads();  // row block ib of U fully read
        if (cb < nblk) tile_store(S + 16 * ib * SLD + 16 * cb, SLD, res, false);
        if (w == CQ_W - 1) tile_store(S + 16 * ib * (SLD + 1), SLD, tile_load(dblk + ib * 256, 16), true);
        __syncthreads();
    }
    // With one block (b <= 16) the loop above does not run: the last diagonal inverse, stored by wavefront
    // 0 alone, must be complete before any wavefront reads S.
    if (nblk == 1) __syncthreads();
}

// acc = sum_{lb = ib..cb} A(ib, lb) * R1(lb, cb) with A from S and R1 row-major b x b in global memory
// (zero past b): the R1 fragments of the whole tile row are loaded before the MFMA chain.
__device__ __forceinline__ d4 tile_times_r1(const double* S, int ib, int cb, const double* __restrict__ r1, int b,
                                            d4* r1_tile = nullptr) {
    const int l = threadIdx.x & 63, i16 = l & 15, g = l >> 4;
    double bv[8][4];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int lb = ib + q;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int kr = 16 * lb + g + 4 * s, col = 16 * cb + i16;
            bv[q][s] = (lb <= cb && kr < b && col < b) ? r1[kr * b + col] : 0.0;
        }
    }
    // R1(ib, cb) itself: fragment step s of lb = ib holds row 16 ib + g + 4s, the accumulator layout
    if (r1_tile) *r1_tile = d4{bv[0][0], bv[0][1], bv[0][2], bv[0][3]};
    d4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int lb = ib + q;
        if (lb <= cb) {
#pragma unroll
            for (int s = 0; s < 4; s++)
                acc[q & 1] = mfma(acc[q & 1], S[(16 * ib + i16) * SLD + 16 * lb + g + 4 * s], bv[q][s]);
        }
    }
    return acc[0] + acc[1];
}

// Upper tile (ib, cb), ib <= cb, of the idx-th entry of the row-by-row upper tile list of nblk blocks.
__device__ __forceinline__ void upper_tile(int idx, int nblk, int& ib, int& cb) {
    int a = 0;
    while (idx >= nblk - a) {
        idx -= nblk - a;
        a++;
    }
    ib = a;
    cb = a + idx;
}

// pass 1: R1 = chol(G) (upper, R1^T R1 = G), X1 = R1^-1, both b x b row-major; flag = 1 on a pivot that
// is not positive and finite (a plain store: pass 1 starts the panel's flag).
// pass 2: flag |= 2 when max|G - I| > 2^-8; R2 = chol(G), r_out = R2 * R1 (R1 row-major from r1) stored
// column-major, X2 = R2^-1 row-major. When G = I + E is close to the identity (max|E| <= 2^-26, the
// usual case: E is the first pass's loss of orthogonality, about kappa^2 u), R2 = I + P is formed
// without elimination: P = U(E - P^T P) (U keeps the strict upper triangle and half the diagonal) from
// P = U(E) (one step, dropped when max|E| <= 2^-40, where P^T P is below 2^-73), and
// X2 = I - P + P^2 (P^2 dropped likewise); the truncation is below 2^-60 in both tiers.
// G is b x b row-major (symmetric; only its upper triangle is read); padded with the identity to a
// multiple of 16. Elimination path: right-looking blocked Cholesky; per 16-row block, 16 steps (one
// barrier each, the block row's updates on the first four wavefronts) factor the block row, then the
// trailing upper tiles take the block's rank-16 update on FP64 MFMA; X by upper_inverse_blocked.
__global__ __launch_bounds__(CQ_NT) void cqr_chol(const double* __restrict__ gm, int b, int pass,
                                                  const double* __restrict__ r1, double* __restrict__ r_out,
                                                  double* __restrict__ x_out, int* __restrict__ flag) {
    __shared__ double rowb[2][128];
    __shared__ double dinv[128];
    __shared__ double wmax[CQ_W];
    __shared__ int bad;
    extern __shared__ double dyn[];
    double* S = dyn;
    double* dblk = dyn + 128 * SLD;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, i16 = l & 15, g = l >> 4;
    const int nb16 = (b + 15) & ~15, nblk = nb16 >> 4;
    const int c = t & 127;
    if (t == 0) bad = 0;
    double dev = 0.0;
    {
        // all 32 loads of a thread in flight at once, then the LDS stores
        double v[32];
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int r = (t >> 7) + 4 * q;
            v[q] = (r < b && c < b && r <= c) ? gm[r * b + c] : (r == c ? 1.0 : 0.0);
        }
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int r = (t >> 7) + 4 * q;
            if (r < b && c < b && r <= c) dev = fmax(dev, fabs(v[q] - (r == c ? 1.0 : 0.0)));
            if (r < nb16 && c < nb16) S[r * SLD + c] = v[q];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) dev = fmax(dev, __shfl_xor(dev, off, 64));  // NaN-free: fmax
    if (l == 0) wmax[w] = dev;
    __syncthreads();
    double devmax = 0.0;
#pragma unroll
    for (int q = 0; q < CQ_W; q++) devmax = fmax(devmax, wmax[q]);
    // a NaN entry leaves dev unchanged through fmax: count it as a failure
    bool nan_in = false;
    for (int r = t >> 7; r < b; r += CQ_NT / 128)
        if (c < b && r <= c && !(S[r * SLD + c] == S[r * SLD + c])) nan_in = true;
    if (pass == 2 && (!(devmax <= 0x1p-8) || nan_in)) bad = 1;
    if (pass == 2 && devmax <= 0x1p-26) {
        const bool second = devmax > 0x1p-40;
        // S <- P = U(E) on the upper triangle (the strictly lower part is zero).
        for (int r = t >> 7; r < nb16; r += CQ_NT / 128)
            if (c < nb16 && r <= c) S[r * SLD + c] = r == c ? 0.5 * (S[r * SLD + c] - 1.0) : S[r * SLD + c];
        __syncthreads();
        const int ntile = nblk * (nblk + 1) / 2;
        if (second) {
            // P <- U(E - P^T P) with E = P + P^T: off the diagonal P - (P^T P), on it P - (P^T P) / 2.
            d4 acc[GR_MAXT];
#pragma unroll
            for (int q = 0; q < GR_MAXT; q++) {
                acc[q] = d4{0.0, 0.0, 0.0, 0.0};
                const int idx = w + CQ_W * q;
                if (idx < ntile) {
                    int ib, cb;
                    upper_tile(idx, nblk, ib, cb);
                    for (int lb = 0; lb <= ib; lb++)
                        acc[q] = tile_mma_tn(acc[q], S + 16 * lb * SLD + 16 * ib, SLD, S + 16 * lb * SLD + 16 * cb,
                                             SLD, 1.0);
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < GR_MAXT; q++) {
                const int idx = w + CQ_W * q;
                if (idx < ntile) {
                    int ib, cb;
                    upper_tile(idx, nblk, ib, cb);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = 16 * ib + g + 4 * r, col = 16 * cb + i16;
                        if (row < col) S[row * SLD + col] -= acc[q][r];
                        else if (row == col) S[row * SLD + col] -= 0.5 * acc[q][r];
                    }
                }
            }
            __syncthreads();
        }
        // X2 = I - P (+ P^2), R = R1 + P R1, upper tiles; lower parts zero.
        for (int idx = w; idx < ntile; idx += CQ_W) {
            int ib, cb;
            upper_tile(idx, nblk, ib, cb);
            d4 pp = {0.0, 0.0, 0.0, 0.0}, pq = {0.0, 0.0, 0.0, 0.0};
            if (second) {
                for (int lb = ib; lb <= cb; lb += 2) {
                    pp = tile_mma(pp, S + 16 * ib * SLD + 16 * lb, SLD, S + 16 * lb * SLD + 16 * cb, SLD, 1.0);
                    if (lb + 1 <= cb)
                        pq = tile_mma(pq, S + 16 * ib * SLD + 16 * (lb + 1), SLD, S + 16 * (lb + 1) * SLD + 16 * cb,
                                      SLD, 1.0);
                }
                pp += pq;
            }
            d4 r1t;
            const d4 pr = tile_times_r1(S, ib, cb, r1, b, &r1t);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * ib + g + 4 * r, col = 16 * cb + i16;
                if (row < b && col < b) {
                    const double pv = row <= col ? S[row * SLD + col] : 0.0;
                    x_out[row * b + col] = (row == col ? 1.0 : 0.0) - pv + pp[r];
                    r_out[col * b + row] = r1t[r] + pr[r];
                }
            }
        }
        for (int r = t >> 7; r < b; r += CQ_NT / 128)
            if (c < r) {
                r_out[c * b + r] = 0.0;
                x_out[r * b + c] = 0.0;
            }
        __syncthreads();
        if (t == 0 && bad) *flag |= 2;
        return;
    }
    // Elimination path. Panel role (first four wavefronts): column c, rows j0 + h + 2i (i < 8) of the block.
    const int h = (t >> 7) & 1;
    const bool panel_wave = t < 256;
    for (int jb = 0; jb < nblk; jb++) {
        const int j0 = 16 * jb;
        // a wavefront whose columns all lie left of the block has nothing to update (those entries are 0)
        const bool work = panel_wave && 64 * (w & 1) + 63 >= j0 && 64 * (w & 1) < nb16;
        double pv[8], cm[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = j0 + h + 2 * i;
            pv[i] = (work && c < nb16) ? S[r * SLD + c] : 0.0;
            cm[i] = c >= r ? 1.0 : 0.0;
        }
#pragma unroll
        for (int jj = 0; jj < 16; jj++) {
            const int j = j0 + jj, p = jj & 1;
            const bool own = work && h == (jj & 1);
            if (own) rowb[p][c] = pv[jj >> 1];
            __syncthreads();
            if (work) {
                // every LDS read of the step issued before the first use; updates are plain FMAs
                const double d = rowb[p][j];
                const double rk = rowb[p][c];
                double rr[8];
#pragma unroll
                for (int i = 0; i < 8; i++) rr[i] = rowb[p][j0 + h + 2 * i];
                if (t == 0 && !(d > 0.0 && d <= 1.7976931348623157e308)) bad = 1;
                const double f = rk * rcp_nr(d);
#pragma unroll
                for (int i = 0; i < 8; i++)
                    if (2 * i + h > jj) pv[i] = fma(-rr[i], f * cm[i], pv[i]);  // condition uniform per wavefront
                if (own) {
                    const double rs = rsq_nr(d);
                    const double rv = c >= j ? rk * rs : 0.0;
                    pv[jj >> 1] = rv;
                    if (c == j) dinv[j] = rcp_nr(rv);
                }
            }
        }
        if (work && c < nb16) {
#pragma unroll
            for (int i = 0; i < 8; i++) S[(j0 + h + 2 * i) * SLD + c] = pv[i];
        }
        __syncthreads();
        // Trailing upper tiles (kb, lb), jb < kb <= lb: S(kb, lb) -= R(jb, kb)^T R(jb, lb).
        const int nt = nblk - jb - 1;
        for (int idx = w; idx < nt * (nt + 1) / 2; idx += CQ_W) {
            int kb, lb;
            upper_tile(idx, nt, kb, lb);
            kb += jb + 1;
            lb += jb + 1;
            double* ct = S + 16 * kb * SLD + 16 * lb;
            d4 acc = tile_load(ct, SLD);
            acc = tile_mma_tn(acc, S + j0 * SLD + 16 * kb, SLD, S + j0 * SLD + 16 * lb, SLD, -1.0);
            tile_store(ct, SLD, acc, false);
        }
        __syncthreads();
    }
    if (pass == 1) {
        for (int r = t >> 7; r < b; r += CQ_NT / 128)
            if (c < b) r_out[r * b + c] = S[r * SLD + c];
    } else {
        // R = R2 R1 tile by tile (upper tiles), stored column-major; zeros below the diagonal.
        for (int idx = w; idx < nblk * (nblk + 1) / 2; idx += CQ_W) {
            int ib, cb;
            upper_tile(idx, nblk, ib, cb);
            const d4 acc = tile_times_r1(S, ib, cb, r1, b);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * ib + g + 4 * r, col = 16 * cb + i16;
                if (row < b && col < b) r_out[col * b + row] = acc[r];
            }
        }
        for (int r = t >> 7; r < b; r += CQ_NT / 128)
            if (c < r) r_out[c * b + r] = 0.0;
    }
    __syncthreads();
    upper_inverse_blocked(S, dblk, dinv, nblk);
    for (int r = t >> 7; r < b; r += CQ_NT / 128)
        if (c < b) x_out[r * b + c] = r <= c ? S[r * SLD + c] : 0.0;
    if (t == 0) {
        if (pass == 1) *flag = bad ? 1 : 0;
        else if (bad) *flag |= 2;
    }
}

// Modified LU of the top b x b block of Q (column-major, ld ldq): Q_top - D = Y1 U, D_ii = -sign(c_i)
// of the pivot c_i the elimination reaches (copysign, so -0.0 counts as negative as in the
// reference's copySign), U_ii = c_i - D_ii (|U_ii| >= 1, no breakdown). Writes into the panel's top
// block (column-major, lda): Y1 below the diagonal, R_H = D R (R column-major b x b) on and above it;
// tau_i = -U_ii D_ii; ui = U^-1 (row-major); spec_i = 1 where Y1's column i is exactly zero below the
// diagonal and c_i >= 0 (a candidate for the reference's tau = 0 case; cqr_special checks Y2).
// Right-looking blocked: for each 16-column block, 16 steps (one barrier each) factor the block column,
// each column of the block row is then solved by one thread, and the trailing tiles take the rank-16
// update on FP64 MFMA; U^-1 as in cqr_chol. Padded with the identity to a multiple of 16.
__global__ __launch_bounds__(CQ_NT) void cqr_lu(const double* __restrict__ qm, int64_t ldq, int b,
                                                const double* __restrict__ rr, double* __restrict__ ap,
                                                int64_t lda, double* __restrict__ tau, double* __restrict__ ui,
                                                int* __restrict__ spec) {
    __shared__ double rowb[2][16];
    __shared__ double colb[2][128];
    __shared__ double dinv[128];
    __shared__ double dsg[128];
    __shared__ double cpv[128];
    __shared__ double upv[128];
    __shared__ int nz[128];
    extern __shared__ double dyn[];
    double* S = dyn;
    double* dblk = dyn + 128 * SLD;
    const int t = threadIdx.x, w = t >> 6;
    const int nb16 = (b + 15) & ~15, nblk = nb16 >> 4;
    // Coalesced load of Q's top block (column contiguous in memory), identity past b.
    {
        const int r = t & 127;
        double v[32];
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int c = (t >> 7) + 4 * q;
            v[q] = (r < b && c < b) ? qm[(int64_t)c * ldq + r] : (r == c ? 1.0 : 0.0);
        }
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int c = (t >> 7) + 4 * q;
            if (r < nb16 && c < nb16) S[r * SLD + c] = v[q];
        }
    }
    if (t < 128) nz[t] = 0;
    __syncthreads();
    // Block-column role (first four wavefronts): column j0 + cc, rows j0 + r0 + 16e, e < 8.
    const int cc = t & 15, r0 = (t >> 4) & 15;
    const bool panel_wave = t < 256;
    for (int jb = 0; jb < nblk; jb++) {
        const int j0 = 16 * jb, ne = nblk - jb;  // row groups e < ne lie inside the matrix
        double pv[8];
#pragma unroll
        for (int e = 0; e < 8; e++) pv[e] = (panel_wave && e < ne) ? S[(j0 + r0 + 16 * e) * SLD + j0 + cc] : 0.0;
#pragma unroll
        for (int ii = 0; ii < 16; ii++) {
            const int i = j0 + ii, p = ii & 1;
            if (panel_wave) {
                if (r0 == ii) rowb[p][cc] = pv[0];
                if (cc == ii) {
#pragma unroll
                    for (int e = 0; e < 8; e++) colb[p][r0 + 16 * e] = pv[e];
                }
            }
            __syncthreads();
            if (panel_wave) {
                const double cv = rowb[p][ii];
                const double rk = rowb[p][cc];
                double col[8];
#pragma unroll
                for (int e = 0; e < 8; e++) col[e] = colb[p][r0 + 16 * e];
                const double d = -copysign(1.0, cv);
                const double u = cv - d;
                const double iu = rcp_nr(u);
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    if (e < ne) {  // uniform
                        const double lv = col[e] * iu;
                        const double upd = fma(-lv, rk, pv[e]);
                        const bool below = e > 0 || r0 > ii;
                        pv[e] = below ? (cc > ii ? upd : (cc == ii ? lv : pv[e])) : pv[e];
                    }
                }
                if (r0 == ii && cc == ii) {
                    pv[0] = u;
                    dinv[i] = iu;
                    dsg[i] = d;
                    cpv[i] = cv;
                    upv[i] = u;
                }
            }
        }
        if (panel_wave) {
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (e < ne) S[(j0 + r0 + 16 * e) * SLD + j0 + cc] = pv[e];
        }
        __syncthreads();
        // Block row: U(i, c) = A(i, c) - sum_{j0 <= q < i} L(i, q) U(q, c), one thread per column c.
        {
            const int c = j0 + 16 + t;
            if (c < nb16) {
                double uv[16];
#pragma unroll
                for (int i = 0; i < 16; i++) uv[i] = S[(j0 + i) * SLD + c];
#pragma unroll
                for (int i = 1; i < 16; i++)
#pragma unroll
                    for (int q = 0; q < i; q++) uv[i] = fma(-S[(j0 + i) * SLD + j0 + q], uv[q], uv[i]);
#pragma unroll
                for (int i = 1; i < 16; i++) S[(j0 + i) * SLD + c] = uv[i];
            }
        }
        __syncthreads();
        // Trailing tiles (rb, cb), rb, cb > jb: S(rb, cb) -= L(rb, jb) U(jb, cb).
        const int nt = nblk - jb - 1;
        for (int idx = w; idx < nt * nt; idx += CQ_W) {
            const int rb = jb + 1 + idx / nt, cb = jb + 1 + idx % nt;
            double* ct = S + 16 * rb * SLD + 16 * cb;
            d4 acc = tile_load(ct, SLD);
            acc = tile_mma(acc, S + 16 * rb * SLD + j0, SLD, S + j0 * SLD + 16 * cb, SLD, -1.0);
            tile_store(ct, SLD, acc, false);
        }
        __syncthreads();
    }
    {
        const int c = t & 127;
        for (int r = t >> 7; r < b; r += 4)
            if (c < b && r > c && S[r * SLD + c] != 0.0) nz[c] = 1;
    }
    __syncthreads();
    if (t < b) {
        tau[t] = -upv[t] * dsg[t];
        spec[t] = (nz[t] == 0 && cpv[t] >= 0.0) ? 1 : 0;
    }
    upper_inverse_blocked(S, dblk, dinv, nblk);
    {
        const int c = t & 127;
        for (int r = t >> 7; r < b; r += 4)
            if (c < b) ui[r * b + c] = r <= c ? S[r * SLD + c] : 0.0;
    }
    __syncthreads();
    // Top block of the factored panel: D R on and above the diagonal (R column-major), Y1 below.
    {
        const int r = t & 127;
        double v[32];
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int c = (t >> 7) + 4 * q;
            v[q] = (r < b && c < b && r <= c) ? rr[c * b + r] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int c = (t >> 7) + 4 * q;
            if (r < b && c < b) ap[(int64_t)c * lda + r] = r > c ? S[r * SLD + c] : dsg[r] * v[q];
        }
    }
}

// Gram matrix G = X^T X of a column-major m x b matrix X (ld ldx, b <= 128), in two launches: workgroup p
// of gram_partial stages rows [p * per, (p + 1) * per) in 64-row chunks through LDS (column c of a chunk
// contiguous, as in memory) and accumulates the upper 16 x 16 tiles of X_p^T X_p on FP64 MFMA, one
// tile per wavefront at a time; gram_reduce sums the partial tiles in a fixed order and writes both
// triangles of G (b x b, row-major).
constexpr int GR_ROWS = 64;
constexpr int GR_LDC = GR_ROWS + 1;

__global__ __launch_bounds__(CQ_NT) void gram_partial(const double* __restrict__ x, int64_t ldx, int m, int b,
                                                      int per, double* __restrict__ part) {
    __shared__ double xs[128 * GR_LDC];
    const int t = threadIdx.x, w = t >> 6, l = t & 63, i16 = l & 15, g = l >> 4;
    const int nblk = (b + 15) >> 4, ntile = nblk * (nblk + 1) / 2;
    const int r0 = blockIdx.x * per, r1 = min(m, r0 + per);
    int ti[GR_MAXT], tj[GR_MAXT];
#pragma unroll
    for (int q = 0; q < GR_MAXT; q++) {
        int idx = w + CQ_W * q, a = 0;
        while (idx >= nblk - a && a < nblk) {
            idx -= nblk - a;
            a++;
        }
        ti[q] = a;
        tj[q] = a + idx;
    }
    d4 acc[GR_MAXT];
#pragma unroll
    for (int q = 0; q < GR_MAXT; q++) acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    // Staging: thread t moves row pairs (2 * (t % 32), +1) of columns t / 32 + 16 j, j < 8, so one
    // 64-row chunk is eight 16-byte loads per thread, issued for the next chunk before this one's MFMAs.
    const int kp = 2 * (t & 31), c0 = t >> 5;
    double2 nxt[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = c0 + 16 * j, k = k0 + kp;
            double2 v = make_double2(0.0, 0.0);
            if (c < b) {
                const double* src = x + (int64_t)c * ldx + k;
                if (k + 1 < r1 && (((uintptr_t)src) & 15) == 0) {
                    v = *reinterpret_cast<const double2*>(src);
                } else {
                    if (k < r1) v.x = src[0];
                    if (k + 1 < r1) v.y = src[1];
                }
            }
            nxt[j] = v;
        }
    };
    if (r0 < r1) fetch(r0);
    for (int k0 = r0; k0 < r1; k0 += GR_ROWS) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = c0 + 16 * j;
            xs[c * GR_LDC + kp] = nxt[j].x;
            xs[c * GR_LDC + kp + 1] = nxt[j].y;
        }
        __syncthreads();
        if (k0 + GR_ROWS < r1) fetch(k0 + GR_ROWS);
#pragma unroll
        for (int q = 0; q < GR_MAXT; q++) {
            if (w + CQ_W * q < ntile) {
                const double* xa = xs + (16 * ti[q] + i16) * GR_LDC + g;
                const double* xb = xs + (16 * tj[q] + i16) * GR_LDC + g;
#pragma unroll
                for (int st = 0; st < GR_ROWS / 4; st++) acc[q] = mfma(acc[q], xa[4 * st], xb[4 * st]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < GR_MAXT; q++) {
        const int idx = w + CQ_W * q;
        if (idx < ntile) {
            double* o = part + ((int64_t)blockIdx.x * ntile + idx) * 256;
#pragma unroll
            for (int r = 0; r < 4; r++) o[(g + 4 * r) * 16 + i16] = acc[q][r];
        }
    }
}

// X^T Y for two column-major m x b matrices: as gram_partial, with both staged side by side and all
// nblk x nblk tiles (tile idx = ti * nblk + tj, up to eight per wavefront).
constexpr int GRC_MAXT = 8;
__global__ __launch_bounds__(CQ_NT) void gram_cross_partial(const double* __restrict__ x, int64_t ldx, int bx,
                                                            const double* __restrict__ y, int64_t ldy, int by, int m,
                                                            int b, int per, double* __restrict__ part) {
    extern __shared__ double xys[];  // two 128 x GR_LDC chunks: X, then Y
    double* xs = xys;
    double* ys = xys + 128 * GR_LDC;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, i16 = l & 15, g = l >> 4;
    const int nblk = (b + 15) >> 4, ntile = nblk * nblk;
    const int r0 = blockIdx.x * per, r1 = min(m, r0 + per);
    d4 acc[GRC_MAXT];
#pragma unroll
    for (int q = 0; q < GRC_MAXT; q++) acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    const int kp = 2 * (t & 31), c0 = t >> 5;
    double2 nx[8], ny[8];
    auto fetch1 = [&](const double* src0, int64_t ld, int nc, int k0, double2 (&o)[8]) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = c0 + 16 * j, k = k0 + kp;
            double2 v = make_double2(0.0, 0.0);
            if (c < nc) {
                const double* src = src0 + (int64_t)c * ld + k;
                if (k + 1 < r1 && (((uintptr_t)src) & 15) == 0) {
                    v = *reinterpret_cast<const double2*>(src);
                } else {
                    if (k < r1) v.x = src[0];
                    if (k + 1 < r1) v.y = src[1];
                }
            }
            o[j] = v;
        }
    };
    if (r0 < r1) {
        fetch1(x, ldx, bx, r0, nx);
        fetch1(y, ldy, by, r0, ny);
    }
    for (int k0 = r0; k0 < r1; k0 += GR_ROWS) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = c0 + 16 * j;
            xs[c * GR_LDC + kp] = nx[j].x;
            xs[c * GR_LDC + kp + 1] = nx[j].y;
            ys[c * GR_LDC + kp] = ny[j].x;
            ys[c * GR_LDC + kp + 1] = ny[j].y;
        }
        __syncthreads();
        if (k0 + GR_ROWS < r1) {
            fetch1(x, ldx, bx, k0 + GR_ROWS, nx);
            fetch1(y, ldy, by, k0 + GR_ROWS, ny);
        }
#pragma unroll
        for (int q = 0; q < GRC_MAXT; q++) {
            const int idx = w + CQ_W * q;
            if (idx < ntile) {
                const double* xa = xs + (16 * (idx / nblk) + i16) * GR_LDC + g;
                const double* yb = ys + (16 * (idx % nblk) + i16) * GR_LDC + g;
#pragma unroll
                for (int st = 0; st < GR_ROWS / 4; st++) acc[q] = mfma(acc[q], xa[4 * st], yb[4 * st]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < GRC_MAXT; q++) {
        const int idx = w + CQ_W * q;
        if (idx < ntile) {
            double* o = part + ((int64_t)blockIdx.x * ntile + idx) * 256;
#pragma unroll
            for (int r = 0; r < 4; r++) o[(g + 4 * r) * 16 + i16] = acc[q][r];
        }
    }
}

// Workgroup (tile, 16-entry group): thread (entry e, lane group pg) sums partials pg, pg + 16, ... of
// its entry, then the 16 sums are added in a fixed order (the result does not depend on timing).
__global__ __launch_bounds__(256) void gram_reduce(const double* __restrict__ part, int nparts, int b, int cross,
                                                   double* __restrict__ gm, int ldg, int bx, int by) {
    __shared__ double red[16][17];
    const int nblk = (b + 15) >> 4, ntile = cross ? nblk * nblk : nblk * (nblk + 1) / 2;
    const int idx = blockIdx.x / 16, eg = blockIdx.x % 16;
    const int e = threadIdx.x & 15, pg = threadIdx.x >> 4;
    const int ent = eg * 16 + e;  // entry of the 16 x 16 tile: row ent / 16, column ent % 16
    double v = 0.0;
    for (int p = pg; p < nparts; p += 16) v += part[((int64_t)p * ntile + idx) * 256 + ent];
    red[pg][e] = v;
    __syncthreads();
    if (threadIdx.x < 16) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < 16; q++) sum += red[q][threadIdx.x];
        int a, cb;
        if (cross) {
            a = idx / nblk;
            cb = idx % nblk;
        } else {
            int rem = idx;
            a = 0;
            while (rem >= nblk - a) {
                rem -= nblk - a;
                a++;
            }
            cb = a + rem;
        }
        const int i = 16 * a + (ent >> 4), j = 16 * cb + (ent & 15);
        if (i < bx && j < by) {
            gm[i * ldg + j] = sum;
            if (!cross) gm[j * ldg + i] = sum;
        }
    }
}

// T of the compact-WY form I - V T V^T of b reflectors (the reference's buildT, HouseholderQR.java:296-320,
// restated as T^-1 = diag(1 / tau) + strictly-upper(V^T V), the inverse of the recurrence T(1:i-1, i) =
// -tau_i T(1:i-1, 1:i-1) V(:, 1:i-1)^T v_i): G = V^T V and tau in, T (b x b row-major, upper) out. A
// reflector with tau_i = 0 (H_i = I) takes no part: its row and column of T^-1 are e_i and T(i, i) = 0.
// Dynamic LDS CQ_DYN.
__global__ __launch_bounds__(CQ_NT) void wy_t_inverse(const double* __restrict__ gm, int b,
                                                      const double* __restrict__ tau, double* __restrict__ tout,
                                                      int ldt) {
    __shared__ double dinv[128];
    __shared__ int zt[128];
    extern __shared__ double dyn[];
    double* S = dyn;
    double* dblk = dyn + 128 * SLD;
    const int t = threadIdx.x;
    const int nb16 = (b + 15) & ~15, nblk = nb16 >> 4;
    if (t < 128) {
        const bool z = t < b && tau[t] == 0.0;
        zt[t] = z ? 1 : 0;
        dinv[t] = t < b ? (z ? 1.0 : tau[t]) : 1.0;  // 1 / U_T(t, t)
    }
    __syncthreads();
    const int c = t & 127;
    // the diagonal entry of column c, 1 / tau_c (the one division of this thread)
    const double diag_c = (c < b && !zt[c]) ? 1.0 / tau[c] : 1.0;
    {
        double v[32];
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int r = (t >> 7) + 4 * q;
            v[q] = (r < c && c < b) ? gm[r * b + c] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int r = (t >> 7) + 4 * q;
            if (r >= nb16 || c >= nb16) continue;
            double x;
            if (r < b && c < b) {
                if (r == c) x = diag_c;
                else x = (r < c && !zt[r] && !zt[c]) ? v[q] : 0.0;
            } else {
                x = r == c ? 1.0 : 0.0;
            }
            S[r * SLD + c] = x;
        }
    }
    __syncthreads();
    upper_inverse_blocked(S, dblk, dinv, nblk);
    for (int r = t >> 7; r < b; r += 4)
        if (c < b) tout[r * ldt + c] = (r <= c && !(r == c && zt[r])) ? S[r * SLD + c] : 0.0;
}

// The reference's tau = 0 case: column i of Y exactly zero below the diagonal (Y1 part in spec, Y2
// rows b..mp-1 checked here) with a non-negative pivot: H_i = I, tau_i = 0 and R's row i keeps the
// pivot's sign (the row is negated back). One workgroup per panel column; the last column of a panel
// has no Y1 part, so about half the panels scan one column.
__global__ __launch_bounds__(1024) void cqr_special(double* __restrict__ ap, int64_t lda, int mp, int b,
                                                    double* __restrict__ tau, const int* __restrict__ spec) {
    const int i = blockIdx.x;
    if (spec[i] == 0) return;
    const double* col = ap + (int64_t)i * lda;
    int any = 0;
#pragma unroll 8
    for (int r = b + (int)threadIdx.x; r < mp; r += 1024) any |= col[r] != 0.0;
    any = __syncthreads_or(any);
    if (any) return;
    if (threadIdx.x == 0) tau[i] = 0.0;
    for (int c = i + (int)threadIdx.x; c < b; c += 1024) ap[(int64_t)c * lda + i] = -ap[(int64_t)c * lda + i];
}

}  // namespace

hipError_t launch_cqr_chol(const double* g, int b, int pass, const double* r1, double* r_out, double* x_out,
                           int* flag, hipStream_t stream) {
    if (b < 1 || b > 128 || (pass != 1 && pass != 2)) return hipErrorInvalidValue;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)cqr_chol, hipFuncAttributeMaxDynamicSharedMemorySize, CQ_DYN);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(cqr_chol, dim3(1), dim3(CQ_NT), CQ_DYN, stream, g, b, pass, r1, r_out, x_out, flag);
    return hipGetLastError();
}

hipError_t launch_cqr_lu(const double* q, int64_t ldq, int b, const double* r, double* ap, int64_t lda, double* tau,
                         double* ui, int* spec, hipStream_t stream) {
    if (b < 1 || b > 128 || ldq < b) return hipErrorInvalidValue;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)cqr_lu, hipFuncAttributeMaxDynamicSharedMemorySize, CQ_DYN);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(cqr_lu, dim3(1), dim3(CQ_NT), CQ_DYN, stream, q, ldq, b, r, ap, lda, tau, ui, spec);
    return hipGetLastError();
}

int gram_parts(int64_t m) {
    const int64_t p = (m + 127) / 128;
    return (int)(p < 1 ? 1 : (p > 256 ? 256 : p));
}

size_t gram_work_doubles(int64_t m) { return (size_t)gram_parts(m) * 36 * 256; }

hipError_t launch_gram(const double* x, int64_t ldx, int m, int b, double* part, double* g, hipStream_t stream) {
    if (b < 1 || b > 128 || m < 1 || ldx < m) return hipErrorInvalidValue;
    // Row ranges of whole 64-row chunks (16-byte aligned loads when X is), at most gram_parts(m) of them.
    const int P0 = gram_parts(m);
    const int per = (int)(((m + P0 - 1) / P0 + GR_ROWS - 1) / GR_ROWS * GR_ROWS);
    const int P = (m + per - 1) / per;
    hipLaunchKernelGGL(gram_partial, dim3(P), dim3(CQ_NT), 0, stream, x, ldx, m, b, per, part);
    const int nblk = (b + 15) >> 4, ntile = nblk * (nblk + 1) / 2;
    hipLaunchKernelGGL(gram_reduce, dim3((unsigned)(16 * ntile)), dim3(256), 0, stream, part, P, b, 0, g, b, b, b);
    return hipGetLastError();
}

size_t gram_cross_work_doubles(int64_t m) { return (size_t)gram_parts(m) * 64 * 256; }

hipError_t launch_gram_cross(const double* x, int64_t ldx, int bx, const double* y, int64_t ldy, int by, int m,
                             double* part, double* g, int ldg, hipStream_t stream) {
    if (bx < 1 || bx > 128 || by < 1 || by > 128 || m < 1 || ldx < m || ldy < m || ldg < by)
        return hipErrorInvalidValue;
    const int b = bx > by ? bx : by;
    const int P0 = gram_parts(m);
    const int per = (int)(((m + P0 - 1) / P0 + GR_ROWS - 1) / GR_ROWS * GR_ROWS);
    const int P = (m + per - 1) / per;
    constexpr int lds = 2 * 128 * GR_LDC * 8;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)gram_cross_partial, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(gram_cross_partial, dim3(P), dim3(CQ_NT), lds, stream, x, ldx, bx, y, ldy, by, m, b, per,
                       part);
    const int nblk = (b + 15) >> 4;
    hipLaunchKernelGGL(gram_reduce, dim3((unsigned)(16 * nblk * nblk)), dim3(256), 0, stream, part, P, b, 1, g, ldg,
                       bx, by);
    return hipGetLastError();
}

hipError_t launch_wy_t(const double* g, int b, const double* tau, double* T, int ldt, hipStream_t stream) {
    if (b < 1 || b > 128 || ldt < b) return hipErrorInvalidValue;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)wy_t_inverse, hipFuncAttributeMaxDynamicSharedMemorySize, CQ_DYN);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(wy_t_inverse, dim3(1), dim3(CQ_NT), CQ_DYN, stream, g, b, tau, T, ldt);
    return hipGetLastError();
}

hipError_t launch_cqr_special(double* ap, int64_t lda, int mp, int b, double* tau, const int* spec,
                              hipStream_t stream) {
    if (b < 1 || b > 128 || mp < b) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cqr_special, dim3(b), dim3(1024), 0, stream, ap, lda, mp, b, tau, spec);
    return hipGetLastError();
}

}  // namespace jlc
