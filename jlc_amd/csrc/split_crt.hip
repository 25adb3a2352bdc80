// CRT split form of DGEMM on INT8 MFMA (gfx950): modular integer products
// recombined by the Chinese remainder theorem (Ozaki, Uchino, Imamura 2025,
// "Ozaki Scheme II: a GEMM-oriented emulation of floating-point matrix
// multiplication using an integer modular technique").
//
//  1. Scaling. Row i of op(A) is mapped to integers a' = rint(a * 2^(pa-ta[i]))
//     with 2^ta[i] a power of two at or above the row's 2-norm (crt_norm_*),
//     so ||a'||_2 <= 2^pa + sqrt(k)/2; column j of op(B) likewise with pb,
//     tb[j]. Power-of-two scaling is exact, and the only rounding of the whole
//     method is this one per input entry (plus the final conversion of the
//     exact integer result to FP64). Scaling by the norm rather than the max
//     lets the moduli count follow Cauchy-Schwarz, |C'_ij| <= ||a'_i|| ||b'_j||,
//     instead of the k * 2^(pa+pb) worst case.
//  2. Residues. For N pairwise coprime moduli m_t <= 256 the symmetric
//     residues a' mod m_t and b' mod m_t fit in int8. They are written in the
//     MFMA fragment order of the slice form: one 1 KiB block per (modulus,
//     32-row block, 32-deep k chunk), lane l holding row l%32, k in
//     [16*(l/32), +16). Within a modulus' plane the blocks are k-major
//     (block (rb, kc) at (kc * row_blocks + rb) KiB), so the operand strips
//     of neighbouring C tiles for one k step are contiguous: with row-block-
//     major planes a tile's 8 row blocks sat kpad KiB apart (512 KiB at
//     K = 16384), on the same few L2 sets, and the sharing between the tiles
//     resident on one XCD was lost to conflict misses.
//  3. Products. Each modulus is one INT8 GEMM with exact INT32 accumulation
//     (|sum| <= k * 2^14, folded modulo m_t every 65536 k when k is larger);
//     the epilogue reduces the sum modulo m_t and stores the int8 residue of
//     C' = A'B' in the accumulator's own lane order (1 KiB per 32 x 32 block).
//     Unlike the slice form there is no operand reuse between moduli, so the
//     GEMM is a plain large-tile INT8 GEMM: 256 x 256 C tile, 8 wavefronts of
//     128 x 64, K staged 64 deep through two LDS buffers, global loads four
//     stages ahead (crt_gemm_i8).
//  4. Reconstruction. With M = prod m_t > 4 max ||a'_i|| ||b'_j||, C' is the unique
//     integer in (-M/2, M/2) with those residues. X = sum_t r_t W_t
//     (W_t = (M/m_t) * ((M/m_t)^-1 mod m_t)) is formed exactly in the 40-bit
//     limbs M needs (every partial sum is an integer below 2^53), q = rint(X/M),
//     C' = X - qM limb by limb (exact fma), the limbs are balanced and summed
//     from the top in double-double and rounded once; then
//     C = alpha * C' * 2^(ta+tb-pa-pb) (+ beta*C through fma).
//
// Rows / columns holding Inf or NaN give NaN outputs, as in the slice form.
#include <math.h>

#include "internal.h"
#include "split_stats.h"

// The reconstruction must round exactly as written: no contraction of a*b+c.
#pragma clang fp contract(off)

namespace jlc {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int NONFINITE = STAT_NONFINITE;
constexpr int CT = 256;        // C tile rows and columns
constexpr int KSTAGE = 64;     // K per LDS stage (two 32-deep MFMA steps)
constexpr int KPAD = 256;      // K padding: a multiple of four stages
constexpr int FOLD_STEPS = 2048;  // 32-deep steps between modular folds (65536 k)

// Symmetric residue of an integer-valued double |y| < 2^53 modulo m.
__device__ __forceinline__ double red(double y, double m, double inv_m, double hi_lim, double lo_lim) {
    const double q = rint(y * inv_m);
    double r = fma(-q, m, y);
    r = r > hi_lim ? r - m : r;
    r = r < lo_lim ? r + m : r;
    return r;
}

// Norm exponents. For row r with max exponent e (frexp; NONFINITE for Inf/NaN
// rows) each entry contributes u(x) = ceil(fl(v*v) * 2^40) + 1 units of 2^-40,
// v = |x| * 2^-e in [0, 1) (0 for x = 0): an upper bound of v^2 * 2^40, since
// fl(v*v) is at most 2^-53 relatively below v*v. The integer sum S bounds
// ||x||^2 <= S * 2^(2e-40) whatever the summation order, so
//   t = e + ceil((L - 40) / 2),  L = bit length of S - 1  (2^L >= S)
// is the same on every device and ||x|| <= 2^t. S < 2^63 while k < 2^22.
// sig = log2(||x|| / sqrt(nnz)) - mean log2|x| (rms over geometric mean of the
// non-zero entries, from S and the spread e - mean log2|x|) feeds the bit
// chooser.
struct NormPart {
    unsigned long long s;
    unsigned int cnt, pad;
};

__device__ __forceinline__ unsigned long long norm_units(double x, int e) {
    const double v = ldexp(fabs(x), -e);
    return v > 0.0 ? (unsigned long long)ceil(v * v * 0x1p40) + 1ull : 0ull;
}

__device__ __forceinline__ void norm_finish(unsigned long long S, unsigned int cnt, int e, float spread, int* t,
                                            float* sig) {
    if (e == NONFINITE || S == 0ull) {
        *t = e == NONFINITE ? NONFINITE : 0;
        *sig = 0.0f;
        return;
    }
    const int L = S == 1ull ? 0 : 64 - __clzll((long long)(S - 1ull));
    *t = e + (L >= 40 ? (L - 39) >> 1 : -((40 - L) >> 1));
    *sig = 0.5f * (log2f((float)S) - 40.0f) + spread - 0.5f * log2f((float)cnt);
}

// k-contiguous rows (element (r, k) at X[r*sr + k]); one wavefront per row.
__global__ __launch_bounds__(256) void crt_norm_rows(const double* __restrict__ X, int64_t sr, int R, int Kd,
                                                     const int* __restrict__ e, const float* __restrict__ spread,
                                                     int* __restrict__ t, float* __restrict__ sig) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= R) return;
    const int ex = e[r];
    unsigned long long S = 0ull;
    unsigned int cnt = 0u;
    if (ex != NONFINITE) {
        const double* row = X + (int64_t)r * sr;
        for (int k = lane; k < Kd; k += 64) {
            const unsigned long long u = norm_units(row[k], ex);
            S += u;
            cnt += u != 0ull;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        S += __shfl_xor(S, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
    }
    if (lane == 0) norm_finish(S, cnt, ex, spread[r], &t[r], &sig[r]);
}

// Rows r contiguous across r (element (r, k) at X[k*sk + r]): pass 1 sums one
// of kparts K ranges for 64 rows per block (wave w takes every fourth k), pass
// 2 adds the ranges. Integer sums: the result does not depend on the split.
__global__ __launch_bounds__(256) void crt_norm_cols_partial(const double* __restrict__ X, int64_t sk, int R, int Kd,
                                                             int krange, const int* __restrict__ e,
                                                             NormPart* __restrict__ part) {
    __shared__ NormPart red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    const int k0 = blockIdx.y * krange;
    const int k1 = min(Kd, k0 + krange);
    unsigned long long S = 0ull;
    unsigned int cnt = 0u;
    const int ex = r < R ? e[r] : NONFINITE;
    if (ex != NONFINITE) {
        const double* col = X + r;
        for (int k = k0 + w; k < k1; k += 4) {
            const unsigned long long u = norm_units(col[(int64_t)k * sk], ex);
            S += u;
            cnt += u != 0ull;
        }
    }
    red[w][lane] = NormPart{S, cnt, 0u};
    __syncthreads();
    if (w == 0 && r < R) {
        NormPart acc = red[0][lane];
#pragma unroll
        for (int i = 1; i < 4; i++) {
            acc.s += red[i][lane].s;
            acc.cnt += red[i][lane].cnt;
        }
        part[(int64_t)blockIdx.y * R + r] = acc;
    }
}

__global__ __launch_bounds__(256) void crt_norm_cols_finish(const NormPart* __restrict__ part, int R, int kparts,
                                                            const int* __restrict__ e,
                                                            const float* __restrict__ spread, int* __restrict__ t,
                                                            float* __restrict__ sig) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    unsigned long long S = 0ull;
    unsigned int cnt = 0u;
    for (int kp = 0; kp < kparts; kp++) {
        S += part[(int64_t)kp * R + r].s;
        cnt += part[(int64_t)kp * R + r].cnt;
    }
    norm_finish(S, cnt, e[r], spread[r], &t[r], &sig[r]);
}

// Statistics and norm exponent of k-contiguous rows in one read of X: one
// workgroup per row, the row held in registers (Kd <= ROW_THREADS * ROW_REG). The
// same per-entry statistics as the split forms' row statistics (max, log2
// sum, count, Inf/NaN) give e and spread, then the integer norm units of the
// same registers give t and sig; the max and the integer sums do not depend on
// the reduction order, so e and t are those of the two-pass kernels.
constexpr int ROW_THREADS = 512, ROW_WAVES = ROW_THREADS / 64, ROW_REG = 32;

__global__ __launch_bounds__(ROW_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void crt_row_stats_norms(const double* __restrict__ X, int64_t sr, int Kd,
                                                           int* __restrict__ e, float* __restrict__ spread,
                                                           int* __restrict__ t, float* __restrict__ sig) {
    __shared__ double smx[ROW_WAVES], sslog[ROW_WAVES], scnt[ROW_WAVES];
    __shared__ int sbad[ROW_WAVES];
    __shared__ unsigned long long sS[ROW_WAVES];
    __shared__ unsigned int sc[ROW_WAVES];
    const int r = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double* row = X + (int64_t)r * sr;
    double v[ROW_REG];
#pragma unroll
    for (int j = 0; j < ROW_REG; j++) {
        const int k = threadIdx.x + ROW_THREADS * j;
        v[j] = k < Kd ? row[k] : 0.0;
    }
    double mx = 0.0, slog = 0.0, cnt = 0.0;
    bool bad = false;
#pragma unroll
    for (int j = 0; j < ROW_REG; j++) stat_add(fabs(v[j]), mx, slog, cnt, bad);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off, 64));
        slog += __shfl_xor(slog, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
        bad |= __shfl_xor((int)bad, off, 64) != 0;
    }
    if (lane == 0) {
        smx[w] = mx;
        sslog[w] = slog;
        scnt[w] = cnt;
        sbad[w] = bad;
    }
    __syncthreads();
    mx = smx[0];
    slog = sslog[0];
    cnt = scnt[0];
    int anybad = sbad[0];
#pragma unroll
    for (int i = 1; i < ROW_WAVES; i++) {
        mx = fmax(mx, smx[i]);
        slog += sslog[i];
        cnt += scnt[i];
        anybad |= sbad[i];
    }
    bad = anybad != 0;
    int ex = 0;
    if (mx > 0.0) frexp(mx, &ex);
    if (bad) ex = NONFINITE;
    unsigned long long S = 0ull;
    unsigned int n = 0u;
    if (ex != NONFINITE) {
#pragma unroll
        for (int j = 0; j < ROW_REG; j++) {
            const unsigned long long u = norm_units(v[j], ex);
            S += u;
            n += u != 0ull;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        S += __shfl_xor(S, off, 64);
        n += __shfl_xor(n, off, 64);
    }
    if (lane == 0) {
        sS[w] = S;
        sc[w] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        S = 0ull;
        n = 0u;
#pragma unroll
        for (int i = 0; i < ROW_WAVES; i++) {
            S += sS[i];
            n += sc[i];
        }
        const float sp = (bad || cnt == 0.0) ? 0.0f : (float)(ex - slog / cnt);
        e[r] = ex;
        spread[r] = sp;
        norm_finish(S, n, ex, sp, &t[r], &sig[r]);
    }
}

// Residues through signed byte digits. An integer image |s| < 2^66 is split
// into three signed 22-bit chunks (s = c2 2^44 + c1 2^22 + c0, each with the
// sign of s) and each chunk into balanced base-256 digits (c = d0 + d1 2^8 +
// d2 2^16, |d| <= 128), nine digits in all: eight packed four to a register,
// the top one (c2's d2) kept as a plain int. Then x = sum_d digit * (2^weight
// mod m) is the top digit's 24-bit product followed by two v_dot4 int8 dot
// products, exact with |x| <= 9 * 128 * 128 < 2^18, and congruent to s mod m.
// The reduction stays in three full-rate instructions:
//   tq = fma(x, fl(1/m), 1.5 * 2^23) = 1.5 * 2^23 + q,  q = rint(x / m),
// since |x * fl(1/m)| < 2^11 keeps the sum in [2^23, 2^24) where the FP32
// spacing is 1; the product is within 2^-14 of x / m, under the 1/(2m)
// distance of x / m from any half integer for odd m (m = 256 is exact), so q
// is rint(x / m). The low 24 bits of tq are 2^22 + q (bit 23 clear), and
//   mad_i24(bits(tq), -m, x) = x - q m - 2^22 m
// has, 2^22 m being a multiple of 256, the two's-complement byte of the
// symmetric residue r = x - q m as its low byte (for m = 256, r = +-128 both
// give the byte of -128). v_perm gathers four such low bytes into a word.
__device__ __forceinline__ void balanced_bytes(int c, int& d0, int& d1, int& d2) {
    d0 = ((c + 128) & 255) - 128;
    c = (c - d0) >> 8;
    d1 = ((c + 128) & 255) - 128;
    d2 = (c - d1) >> 8;
}

__device__ __forceinline__ int byte4(int a, int b, int c, int d) {
    return (int)((unsigned)(a & 255) | ((unsigned)(b & 255) << 8) | ((unsigned)(c & 255) << 16) |
                 ((unsigned)(d & 255) << 24));
}

// Digits of one image: D0 = (c0: d0 d1 d2, c1: d0), D1 = (c1: d1 d2, c2: d0 d1) packed, E2 = c2's d2.
__device__ __forceinline__ void image_digits(int c2, int c1, int c0, int& D0, int& D1, int& E2) {
    int a0, a1, a2, b0, b1, b2, e0, e1;
    balanced_bytes(c0, a0, a1, a2);
    balanced_bytes(c1, b0, b1, b2);
    balanced_bytes(c2, e0, e1, E2);
    D0 = byte4(a0, a1, a2, b0);
    D1 = byte4(b1, b2, e0, e1);
}

// Low bytes of a, b, c, d as one word (a in the lowest byte). Selector bytes
// 0-3 pick bytes of the second operand, 4-7 of the first, 0x0c gives zero.
__device__ __forceinline__ int low_bytes4(int a, int b, int c, int d) {
    const unsigned lo = __builtin_amdgcn_perm((unsigned)b, (unsigned)a, 0x0c0c0400u);
    const unsigned hi = __builtin_amdgcn_perm((unsigned)d, (unsigned)c, 0x0c0c0400u);
    return (int)__builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

__device__ __forceinline__ v4i residues16(const int (&D0)[16], const int (&D1)[16], const int (&E2)[16],
                                          const CrtTables& T, int t) {
    const int K0 = T.kdig[t][0], K1 = T.kdig[t][1];
    const int w2 = (int)(signed char)(T.kdig[t][2] & 255);
    const int negm = -(int)T.mod[t];
    const float invf = T.invf[t];
    int q[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int x = __builtin_amdgcn_sdot4(D0[j], K0, __builtin_amdgcn_sdot4(D1[j], K1, __mul24(E2[j], w2), false),
                                             false);
        const float tq = fmaf((float)x, invf, 12582912.0f);
        q[j] = __mul24(__float_as_int(tq), negm) + x;
    }
    v4i w;
    w.x = low_bytes4(q[0], q[1], q[2], q[3]);
    w.y = low_bytes4(q[4], q[5], q[6], q[7]);
    w.z = low_bytes4(q[8], q[9], q[10], q[11]);
    w.w = low_bytes4(q[12], q[13], q[14], q[15]);
    return w;
}

// Residue planes of op(X) (element (r, k) at X[r*sr + k*sk], R x Kd, zero
// padded): one wavefront per 1 KiB block, taken row block by row block (the
// reads of neighbouring wavefronts are neighbours in X); lane l handles row
// l%32, k in [16*(l/32), +16) and writes 16 bytes per modulus to block
// (rb, kc) at kc * row_blocks + rb of the k-major plane.
// VEC: k-contiguous rows (sk == 1), 16-byte aligned, Kd a multiple of 16:
// each lane reads its 16 doubles as eight 16-byte loads.
template <bool VEC>
__global__ __launch_bounds__(256) void crt_pack(const double* __restrict__ X, int64_t sr, int64_t sk, int R, int Kd,
                                                int kchunks, int nblocks, const int* __restrict__ e, int p,
                                                CrtTables T, v4i* __restrict__ out, int64_t plane_stride) {
    const int lane = threadIdx.x & 63;
    const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blk >= nblocks) return;
    const int rb = blk / kchunks, kc = blk % kchunks;
    const int r = rb * 32 + (lane & 31);
    const int k0 = kc * 32 + 16 * (lane >> 5);
    const int ex = r < R ? e[r] : 0;
    const bool live = r < R && ex != NONFINITE;
    // Loads use clamped, always-valid indices and the padding is selected to
    // zero afterwards, so the element loop has no branches.
    const int64_t rowoff = (int64_t)(r < R ? r : R - 1) * sr;
    double xv[16];
    if (VEC) {
        const int kb = k0 < Kd ? k0 : Kd - 16;
#pragma unroll
        for (int j = 0; j < 16; j += 2) {
            const double2 v = *reinterpret_cast<const double2*>(X + rowoff + kb + j);
            xv[j] = v.x;
            xv[j + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int k = k0 + j;
            xv[j] = X[rowoff + (int64_t)(k < Kd ? k : Kd - 1) * sk];
        }
    }
    int D0[16], D1[16], E2[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int k = k0 + j;
        double s = ldexp(xv[j], p - ex);
        s = fabs(s) < 9007199254740992.0 ? rint(s) : s;
        s = (live && k < Kd) ? s : 0.0;
        // s = c2 * 2^44 + c1 * 2^22 + c0 exactly; every chunk has the sign of s and |chunk| < 2^22.
        const double h = trunc(s * 0x1p-44);
        const double rest = fma(-h, 0x1p44, s);
        const double mid = trunc(rest * 0x1p-22);
        image_digits((int)h, (int)mid, (int)fma(-mid, 0x1p22, rest), D0[j], D1[j], E2[j]);
    }
    const int rblocks = nblocks / kchunks;
    const int64_t off = ((int64_t)kc * rblocks + rb) * 64 + lane;
    for (int t = 0; t < T.nmod; t++) out[t * plane_stride + off] = residues16(D0, D1, E2, T, t);
}

// Column strips of `group` tile rows (256 x 256 tiles), planes in blockIdx.y.
// With xcd set (and a tile count divisible by 8), workgroup b, which the
// dispatcher places on XCD b % 8, takes tile (b % 8) * (count / 8) + b / 8 of
// that order, so the workgroups resident on one XCD (and sharing its L2) work
// on neighbouring tiles.
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int& tm, int& tn, int group = 16,
                                            bool xcd = false) {
    const int count = tiles_m * tiles_n;
    if (xcd && count % 8 == 0) bid = (bid & 7) * (count >> 3) + (bid >> 3);
    const int GROUP = group;
    const int per_group = GROUP * tiles_n;
    const int g = bid / per_group;
    const int first_m = g * GROUP;
    const int gsize = min(tiles_m - first_m, GROUP);
    const int in = bid % per_group;
    tm = first_m + in % gsize;
    tn = in / gsize;
}

__device__ __forceinline__ v16i reduce_acc(v16i acc, double m, double inv_m, double hl, double ll) {
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = (int)red((double)acc[r], m, inv_m, hl, ll);
    return acc;
}

// One INT8 GEMM per modulus (blockIdx.y); C residues out in accumulator order.
// Wavefront w: rows (w>>2)*128 + [0,128) as 4 blocks, columns (w&3)*64 + [0,64) as 2 blocks.
// K runs in 64-deep stages (a multiple of four of them: K is padded to 256).
// Global loads run four stages ahead in four register sets, so a load has
// four stages of MFMA work to arrive (two stages left the MFMA pipe idle on
// load latency: MFMA busy 58% of cycles against 64% with four); stage s is
// multiplied from one of two 32 KiB LDS buffers while stage s+1 is written to
// the other, one barrier per stage. The loop is unrolled by four so that the
// register sets have static names, and the modular folds sit between
// 65536-deep K ranges, outside it.
constexpr int NBUF_LDS = 2;
constexpr int STAGE_V4 = 32 * 64;  // v4i per stage (32 blocks of 1 KiB)
constexpr int PREFETCH = 4;        // register sets (stages in flight)

__device__ __forceinline__ void fold_acc(v16i& acc, int m, float invf) {
    // acc -= rint(acc / m) * m with an FP32 quotient (off by a few units at
    // most): |acc| < 4m afterwards, far inside the INT32 headroom the next
    // 65536-deep range needs (k * 2^14 = 2^30).
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] -= (int)rintf((float)acc[r] * invf) * m;
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void crt_gemm_i8(
    int kchunks, const v4i* __restrict__ Ap, int64_t a_plane, const v4i* __restrict__ Bp, int64_t b_plane,
    v4i* __restrict__ Cr, int64_t c_plane, int cblocks_n, int tiles_m, int tiles_n, int group, int xcd, CrtTables T) {
    // row blocks of the A and B planes (k-major: block (rb, kc) at kc * row_blocks + rb)
    const int rba = tiles_m * 8, rbb = tiles_n * 8;
    extern __shared__ v4i lds[];  // [NBUF_LDS][32 blocks: A kc*8+rb (0..15), B 16 + kc*8+cb][64 lanes]
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w >> 2, wn = w & 3;
    const int t = blockIdx.y;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn, group, xcd != 0);
    const v4i* A = Ap + t * a_plane + (int64_t)tm * 8 * 64;
    const v4i* B = Bp + t * b_plane + (int64_t)tn * 8 * 64;
    // Loads of a stage: 32 blocks of 1 KiB; wavefront w moves blocks w, w+8 (A) and w+16, w+24 (B).
    const v4i* src[4];
    int dst[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int b = w + 8 * i;  // 0..15 A, 16..31 B; block b: row/col block (b&15)>>1, kc b&1
        const int rb = (b & 15) >> 1, kc = b & 1;
        src[i] = (i < 2 ? A : B) + ((int64_t)kc * (i < 2 ? rba : rbb) + rb) * 64 + lane;
        dst[i] = ((b & 16) + kc * 8 + rb) * 64 + lane;
    }
    const int stages = kchunks / 2;
    const int64_t sa = (int64_t)2 * rba * 64, sb = (int64_t)2 * rbb * 64;  // one stage (two k chunks)
    v4i r[PREFETCH][4];
    auto load = [&](int s, v4i (&x)[4]) {
#pragma unroll
        for (int i = 0; i < 4; i++) x[i] = src[i][(int64_t)s * (i < 2 ? sa : sb)];
    };
    auto store = [&](int buf, const v4i (&x)[4]) {
        v4i* base = lds + buf * STAGE_V4;
#pragma unroll
        for (int i = 0; i < 4; i++) base[dst[i]] = x[i];
    };
    v16i acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double m = T.mod[t], inv_m = T.inv[t], hl = T.hi_lim[t], ll = T.lo_lim[t];

    auto compute = [&](int buf) {
        const v4i* base = lds + buf * STAGE_V4 + lane;
#pragma unroll
        for (int kc = 0; kc < 2; kc++) {
            v4i af[4], bf[2];
#pragma unroll
            for (int i = 0; i < 4; i++) af[i] = base[(kc * 8 + wm * 4 + i) * 64];
#pragma unroll
            for (int j = 0; j < 2; j++) bf[j] = base[(16 + kc * 8 + wn * 2 + j) * 64];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    };

    // Stage q waits in register set q % 4 and is multiplied from LDS buffer
    // q % 2. Iteration s: multiply stage s, write stage s+1 to LDS, refill
    // its register set with stage s+5 (clamped: the repeats of the last stage
    // land in a set nobody stores), barrier. Branch-free, so the compiler's
    // wait counts keep the later sets' loads in flight.
    const int last = stages - 1;
#pragma unroll
    for (int q = 0; q < PREFETCH; q++) load(min(q, last), r[q]);
    store(0, r[0]);
    load(min(PREFETCH, last), r[0]);
    __syncthreads();
    constexpr int FOLD_STAGES = FOLD_STEPS / 2;  // a multiple of 4
    const int mi = (int)m;
    const float invf = T.invf[t];
    for (int s0 = 0; s0 < stages; s0 += FOLD_STAGES) {
        const int s1 = min(s0 + FOLD_STAGES, stages);
        for (int s = s0; s < s1; s += PREFETCH) {
#pragma unroll
            for (int u = 0; u < PREFETCH; u++) {
                compute(u & 1);
                store((u + 1) & 1, r[(u + 1) & 3]);
                load(min(s + u + PREFETCH + 1, last), r[(u + 1) & 3]);
                __syncthreads();
            }
        }
        if (s1 < stages) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) fold_acc(acc[i][j], mi, invf);
        }
    }
    // Residues of C' in accumulator order: block (row block, col block) of the plane.
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const v16i r = reduce_acc(acc[i][j], m, inv_m, hl, ll);
            v4i o;
            o.x = (r[0] & 255) | ((r[1] & 255) << 8) | ((r[2] & 255) << 16) | ((r[3] & 255) << 24);
            o.y = (r[4] & 255) | ((r[5] & 255) << 8) | ((r[6] & 255) << 16) | ((r[7] & 255) << 24);
            o.z = (r[8] & 255) | ((r[9] & 255) << 8) | ((r[10] & 255) << 16) | ((r[11] & 255) << 24);
            o.w = (r[12] & 255) | ((r[13] & 255) << 8) | ((r[14] & 255) << 16) | ((r[15] & 255) << 24);
            const int64_t rbc = tm * 8 + wm * 4 + i, cbc = tn * 8 + wn * 2 + j;
            Cr[t * c_plane + (rbc * cblocks_n + cbc) * 64 + lane] = o;
        }
}

constexpr int W4_NBUF = 3;  // LDS ring of the four-wavefront kernel

// The same product with four wavefronts of 128 x 128 (2 x 2 over the 256 x 256
// tile, one wavefront per SIMD): each fragment read from LDS feeds four MFMAs
// instead of two or four, halving the LDS read traffic per operation (the
// 256 accumulator registers live in the accumulation register file). Stage
// loads: 32 blocks of 1 KiB, wavefront w moves blocks w + 4i (i < 4: A, else B).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void crt_gemm_i8_w4(
    int kchunks, const v4i* __restrict__ Ap, int64_t a_plane, const v4i* __restrict__ Bp, int64_t b_plane,
    v4i* __restrict__ Cr, int64_t c_plane, int cblocks_n, int tiles_m, int tiles_n, CrtTables T) {
    extern __shared__ v4i lds[];  // [W4_NBUF][32 blocks: A kc*8+rb (0..15), B 16 + kc*8+cb][64 lanes]
    // An (empty) use of the accumulation registers: without one the compiler
    // assumes the kernel needs none and keeps the 256 accumulators in the
    // 256 ordinary vector registers, spilling everything else.
    asm volatile("" ::: "a0");
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w >> 1, wn = w & 1;
    const int t = blockIdx.y;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    // block w + 4i of a stage: operand i >> 2, row/col block (w >> 1) + 2 (i & 3), k chunk w & 1; planes
    // k-major (block (rb, kc) at kc * row_blocks + rb)
    const int rba = tiles_m * 8, rbb = tiles_n * 8;
    constexpr int64_t blk_stride = 2 * 64;  // two row blocks apart
    const v4i* srcA = Ap + t * a_plane + ((int64_t)(w & 1) * rba + tm * 8 + (w >> 1)) * 64 + lane;
    const v4i* srcB = Bp + t * b_plane + ((int64_t)(w & 1) * rbb + tn * 8 + (w >> 1)) * 64 + lane;
    const int dst0 = ((w & 1) * 8 + (w >> 1)) * 64 + lane;  // + 2i blocks; B at + 16 blocks
    const int stages = kchunks / 2;
    const int64_t sa = (int64_t)2 * rba * 64, sb = (int64_t)2 * rbb * 64;  // one stage (two k chunks)
    v4i r0[8], r1[8];
    auto load = [&](int s, v4i (&r)[8]) {
        const int64_t oa = (int64_t)s * sa, ob = (int64_t)s * sb;
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = srcA[oa + i * blk_stride];
#pragma unroll
        for (int i = 0; i < 4; i++) r[4 + i] = srcB[ob + i * blk_stride];
    };
    auto store = [&](int buf, const v4i (&r)[8]) {
        v4i* base = lds + buf * STAGE_V4 + dst0;
#pragma unroll
        for (int i = 0; i < 4; i++) base[i * 128] = r[i];
#pragma unroll
        for (int i = 0; i < 4; i++) base[(16 + 2 * i) * 64] = r[4 + i];
    };
    v16i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double m = T.mod[t], inv_m = T.inv[t], hl = T.hi_lim[t], ll = T.lo_lim[t];

    auto compute = [&](int buf) {
        const v4i* base = lds + buf * STAGE_V4 + lane;
#pragma unroll
        for (int kc = 0; kc < 2; kc++) {
            v4i af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; i++) af[i] = base[(kc * 8 + wm * 4 + i) * 64];
#pragma unroll
            for (int j = 0; j < 4; j++) bf[j] = base[(16 + kc * 8 + wn * 4 + j) * 64];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    };
    // Fold between 65536-deep K ranges: acc -= rint(acc / m) * m in integer
    // arithmetic (the FP32 quotient may be off by a few units, leaving
    // |acc| < 4m, far inside the INT32 headroom the next range needs).
    const int mi = (int)m;
    const float invf = T.invf[t];
    auto fold = [&]() {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[i][j][r] -= (int)rintf((float)acc[i][j][r] * invf) * mi;
    };

    // Same branch-free three-buffer pipeline as crt_gemm_i8, with the folds
    // outside the inner loop.
    load(0, r0);
    store(0, r0);
    load(1, r1);
    __syncthreads();
    const int last = stages - 1;
    constexpr int FOLD_STAGES = FOLD_STEPS / 2;
    int buf = 0;
    for (int s0 = 0; s0 < stages; s0 += FOLD_STAGES) {
        const int s1 = min(s0 + FOLD_STAGES, stages);
        for (int s = s0; s < s1; s += 2) {
            load(min(s + 2, last), r0);
            compute(buf);
            int nb = buf == W4_NBUF - 1 ? 0 : buf + 1;
            store(nb, r1);
            __syncthreads();
            buf = nb;
            load(min(s + 3, last), r1);
            compute(buf);
            nb = buf == W4_NBUF - 1 ? 0 : buf + 1;
            store(nb, r0);
            __syncthreads();
            buf = nb;
        }
        if (s1 < stages) fold();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const v16i r = reduce_acc(acc[i][j], m, inv_m, hl, ll);
            v4i o;
            o.x = (r[0] & 255) | ((r[1] & 255) << 8) | ((r[2] & 255) << 16) | ((r[3] & 255) << 24);
            o.y = (r[4] & 255) | ((r[5] & 255) << 8) | ((r[6] & 255) << 16) | ((r[7] & 255) << 24);
            o.z = (r[8] & 255) | ((r[9] & 255) << 8) | ((r[10] & 255) << 16) | ((r[11] & 255) << 24);
            o.w = (r[12] & 255) | ((r[13] & 255) << 8) | ((r[14] & 255) << 16) | ((r[15] & 255) << 24);
            const int64_t rbc = tm * 8 + wm * 4 + i, cbc = tn * 8 + wn * 4 + j;
            Cr[t * c_plane + (rbc * cblocks_n + cbc) * 64 + lane] = o;
        }
}

// 40-bit limbs that hold M = prod of the first NM moduli (M < 2^(40 L):
// log2 M is 39.9, 79.2, 117.8 and 155.4 for 5, 10, 15 and 20 moduli); the
// limbs of M and of every W_t at and above L are zero, so the reconstruction
// over L limbs gives the same bits as over four.
constexpr int crt_limbs(int nm) { return nm <= 5 ? 1 : nm <= 10 ? 2 : nm <= 15 ? 3 : 4; }

// CRT reconstruction of one residue vector (exact up to the final rounding).
template <int NM>
__device__ __forceinline__ double crt_value(const int (&r)[NM], const CrtTables& T) {
    constexpr int L = crt_limbs(NM);
    double X[L];
#pragma unroll
    for (int j = 0; j < L; j++) X[j] = 0.0;
    // Every product and partial sum here is an integer below 2^53, so the fused
    // form is exact and equal to the separate one.
#pragma unroll
    for (int t = 0; t < NM; t++) {
        const double rt = (double)r[t];
#pragma unroll
        for (int j = 0; j < L; j++) X[j] = fma(rt, T.w[t][j], X[j]);
    }
    // X/M lies within 1/4 of the integer q, so the rounding of this estimate cannot change q.
    double qf = 0.0;
#pragma unroll
    for (int j = L - 1; j >= 0; j--) qf = fma(X[j], T.finv[j], qf);
    const double q = rint(qf);
    double D[L];
#pragma unroll
    for (int j = 0; j < L; j++) D[j] = fma(-q, T.mlimb[j], X[j]);
    const double base = 0x1p40;
#pragma unroll
    for (int j = 0; j < L - 1; j++) {
        const double c = rint(D[j] * 0x1p-40);
        D[j] = D[j] - c * base;
        D[j + 1] = D[j + 1] + c;
    }
    double h = D[L - 1], l = 0.0;
#pragma unroll
    for (int j = L - 2; j >= 0; j--) {
        const double hb = h * base;
        const double s = hb + D[j];
        const double bb = s - hb;
        const double e = (hb - (s - bb)) + (D[j] - bb);
        l = l * base + e;
        h = s;
    }
    return h + l;
}

// One thread per (32 x 32 block, lane): 16 outputs of column l%32.
template <int NM>
__global__ __launch_bounds__(256) void crt_combine(const v4i* __restrict__ Cr, int64_t c_plane, int cblocks_n,
                                                   int nblocks, int M, int N, const int* __restrict__ ea,
                                                   const int* __restrict__ eb, int pa, int pb, double alpha,
                                                   double beta, double* __restrict__ C, int64_t ldc, CrtTables T) {
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int blk = (int)(gid >> 6), lane = (int)(gid & 63);
    if (blk >= nblocks) return;
    const int rbc = blk / cblocks_n, cbc = blk % cblocks_n;
    const int col = cbc * 32 + (lane & 31);
    const int rbase = rbc * 32 + 4 * (lane >> 5);
    if (col >= N || rbase >= M) return;
    v4i vv[NM];
#pragma unroll
    for (int t = 0; t < NM; t++) vv[t] = Cr[t * c_plane + (int64_t)blk * 64 + lane];
    const int fj = eb[col];
#pragma unroll
    for (int b = 0; b < 16; b++) {
        const int row = rbase + 8 * (b >> 2) + (b & 3);
        if (row >= M) continue;
        int r[NM];
#pragma unroll
        for (int t = 0; t < NM; t++) r[t] = (signed char)((vv[t][b >> 2] >> (8 * (b & 3))) & 255);
        const int ei = ea[row];
        const double v = (ei == NONFINITE || fj == NONFINITE) ? __builtin_nan("")
                                                              : ldexp(crt_value<NM>(r, T), ei + fj - pa - pb);
        double* p = C + (int64_t)row * ldc + col;
        *p = (beta == 0.0) ? alpha * v : fma(alpha, v, beta * *p);
    }
}

// Little-endian base-2^32 integers of 8 words for the host tables.
struct Big {
    uint32_t w[8] = {};
    void mul(uint32_t v) {
        uint64_t carry = 0;
        for (int i = 0; i < 8; i++) {
            const uint64_t x = (uint64_t)w[i] * v + carry;
            w[i] = (uint32_t)x;
            carry = x >> 32;
        }
    }
    uint32_t mod(uint32_t m) const {
        uint64_t r = 0;
        for (int i = 7; i >= 0; i--) r = ((r << 32) | w[i]) % m;
        return (uint32_t)r;
    }
    double limb(int j) const {  // bits [40j, 40j+40)
        double v = 0.0;
        for (int b = 39; b >= 0; b--) {
            const int bit = 40 * j + b;
            v = v * 2.0 + (double)((w[bit >> 5] >> (bit & 31)) & 1u);
        }
        return v;
    }
    double value() const {
        double v = 0.0;
        for (int i = 7; i >= 0; i--) v = v * 4294967296.0 + (double)w[i];
        return v;
    }
};

}  // namespace

const int CRT_MODULI[CRT_MAX_MODULI] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                        223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

hipError_t launch_crt_stats_norms(int layout_a, int layout_b, const double* a, int64_t lda, const double* b,
                                  int64_t ldb, const SplitPlan& sp, int* ta, int* tb, float* sig_a, float* sig_b,
                                  hipStream_t stream) {
    const int K = (int)sp.k;
    NormPart* part = reinterpret_cast<NormPart*>(sp.stats_part);
    auto one = [&](bool kcontig, const double* x, int64_t s, int R, int* e, float* spread, int* t, float* sig) {
        // Rows long enough to fill a workgroup and short enough for its registers: one read.
        if (kcontig && K >= 2048 && K <= ROW_THREADS * ROW_REG) {
            hipLaunchKernelGGL(crt_row_stats_norms, dim3(R), dim3(ROW_THREADS), 0, stream, x, s, K, e, spread, t,
                               sig);
            return hipGetLastError();
        }
        hipError_t err = launch_operand_stats(kcontig, x, s, R, K, sp.stats_part, e, spread, stream);
        if (err != hipSuccess) return err;
        if (kcontig) {
            hipLaunchKernelGGL(crt_norm_rows, dim3((R + 3) / 4), dim3(256), 0, stream, x, s, R, K, e, spread, t,
                               sig);
            return hipGetLastError();
        }
        const int kparts = split_stats_kparts(R, K);
        const int krange = (K + kparts - 1) / kparts;
        hipLaunchKernelGGL(crt_norm_cols_partial, dim3((R + 63) / 64, kparts), dim3(256), 0, stream, x, s, R, K,
                           krange, e, part);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL(crt_norm_cols_finish, dim3((R + 255) / 256), dim3(256), 0, stream, part, R, kparts, e,
                           spread, t, sig);
        return hipGetLastError();
    };
    // Rows of op(A): ROW -> element (i,k) at a[i*lda + k]; COL -> a[k*lda + i].
    hipError_t err = one(layout_a == JLC_LAYOUT_ROW, a, lda, (int)sp.m, sp.ea, sp.spread_a, ta, sig_a);
    if (err != hipSuccess) return err;
    // Columns of op(B): COL -> element (k,j) at b[j*ldb + k]; ROW -> b[k*ldb + j].
    return one(layout_b == JLC_LAYOUT_COL, b, ldb, (int)sp.n, sp.eb, sp.spread_b, tb, sig_b);
}

double crt_log2_product(int nmod) {
    double s = 0.0;
    for (int t = 0; t < nmod && t < CRT_MAX_MODULI; t++) s += std::log2((double)CRT_MODULI[t]);
    return s;
}

void crt_tables(int nmod, CrtTables* T) {
    *T = CrtTables{};
    T->nmod = nmod;
    Big M;
    M.w[0] = 1;
    for (int t = 0; t < nmod; t++) M.mul((uint32_t)CRT_MODULI[t]);
    for (int t = 0; t < nmod; t++) {
        const int m = CRT_MODULI[t];
        T->mod[t] = m;
        T->inv[t] = 1.0 / m;
        T->hi_lim[t] = (double)((m - 1) / 2);
        T->lo_lim[t] = -(double)(m / 2);
        // 2^22 and 2^44 mod m in the symmetric range
        auto sym = [m](long long v) {
            int c = (int)(v % m);
            return c > (m - 1) / 2 ? c - m : c;
        };
        // digit weights 2^(22c + 8i), packed in the order of image_digits
        int kw[9];
        for (int c = 0; c < 3; c++)
            for (int i = 0; i < 3; i++) kw[3 * c + i] = sym(1LL << (22 * c + 8 * i));
        auto pk = [](int a, int b, int c, int d) {
            return (int)((uint32_t)(a & 255) | ((uint32_t)(b & 255) << 8) | ((uint32_t)(c & 255) << 16) |
                         ((uint32_t)(d & 255) << 24));
        };
        T->kdig[t][0] = pk(kw[0], kw[1], kw[2], kw[3]);
        T->kdig[t][1] = pk(kw[4], kw[5], kw[6], kw[7]);
        T->kdig[t][2] = pk(kw[8], 0, 0, 0);
        T->invf[t] = 1.0f / (float)m;
        Big Mi;
        Mi.w[0] = 1;
        for (int u = 0; u < nmod; u++)
            if (u != t) Mi.mul((uint32_t)CRT_MODULI[u]);
        const uint32_t r = Mi.mod((uint32_t)m);
        uint32_t y = 0;
        for (uint32_t cand = 1; cand < (uint32_t)m; cand++)
            if ((r * cand) % (uint32_t)m == 1) {
                y = cand;
                break;
            }
        Mi.mul(y);
        for (int j = 0; j < 4; j++) T->w[t][j] = Mi.limb(j);
    }
    const double md = M.value();
    for (int j = 0; j < 4; j++) {
        T->mlimb[j] = M.limb(j);
        T->finv[j] = std::ldexp(1.0, 40 * j) / md;
    }
}

void crt_plan_sizes(int64_t m, int64_t n, int64_t k, int nmod, CrtPlan* p) {
    p->m = m;
    p->n = n;
    p->k = k;
    p->nmod = nmod;
    p->mpad = (m + CT - 1) / CT * CT;
    p->npad = (n + CT - 1) / CT * CT;
    p->kpad = (k + KPAD - 1) / KPAD * KPAD;
    p->a_plane_bytes = p->mpad * p->kpad;
    p->b_plane_bytes = p->npad * p->kpad;
    p->c_plane_bytes = p->mpad * p->npad;
}

hipError_t launch_crt(int layout_a, int layout_b, const double* a, int64_t lda, const double* b, int64_t ldb,
                      const CrtPlan& p, double alpha, double beta, double* c, int64_t ldc, hipStream_t stream,
                      TileInfo* info) {
    const int M = (int)p.m, N = (int)p.n, K = (int)p.k;
    CrtTables T;
    crt_tables(p.nmod, &T);
    const int64_t asr = layout_a == JLC_LAYOUT_ROW ? lda : 1, ask = layout_a == JLC_LAYOUT_ROW ? 1 : lda;
    const int64_t bsr = layout_b == JLC_LAYOUT_COL ? ldb : 1, bsk = layout_b == JLC_LAYOUT_COL ? 1 : ldb;
    const int kchunks = (int)(p.kpad / 32);
    const int ablocks = (int)(p.mpad / 32) * kchunks, bblocks = (int)(p.npad / 32) * kchunks;
    v4i* ap = static_cast<v4i*>(p.a_pack);
    v4i* bp = static_cast<v4i*>(p.b_pack);
    v4i* cr = static_cast<v4i*>(p.c_res);
    auto pack = [&](const double* x, int64_t sr, int64_t sk, int R, int blocks, const int* e, int bits, v4i* o,
                    int64_t plane) {
        const bool vec = sk == 1 && K % 16 == 0 && sr % 2 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0;
        if (vec)
            hipLaunchKernelGGL(crt_pack<true>, dim3((blocks + 3) / 4), dim3(256), 0, stream, x, sr, sk, R, K,
                               kchunks, blocks, e, bits, T, o, plane);
        else
            hipLaunchKernelGGL(crt_pack<false>, dim3((blocks + 3) / 4), dim3(256), 0, stream, x, sr, sk, R, K,
                               kchunks, blocks, e, bits, T, o, plane);
        return hipGetLastError();
    };
    hipError_t err = pack(a, asr, ask, M, ablocks, p.ea, p.pa, ap, p.a_plane_bytes / 16);
    if (err != hipSuccess) return err;
    if ((err = pack(b, bsr, bsk, N, bblocks, p.eb, p.pb, bp, p.b_plane_bytes / 16)) != hipSuccess) return err;
    const int tiles_m = (int)(p.mpad / CT), tiles_n = (int)(p.npad / CT);
    const int cblocks_n = (int)(p.npad / 32);
    const TuneTable tt = tune_snapshot();
    const bool w4 = tt.crt_waves == 4;
    const size_t lds = (size_t)(w4 ? W4_NBUF : NBUF_LDS) * STAGE_V4 * sizeof(v4i);
    const int group = tt.crt_group > 0 ? tt.crt_group : 16, xcd = tt.crt_xcd;
    const void* kfn = w4 ? (const void*)crt_gemm_i8_w4 : (const void*)crt_gemm_i8;
    (void)hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    mark_main(info, true, stream);
    if (w4)
        hipLaunchKernelGGL(crt_gemm_i8_w4, dim3(tiles_m * tiles_n, p.nmod), dim3(256), lds, stream, kchunks, ap,
                           p.a_plane_bytes / 16, bp, p.b_plane_bytes / 16, cr, p.c_plane_bytes / 16, cblocks_n,
                           tiles_m, tiles_n, T);
    else
        hipLaunchKernelGGL(crt_gemm_i8, dim3(tiles_m * tiles_n, p.nmod), dim3(512), lds, stream, kchunks, ap,
                           p.a_plane_bytes / 16, bp, p.b_plane_bytes / 16, cr, p.c_plane_bytes / 16, cblocks_n,
                           tiles_m, tiles_n, group, xcd, T);
    mark_main(info, false, stream);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    const int cblocks = (int)(p.mpad / 32) * cblocks_n;
    const dim3 cgrid((int)(((int64_t)cblocks * 64 + 255) / 256));
    switch (p.nmod) {
#define JLC_CRT_CASE(NM_)                                                                                         \
    case NM_:                                                                                                     \
        hipLaunchKernelGGL(crt_combine<NM_>, cgrid, dim3(256), 0, stream, cr, p.c_plane_bytes / 16, cblocks_n,    \
                           cblocks, M, N, p.ea, p.eb, p.pa, p.pb, alpha, beta, c, ldc, T);                        \
        break;
        JLC_CRT_CASE(2) JLC_CRT_CASE(3) JLC_CRT_CASE(4) JLC_CRT_CASE(5) JLC_CRT_CASE(6) JLC_CRT_CASE(7)
        JLC_CRT_CASE(8) JLC_CRT_CASE(9) JLC_CRT_CASE(10) JLC_CRT_CASE(11) JLC_CRT_CASE(12) JLC_CRT_CASE(13)
        JLC_CRT_CASE(14) JLC_CRT_CASE(15) JLC_CRT_CASE(16) JLC_CRT_CASE(17) JLC_CRT_CASE(18) JLC_CRT_CASE(19)
        JLC_CRT_CASE(20)
#undef JLC_CRT_CASE
        default:
            return hipErrorInvalidValue;
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if (info) {
        info->tile_m = CT;
        info->tile_n = CT;
        info->tile_k = KSTAGE;
        info->slices = p.nmod;
        info->launches += 4;
    }
    return hipSuccess;
}

}  // namespace jlc
