// Split-precision DGEMM on INT8 MFMA (gfx950).
//
// C = alpha * op(A) * op(B) + beta * C computed from INT8 slices of A and B,
// an Ozaki-style error-free splitting (Ozaki, Ogita, Oishi, Rump 2012;
// Ootomo, Ozaki, Yokota 2024 for integer matrix units):
//
//  1. Scaling. Each row i of op(A) is scaled by 2^-ea[i] and each column j of
//     op(B) by 2^-eb[j], where 2^e is the power of two just above the row /
//     column max |x| (frexp exponent), so scaled entries lie in (-1, 1).
//     Power-of-two scaling is exact.
//  2. Slicing. A scaled entry v is cut into S signed 7-bit digits:
//     v <- v*128, q = trunc(v), v <- v - q, repeated S times. Every step is
//     exact in FP64, all digits share the sign of the entry and |q| <= 127, so
//     x = 2^e * (sum_t q_t 2^-7t + r) with |r| < 2^-7S.
//  3. Products. For slice indices t, u (0-based) with t + u <= S-1 (the
//     triangular truncation), the integer products A_t*B_u are accumulated per
//     diagonal d = t + u in INT32 on v_mfma_i32_32x32x32_i8. A diagonal holds
//     at most S products of K terms of magnitude <= 127^2, so K is processed in
//     chunks with S*chunk*127^2 < 2^31: every INT32 sum is exact.
//  4. Recombination in FP64: sum_d acc_d * 2^-7(d+2) from the smallest
//     diagonal up, scaled by 2^(ea[i]+eb[j]), then alpha / beta applied as
//     fma(alpha, v, beta*c) (first chunk) or fma(alpha, v, c) (later chunks).
//
// Slices are written in MFMA fragment order: one 1 KiB block per (slice,
// 32-row block, 32-deep k chunk) holds lane l's 16 bytes at l*16, lane l
// being row l%32, k in [16*(l/32), 16*(l/32)+16). The GEMM kernel therefore
// moves whole 1 KiB blocks (global -> registers -> LDS -> registers) with
// linear 16-byte accesses and no bank conflicts.
//
// GEMM kernel: 512 threads (8 wavefronts, 2 per SIMD), 128 x 64 C tile, each
// wavefront one 32 x 32 block with S INT32 accumulator sets (one per
// diagonal). Per 32-deep K step the whole stage (S slices of 4 A blocks and
// 2 B blocks) sits in LDS, so every loaded A_t fragment feeds S - t MFMAs:
// the pair products reuse operands instead of running S(S+1)/2 separate GEMMs.
//
// Non-finite inputs: a row or column containing Inf/NaN is marked and its
// outputs are NaN (the slices cannot carry them).
#include <math.h>

#include "internal.h"
#include "split_stats.h"

namespace jlc {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int NONFINITE = STAT_NONFINITE;
constexpr int TM = 128;             // C tile rows
constexpr int TN = 64;              // C tile cols
constexpr int KS = 32;              // K per MFMA step

// Row statistics of op(X) where element (r, k) is X[r*sr + k*sk]:
// e[r] = frexp exponent of max_k |x| (0 for an all-zero row, NONFINITE if any
// entry is Inf/NaN) and spread[r] = e[r] - mean_k log2|x| over the non-zero
// entries: the bits between the row's scale and its typical entry (geometric
// mean, so heavy-tailed rows show their full width), read by the slice-count
// and moduli-count choosers.
__global__ __launch_bounds__(256) void row_stats_kcontig(const double* __restrict__ X, int64_t sr, int R, int Kd,
                                                         int* __restrict__ e, float* __restrict__ spread) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= R) return;
    const double* row = X + (int64_t)r * sr;
    double mx = 0.0, slog = 0.0, cnt = 0.0;
    bool bad = false;
    for (int k = lane; k < Kd; k += 64) stat_add(fabs(row[k]), mx, slog, cnt, bad);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off, 64));
        slog += __shfl_xor(slog, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
        bad |= __shfl_xor((int)bad, off, 64) != 0;
    }
    if (lane == 0) {
        int ex = 0;
        if (mx > 0.0) frexp(mx, &ex);
        e[r] = bad ? NONFINITE : ex;
        spread[r] = (bad || cnt == 0.0) ? 0.0f : (float)(ex - slog / cnt);
    }
}

// Same statistics when consecutive rows r are contiguous (element (r, k) at
// X[k*sk + r]). Two passes so that a long K fills the chip: pass 1 gives each
// block 64 rows and one of kparts K ranges (wave w of the block takes every
// fourth k of the range), and writes that range's {max or -1 for
// non-finite, sum of log2|x|, non-zero count} to part[kp*R + r]; pass 2
// combines the ranges in order, so the result does not depend on scheduling.
struct StatPart {
    double mx, slog, cnt, pad;
};

__global__ __launch_bounds__(256) void col_stats_partial(const double* __restrict__ X, int64_t sk, int R, int Kd,
                                                         int krange, StatPart* __restrict__ part) {
    __shared__ StatPart red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    const int k0 = blockIdx.y * krange;
    const int k1 = min(Kd, k0 + krange);
    double mx = 0.0, slog = 0.0, cnt = 0.0;
    bool bad = false;
    if (r < R) {
        const double* col = X + r;
        int k = k0 + w;
        for (; k + 12 < k1; k += 16) {
            const double v0 = fabs(col[(int64_t)k * sk]), v1 = fabs(col[(int64_t)(k + 4) * sk]);
            const double v2 = fabs(col[(int64_t)(k + 8) * sk]), v3 = fabs(col[(int64_t)(k + 12) * sk]);
            stat_add(v0, mx, slog, cnt, bad);
            stat_add(v1, mx, slog, cnt, bad);
            stat_add(v2, mx, slog, cnt, bad);
            stat_add(v3, mx, slog, cnt, bad);
        }
        for (; k < k1; k += 4) stat_add(fabs(col[(int64_t)k * sk]), mx, slog, cnt, bad);
    }
    red[w][lane] = StatPart{bad ? -1.0 : mx, slog, cnt, 0.0};
    __syncthreads();
    if (w == 0 && r < R) {
        StatPart acc = red[0][lane];
#pragma unroll
        for (int i = 1; i < 4; i++) {
            const StatPart t = red[i][lane];
            acc.mx = (acc.mx < 0.0 || t.mx < 0.0) ? -1.0 : fmax(acc.mx, t.mx);
            acc.slog += t.slog;
            acc.cnt += t.cnt;
        }
        part[(int64_t)blockIdx.y * R + r] = acc;
    }
}

__global__ __launch_bounds__(256) void col_stats_finish(const StatPart* __restrict__ part, int R, int kparts,
                                                        int* __restrict__ e, float* __restrict__ spread) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    double mx = 0.0, slog = 0.0, cnt = 0.0;
    bool bad = false;
    for (int kp = 0; kp < kparts; kp++) {
        const StatPart t = part[(int64_t)kp * R + r];
        bad |= t.mx < 0.0;
        mx = fmax(mx, t.mx);
        slog += t.slog;
        cnt += t.cnt;
    }
    int ex = 0;
    if (mx > 0.0) frexp(mx, &ex);
    e[r] = bad ? NONFINITE : ex;
    spread[r] = (bad || cnt == 0.0) ? 0.0f : (float)(ex - slog / cnt);
}

// Writes the S slices of op(X) (element (r, k) at X[r*sr + k*sk], R x Kd,
// zero padded to Rpad x Kpad) in fragment order. One wavefront per 1 KiB block
// (32 rows x 32 k); lane l slices row l%32, k in [16*(l/32), +16).
template <int S>
__global__ __launch_bounds__(256) void pack_slices(const double* __restrict__ X, int64_t sr, int64_t sk, int R, int Kd,
                                                   int kchunks, int nblocks, const int* __restrict__ e,
                                                   v4i* __restrict__ out, int64_t slice_stride) {
    const int lane = threadIdx.x & 63;
    const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blk >= nblocks) return;
    const int rb = blk / kchunks, kc = blk % kchunks;
    const int r = rb * 32 + (lane & 31);
    const int k0 = kc * KS + 16 * (lane >> 5);
    unsigned char q[S][16];
    int ex = r < R ? e[r] : 0;
    const bool live = r < R && ex != NONFINITE;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int k = k0 + j;
        double v = 0.0;
        if (live && k < Kd) v = ldexp(X[(int64_t)r * sr + (int64_t)k * sk], -ex);
#pragma unroll
        for (int t = 0; t < S; t++) {
            v *= 128.0;
            const double d = trunc(v);
            v -= d;
            q[t][j] = (unsigned char)(signed char)(int)d;
        }
    }
    const int64_t off = (int64_t)blk * 64 + lane;
#pragma unroll
    for (int t = 0; t < S; t++) {
        v4i w;
        w.x = q[t][0] | (q[t][1] << 8) | (q[t][2] << 16) | (q[t][3] << 24);
        w.y = q[t][4] | (q[t][5] << 8) | (q[t][6] << 16) | (q[t][7] << 24);
        w.z = q[t][8] | (q[t][9] << 8) | (q[t][10] << 16) | (q[t][11] << 24);
        w.w = q[t][12] | (q[t][13] << 8) | (q[t][14] << 16) | (q[t][15] << 24);
        out[t * slice_stride + off] = w;
    }
}

// Tile order: ids regrouped per XCD, then column strips of 4 tile rows (with
// 128 x 64 tiles this keeps the 32 resident tiles of an XCD on 4 A and 8 B panels).
__device__ __forceinline__ void tile_coords(int bid, int nblocks, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int XCDS = 8;
    constexpr int GROUP = 4;
    int full = nblocks - nblocks % XCDS;
    int lid = bid;
    if (bid < full) lid = (bid % XCDS) * (full / XCDS) + bid / XCDS;
    int per_group = GROUP * tiles_n;
    int g = lid / per_group;
    int first_m = g * GROUP;
    int gsize = min(tiles_m - first_m, GROUP);
    int in = lid % per_group;
    tm = first_m + in % gsize;
    tn = in / gsize;
}

template <int S>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void split_gemm_i8(
    int M, int N, int ksteps, int chunk, const v4i* __restrict__ Ap, int64_t a_slice, const v4i* __restrict__ Bp,
    int64_t b_slice, const int* __restrict__ ea, const int* __restrict__ eb, double alpha, double beta,
    double* __restrict__ C, int64_t ldc, int tiles_m, int tiles_n) {
    extern __shared__ v4i lds[];
    constexpr int NA = 4 * S;            // A blocks per stage: slice t, row block j -> t*4 + j
    constexpr int NBLK = 6 * S;          // + B blocks: slice u, col block j -> NA + u*2 + j
    constexpr int PER = (NBLK + 7) / 8;  // blocks each wavefront moves per stage
    const int lane = threadIdx.x & 63;
    // Wave-uniform id: block addresses below live in scalar registers.
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w >> 1, wn = w & 1;
    int tm, tn;
    tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
    const int mb0 = tm * (TM / 32), nb0 = tn * (TN / 32);

    // Block moves are unconditional: a wavefront whose share runs past NBLK
    // repeats block NBLK-1, writing the same bytes to the same LDS slot as the
    // wavefront that owns it. Addresses are rebuilt from wave-uniform scalars at
    // each use, so no per-lane pointer array stays live across the K loop.
    auto blk_of = [&](int i) {
        const int b = w + 8 * i;
        return b < NBLK ? b : NBLK - 1;
    };
    auto src_of = [&](int b) {
        return b < NA ? Ap + (b / 4) * a_slice + (int64_t)(mb0 + (b % 4)) * ksteps * 64
                      : Bp + ((b - NA) / 2) * b_slice + (int64_t)(nb0 + ((b - NA) % 2)) * ksteps * 64;
    };
    v4i pre[PER];
    auto load = [&](int kc) {
#pragma unroll
        for (int i = 0; i < PER; i++) pre[i] = src_of(blk_of(i))[(int64_t)kc * 64 + lane];
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < PER; i++) lds[(buf * NBLK + blk_of(i)) * 64 + lane] = pre[i];
    };

    v16i acc[S];
#pragma unroll
    for (int d = 0; d < S; d++) acc[d] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

    // Accumulator element r of lane l holds C(8*(r/4) + 4*(l/32) + r%4, l%32) of its 32 x 32 block.
    const int col = tn * TN + wn * 32 + (lane & 31);
    const int rbase = tm * TM + wm * 32 + 4 * (lane >> 5);
    const int fj = col < N ? eb[col] : 0;
    auto flush = [&](bool first) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = rbase + 8 * (r >> 2) + (r & 3);
            if (row < M && col < N) {
                double sum = 0.0;
#pragma unroll
                for (int d = S - 1; d >= 0; d--) sum += (double)acc[d][r] * ldexp(1.0, -7 * (d + 2));
                const int ei = ea[row];
                const double v = (ei == NONFINITE || fj == NONFINITE) ? __builtin_nan("") : ldexp(sum, ei + fj);
                double* p = C + (int64_t)row * ldc + col;
                if (first) *p = (beta == 0.0) ? alpha * v : fma(alpha, v, beta * *p);
                else *p = fma(alpha, v, *p);
            }
        }
    };

    load(0);
    store(0);
    __syncthreads();
    // The K loop is branch-free inside a chunk (the prefetch index is clamped and
    // the LDS store is unconditional; the repeat of the last step lands in the
    // buffer nobody reads), so the compiler's wait counts keep the next step's
    // loads in flight across the MFMAs instead of draining them at a join.
    const int last = ksteps - 1;
    for (int c0 = 0; c0 < ksteps; c0 += chunk) {
        const int c1 = min(ksteps, c0 + chunk);
        for (int kc = c0; kc < c1; kc++) {
            const int cur = kc & 1;
            load(min(kc + 1, last));
            const v4i* st = lds + cur * NBLK * 64 + lane;
            v4i bf[S];
#pragma unroll
            for (int u = 0; u < S; u++) bf[u] = st[(NA + u * 2 + wn) * 64];
            // t outer, u descending: consecutive MFMAs always target different diagonals.
            // The next stage's LDS store sits after the first two slice rows of MFMAs.
#pragma unroll
            for (int t = 0; t < S; t++) {
                if (t == 2) store(cur ^ 1);
                const v4i af = st[(t * 4 + wm) * 64];
#pragma unroll
                for (int u = S - 1 - t; u >= 0; u--)
                    acc[t + u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[u], acc[t + u], 0, 0, 0);
            }
            __syncthreads();
        }
        flush(c0 == 0);
#pragma unroll
        for (int d = 0; d < S; d++) acc[d] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    }
}

template <int S>
hipError_t launch_pack(const double* X, int64_t sr, int64_t sk, int R, int Kd, int Rpad, int Kpad, const int* e,
                       void* out, int64_t slice_bytes, hipStream_t stream) {
    const int kchunks = Kpad / KS;
    const int nblocks = (Rpad / 32) * kchunks;
    hipLaunchKernelGGL(pack_slices<S>, dim3((nblocks + 3) / 4), dim3(256), 0, stream, X, sr, sk, R, Kd, kchunks,
                       nblocks, e, static_cast<v4i*>(out), slice_bytes / 16);
    return hipGetLastError();
}

template <int S>
hipError_t launch_core(const SplitPlan& p, double alpha, double beta, double* C, int64_t ldc, hipStream_t stream) {
    const int tiles_m = p.mpad / TM, tiles_n = p.npad / TN;
    const int ksteps = p.kpad / KS;
    // S * chunk * 32 * 127^2 < 2^31: every diagonal's INT32 sum is exact.
    int chunk = (int)(2147483647LL / ((long long)S * 32 * 127 * 127));
    if (chunk > ksteps) chunk = ksteps;
    const size_t lds = (size_t)2 * 6 * S * 1024;
    (void)hipFuncSetAttribute((const void*)split_gemm_i8<S>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(split_gemm_i8<S>, dim3(tiles_m * tiles_n), dim3(512), lds, stream, (int)p.m, (int)p.n, ksteps,
                       chunk, static_cast<const v4i*>(p.a_pack), p.a_slice_bytes / 16,
                       static_cast<const v4i*>(p.b_pack), p.b_slice_bytes / 16, p.ea, p.eb, alpha, beta, C, ldc,
                       tiles_m, tiles_n);
    return hipGetLastError();
}

}  // namespace

void split_plan_sizes(int64_t m, int64_t n, int64_t k, int slices, SplitPlan* p) {
    p->m = m;
    p->n = n;
    p->k = k;
    p->slices = slices;
    p->mpad = (m + TM - 1) / TM * TM;
    p->npad = (n + TN - 1) / TN * TN;
    p->kpad = (k + KS - 1) / KS * KS;
    p->a_slice_bytes = p->mpad * p->kpad;
    p->b_slice_bytes = p->npad * p->kpad;
}

int split_stats_kparts(int64_t rows, int64_t k) {
    const int64_t col_blocks = (rows + 63) / 64;
    int64_t kp = (2048 + col_blocks - 1) / col_blocks;       // aim for >= 2048 workgroups
    kp = std::min<int64_t>(kp, std::max<int64_t>(1, k / 256));  // at least 256 k per range
    return (int)std::max<int64_t>(1, std::min<int64_t>(kp, 256));
}

int64_t split_stats_part_elems(int64_t m, int64_t n, int64_t k) {
    return std::max(split_stats_kparts(m, k) * m, split_stats_kparts(n, k) * n);
}

namespace {
hipError_t launch_col_stats(const double* x, int64_t sk, int R, int K, double* part, int* e, float* spread,
                            hipStream_t stream) {
    const int kparts = split_stats_kparts(R, K);
    const int krange = (K + kparts - 1) / kparts;
    StatPart* pp = reinterpret_cast<StatPart*>(part);
    hipLaunchKernelGGL(col_stats_partial, dim3((R + 63) / 64, kparts), dim3(256), 0, stream, x, sk, R, K, krange,
                       pp);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(col_stats_finish, dim3((R + 255) / 256), dim3(256), 0, stream, pp, R, kparts, e, spread);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_operand_stats(bool kcontig, const double* x, int64_t stride, int R, int K, double* part, int* e,
                                float* spread, hipStream_t stream) {
    if (!kcontig) return launch_col_stats(x, stride, R, K, part, e, spread, stream);
    hipLaunchKernelGGL(row_stats_kcontig, dim3((R + 3) / 4), dim3(256), 0, stream, x, stride, R, K, e, spread);
    return hipGetLastError();
}

hipError_t launch_split_stats(int layout_a, int layout_b, const double* a, int64_t lda, const double* b, int64_t ldb,
                              SplitPlan* p, hipStream_t stream) {
    const int M = (int)p->m, N = (int)p->n, K = (int)p->k;
    // Rows of op(A): ROW -> element (i,k) at a[i*lda + k]; COL -> a[k*lda + i].
    if (layout_a == JLC_LAYOUT_ROW)
        hipLaunchKernelGGL(row_stats_kcontig, dim3((M + 3) / 4), dim3(256), 0, stream, a, lda, M, K, p->ea,
                           p->spread_a);
    else
    {
        hipError_t err = launch_col_stats(a, lda, M, K, p->stats_part, p->ea, p->spread_a, stream);
        if (err != hipSuccess) return err;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    // Columns of op(B): ROW -> element (k,j) at b[k*ldb + j]; COL -> b[j*ldb + k].
    if (layout_b == JLC_LAYOUT_COL)
        hipLaunchKernelGGL(row_stats_kcontig, dim3((N + 3) / 4), dim3(256), 0, stream, b, ldb, N, K, p->eb,
                           p->spread_b);
    else
        return launch_col_stats(b, ldb, N, K, p->stats_part, p->eb, p->spread_b, stream);
    return hipGetLastError();
}

hipError_t launch_split_pack_and_gemm(int layout_a, int layout_b, const double* a, int64_t lda, const double* b,
                                      int64_t ldb, const SplitPlan& p, double alpha, double beta, double* c,
                                      int64_t ldc, hipStream_t stream, TileInfo* info) {
    const int M = (int)p.m, N = (int)p.n, K = (int)p.k;
    const int64_t asr = layout_a == JLC_LAYOUT_ROW ? lda : 1, ask = layout_a == JLC_LAYOUT_ROW ? 1 : lda;
    const int64_t bsr = layout_b == JLC_LAYOUT_COL ? ldb : 1, bsk = layout_b == JLC_LAYOUT_COL ? 1 : ldb;
    hipError_t err = hipSuccess;
    switch (p.slices) {
#define JLC_SPLIT_CASE(S_)                                                                                      \
    case S_:                                                                                                    \
        err = launch_pack<S_>(a, asr, ask, M, K, (int)p.mpad, (int)p.kpad, p.ea, p.a_pack, p.a_slice_bytes,     \
                              stream);                                                                          \
        if (err == hipSuccess)                                                                                  \
            err = launch_pack<S_>(b, bsr, bsk, N, K, (int)p.npad, (int)p.kpad, p.eb, p.b_pack, p.b_slice_bytes, \
                                  stream);                                                                      \
        mark_main(info, true, stream);                                                                          \
        if (err == hipSuccess) err = launch_core<S_>(p, alpha, beta, c, ldc, stream);                           \
        mark_main(info, false, stream);                                                                         \
        break;
        JLC_SPLIT_CASE(3)
        JLC_SPLIT_CASE(4)
        JLC_SPLIT_CASE(5)
        JLC_SPLIT_CASE(6)
        JLC_SPLIT_CASE(7)
        JLC_SPLIT_CASE(8)
        JLC_SPLIT_CASE(9)
#undef JLC_SPLIT_CASE
        default:
            return hipErrorInvalidValue;
    }
    if (info) {
        info->tile_m = TM;
        info->tile_n = TN;
        info->tile_k = KS;
        info->launches += 3;
    }
    return err;
}

}  // namespace jlc
