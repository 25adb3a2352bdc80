// Split-precision DGEMM on BF16 MFMA (gfx950): the BF16-slice variant of the
// split form in split_int8.hip.
//
// Same scaling (row / column powers of two from the exponent pass) but 8-bit
// digits: v <- v*256, q = trunc(v), v <- v - q; |q| <= 255 is exactly
// representable in BF16 (8 significant bits). Slice products run on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation, which is exact while every
// partial sum stays an integer of magnitude <= 2^24: a diagonal of at most S
// pairs accumulates KC K-values with S*KC*255^2 <= 2^24, then its FP32 sums are
// converted (exactly) into INT32 diagonal accumulators and cleared. The INT32
// accumulators are flushed into FP64 C every `chunk` K steps with
// S*chunk*16*255^2 < 2^31 and recombined as sum_d acc_d 2^-8(d+2) from the
// smallest diagonal up, exactly as the INT8 form (so the oracle restatement
// with 8-bit digits reproduces the result bit for bit).
//
// Fragment blocks: 1 KiB per (slice, 32-row block, 16-deep k chunk); lane l
// holds row l%32, k in [8*(l/32), +8) as 8 BF16 values. Kernel: 256 threads
// (4 wavefronts in 2 x 2, each one 32 x 32 block), 64 x 64 C tile, one
// wavefront per SIMD (the FP32 + INT32 accumulator sets need the registers).
#include <math.h>

#include "internal.h"
#include "split_stats.h"

namespace jlc {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int NONFINITE = STAT_NONFINITE;  // exponent marker of rows / columns holding Inf or NaN
constexpr int TM = 64, TN = 64, KS = 16;

__device__ __forceinline__ unsigned short bf16_of_int(int q) {
    // |q| <= 255: the float is exact and its low 16 bits are zero.
    return (unsigned short)(__float_as_uint((float)q) >> 16);
}

template <int S>
__global__ __launch_bounds__(256) void pack_slices_bf16(const double* __restrict__ X, int64_t sr, int64_t sk, int R,
                                                        int Kd, int kchunks, int nblocks, const int* __restrict__ e,
                                                        v4i* __restrict__ out, int64_t slice_stride) {
    const int lane = threadIdx.x & 63;
    const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blk >= nblocks) return;
    const int rb = blk / kchunks, kc = blk % kchunks;
    const int r = rb * 32 + (lane & 31);
    const int k0 = kc * KS + 8 * (lane >> 5);
    unsigned short q[S][8];
    const int ex = r < R ? e[r] : 0;
    const bool live = r < R && ex != NONFINITE;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int k = k0 + j;
        double v = 0.0;
        if (live && k < Kd) v = ldexp(X[(int64_t)r * sr + (int64_t)k * sk], -ex);
#pragma unroll
        for (int t = 0; t < S; t++) {
            v *= 256.0;
            const double d = trunc(v);
            v -= d;
            q[t][j] = bf16_of_int((int)d);
        }
    }
    const int64_t off = (int64_t)blk * 64 + lane;
#pragma unroll
    for (int t = 0; t < S; t++) {
        v4i w;
        w.x = q[t][0] | (q[t][1] << 16);
        w.y = q[t][2] | (q[t][3] << 16);
        w.z = q[t][4] | (q[t][5] << 16);
        w.w = q[t][6] | (q[t][7] << 16);
        out[t * slice_stride + off] = w;
    }
}

__device__ __forceinline__ void tile_coords(int bid, int nblocks, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int XCDS = 8, GROUP = 8;
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
__global__ __launch_bounds__(256) void split_gemm_bf16(int M, int N, int ksteps, int fsteps, int chunk,
                                                       const v4i* __restrict__ Ap, int64_t a_slice,
                                                       const v4i* __restrict__ Bp, int64_t b_slice,
                                                       const int* __restrict__ ea, const int* __restrict__ eb,
                                                       double alpha, double beta, double* __restrict__ C,
                                                       int64_t ldc, int tiles_m, int tiles_n) {
    extern __shared__ v4i lds[];
    constexpr int NA = 2 * S;     // A blocks per stage: slice t, row block j -> t*2 + j
    constexpr int NBLK = 4 * S;   // + B blocks: NA + u*2 + j
    constexpr int PER = (NBLK + 3) / 4;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w >> 1, wn = w & 1;
    int tm, tn;
    tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
    const int mb0 = tm * 2, nb0 = tn * 2;
    // Unconditional block moves (a share past NBLK repeats block NBLK-1, same
    // bytes to the same slot) with addresses rebuilt from wave-uniform scalars.
    auto blk_of = [&](int i) {
        const int b = w + 4 * i;
        return b < NBLK ? b : NBLK - 1;
    };
    auto src_of = [&](int b) {
        return b < NA ? Ap + (b / 2) * a_slice + (int64_t)(mb0 + (b % 2)) * ksteps * 64
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
    v16f accf[S];
    v16i acci[S];
#pragma unroll
    for (int d = 0; d < S; d++) {
        accf[d] = v16f{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        acci[d] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    }
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
                for (int d = S - 1; d >= 0; d--) sum += (double)acci[d][r] * ldexp(1.0, -8 * (d + 2));
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
    // Branch-free K steps inside each exact-FP32 run (clamped prefetch,
    // unconditional LDS store of the next step), so the wait counts keep the
    // next step's loads in flight; chunk is a multiple of fsteps.
    const int last = ksteps - 1;
    for (int c0 = 0; c0 < ksteps; c0 += chunk) {
        const int c1 = min(ksteps, c0 + chunk);
        for (int f0 = c0; f0 < c1; f0 += fsteps) {
            const int f1 = min(c1, f0 + fsteps);
            for (int kc = f0; kc < f1; kc++) {
                const int cur = kc & 1;
                load(min(kc + 1, last));
                const v4i* st = lds + cur * NBLK * 64 + lane;
                bf8 bf[S];
#pragma unroll
                for (int u = 0; u < S; u++) bf[u] = __builtin_bit_cast(bf8, st[(NA + u * 2 + wn) * 64]);
#pragma unroll
                for (int t = 0; t < S; t++) {
                    const bf8 af = __builtin_bit_cast(bf8, st[(t * 2 + wm) * 64]);
#pragma unroll
                    for (int u = S - 1 - t; u >= 0; u--)
                        accf[t + u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, bf[u], accf[t + u], 0, 0, 0);
                }
                store(cur ^ 1);
                __syncthreads();
            }
            // FP32 diagonal sums are exact integers here: move them into INT32.
#pragma unroll
            for (int d = 0; d < S; d++) {
#pragma unroll
                for (int r = 0; r < 16; r++) acci[d][r] += (int)accf[d][r];
                accf[d] = v16f{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            }
        }
        flush(c0 == 0);
#pragma unroll
        for (int d = 0; d < S; d++) acci[d] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    }
}

template <int S>
hipError_t run_bf16(int layout_a, int layout_b, const double* a, int64_t lda, const double* b, int64_t ldb,
                    const SplitPlan& p, double alpha, double beta, double* c, int64_t ldc, hipStream_t stream,
                    const TileInfo* info) {
    const int M = (int)p.m, N = (int)p.n, K = (int)p.k;
    const int64_t asr = layout_a == JLC_LAYOUT_ROW ? lda : 1, ask = layout_a == JLC_LAYOUT_ROW ? 1 : lda;
    const int64_t bsr = layout_b == JLC_LAYOUT_COL ? ldb : 1, bsk = layout_b == JLC_LAYOUT_COL ? 1 : ldb;
    const int kchunks = (int)(p.kpad / KS);
    int nblocks = (int)(p.mpad / 32) * kchunks;
    hipLaunchKernelGGL(pack_slices_bf16<S>, dim3((nblocks + 3) / 4), dim3(256), 0, stream, a, asr, ask, M, K, kchunks,
                       nblocks, p.ea, static_cast<v4i*>(p.a_pack), p.a_slice_bytes / 16);
    nblocks = (int)(p.npad / 32) * kchunks;
    hipLaunchKernelGGL(pack_slices_bf16<S>, dim3((nblocks + 3) / 4), dim3(256), 0, stream, b, bsr, bsk, N, K, kchunks,
                       nblocks, p.eb, static_cast<v4i*>(p.b_pack), p.b_slice_bytes / 16);
    const int ksteps = kchunks;
    // FP32 exactness: S * KC * 255^2 <= 2^24 with KC = fsteps * 16.
    int fsteps = (int)((1LL << 24) / ((long long)S * 255 * 255 * KS));
    if (fsteps < 1) fsteps = 1;
    // INT32 exactness: S * chunk * 16 * 255^2 < 2^31.
    int chunk = (int)(2147483647LL / ((long long)S * KS * 255 * 255));
    chunk = chunk / fsteps * fsteps;
    if (chunk > ksteps) chunk = ksteps;
    const int tiles_m = (int)(p.mpad / TM), tiles_n = (int)(p.npad / TN);
    const size_t lds = (size_t)2 * 4 * S * 1024;
    (void)hipFuncSetAttribute((const void*)split_gemm_bf16<S>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    mark_main(info, true, stream);
    hipLaunchKernelGGL(split_gemm_bf16<S>, dim3(tiles_m * tiles_n), dim3(256), lds, stream, M, N, ksteps, fsteps,
                       chunk, static_cast<const v4i*>(p.a_pack), p.a_slice_bytes / 16,
                       static_cast<const v4i*>(p.b_pack), p.b_slice_bytes / 16, p.ea, p.eb, alpha, beta, c, ldc,
                       tiles_m, tiles_n);
    mark_main(info, false, stream);
    return hipGetLastError();
}

}  // namespace

void split_bf16_plan_sizes(int64_t m, int64_t n, int64_t k, int slices, SplitPlan* p) {
    split_plan_sizes(m, n, k, slices, p);  // 128 x 64 padding covers the 64 x 64 tile
    p->kpad = (k + KS - 1) / KS * KS;
    p->a_slice_bytes = p->mpad * p->kpad * 2;
    p->b_slice_bytes = p->npad * p->kpad * 2;
}

int64_t split_bf16_chunk_k(int slices, int64_t k) {
    int fsteps = (int)((1LL << 24) / ((long long)slices * 255 * 255 * KS));
    if (fsteps < 1) fsteps = 1;
    int64_t chunk = 2147483647LL / ((long long)slices * KS * 255 * 255);
    chunk = chunk / fsteps * fsteps;
    const int64_t ksteps = (k + KS - 1) / KS;
    if (chunk > ksteps) chunk = ksteps;
    return chunk * KS;
}

hipError_t launch_split_bf16(int layout_a, int layout_b, const double* a, int64_t lda, const double* b, int64_t ldb,
                             const SplitPlan& p, double alpha, double beta, double* c, int64_t ldc,
                             hipStream_t stream, TileInfo* info) {
    hipError_t err;
    switch (p.slices) {
        case 3: err = run_bf16<3>(layout_a, layout_b, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info); break;
        case 4: err = run_bf16<4>(layout_a, layout_b, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info); break;
        case 5: err = run_bf16<5>(layout_a, layout_b, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info); break;
        case 6: err = run_bf16<6>(layout_a, layout_b, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info); break;
        case 7: err = run_bf16<7>(layout_a, layout_b, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info); break;
        case 8: err = run_bf16<8>(layout_a, layout_b, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info); break;
        default: return hipErrorInvalidValue;
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
