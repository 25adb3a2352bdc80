// Grouped (batched, uneven) FP64 GEMM: many independent problems in one
// persistent launch.
//
// This serves the reference's TINY / SMALL regimes in bulk
// (SpecializedKernels.tinyGemm / smallGemm, src/main/java/net/faulj/compute/
// SpecializedKernels.java:79-158, and the GemmDispatch thresholds that route
// 512-flop and 50k-flop problems there, GemmDispatch.java:30-33, :76-106)
// and the many small trailing updates of batched QR factorisations.
//
// Work decomposition. Every problem is cut into 32 x 32 tiles of C, or 16-row
// (16-column) tiles where m (n) is at most a tuned bound, so that small
// problems do not pad their MFMA work to 32; the host builds one 64-bit entry
// per tile (problem index, tile shape, tile row, tile column), longest K first. A persistent grid of wavefronts deals the list round-robin
// (wavefront w takes entries w, w + W, w + 2W, ...): no atomics, no counters
// to clear, so the launch is safe to capture in a HIP graph and to replay. The
// host places the entries so that the tiles of one problem (in chunks of whole
// tile rows) fall to the wavefronts of one XCD, whose L2 then serves the A row
// blocks and B column blocks the tiles share; slots left over by that placement
// hold a skip entry.
//
// Tile kernel. One wavefront owns a tile as 1 x 1 to 2 x 2 v_mfma_f64_16x16x4
// blocks and reads its operands straight from global memory (L2): per 16-deep
// k block each lane fetches 4 consecutive k of its row of op(A) and column of
// op(B), lane group g taking k = k0 + 4g + j in MFMA step j (the same map for A
// and B, so the sums are the right ones). The next k block is fetched while
// the current one is multiplied. No LDS and no barriers inside a tile.
// Edges are masked; alpha == 0 or k == 0 reads neither A nor B, beta == 0
// never reads C (BLAS semantics).
#include "internal.h"

namespace jlc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
// Global-address-space views: the descriptor's pointers would otherwise compile
// to flat accesses, which also occupy the LDS counter and path.
typedef const __attribute__((address_space(1))) double* gcptr;
typedef __attribute__((address_space(1))) double* gptr;

constexpr int WAVES_PER_GROUP = GROUPED_WAVES_PER_GROUP;

// acc += op(A)[i0:+16TI, :] * op(B)[:, j0:+16TJ] for one tile. Every lane keeps
// one base pointer per 16-row block of op(A) and per 16-column block of op(B),
// clamped to the last row / column: rows past m and columns past n then read
// real entries, which only reach C rows and columns that are never stored, so
// they need no mask. k past K does need one (both operands read as zero, so
// that 0 * Inf cannot turn into NaN), and only the last k block has such k.
template <bool A_ROW, bool B_ROW, int TI, int TJ>
__device__ __forceinline__ void tile_product(const GroupedDesc& d, int i0, int j0, int r16, int g,
                                             d4 (&acc)[TI][TJ]) {
    const int K = d.k;
    const int64_t ask = A_ROW ? 1 : d.lda;  // k stride of op(A)
    const int64_t bsk = B_ROW ? d.ldb : 1;  // k stride of op(B)
    gcptr pa[TI], pb[TJ];
#pragma unroll
    for (int i = 0; i < TI; i++) {
        const int row = min(i0 + 16 * i + r16, d.m - 1);
        pa[i] = (gcptr)d.a + (A_ROW ? (int64_t)row * d.lda : (int64_t)row);
    }
#pragma unroll
    for (int j = 0; j < TJ; j++) {
        const int col = min(j0 + 16 * j + r16, d.n - 1);
        pb[j] = (gcptr)d.b + (B_ROW ? (int64_t)col : (int64_t)col * d.ldb);
    }
    // Lane group g takes k = k0 + 4g + s in MFMA step s, for A and B alike.
    double na[TI][4], nb[TJ][4];
    auto fetch_full = [&](int k0) {
        const int kb = k0 + 4 * g;
#pragma unroll
        for (int i = 0; i < TI; i++) {
            const gcptr qa = pa[i] + kb * ask;
#pragma unroll
            for (int s = 0; s < 4; s++) na[i][s] = qa[s * ask];
        }
#pragma unroll
        for (int j = 0; j < TJ; j++) {
            const gcptr qb = pb[j] + kb * bsk;
#pragma unroll
            for (int s = 0; s < 4; s++) nb[j][s] = qb[s * bsk];
        }
    };
    auto fetch_tail = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = k0 + 4 * g + s;
            const int kc = min(k, K - 1);
#pragma unroll
            for (int i = 0; i < TI; i++) {
                const double va = pa[i][kc * ask];
                na[i][s] = k < K ? va : 0.0;
            }
#pragma unroll
            for (int j = 0; j < TJ; j++) {
                const double vb = pb[j][kc * bsk];
                nb[j][s] = k < K ? vb : 0.0;
            }
        }
    };
    auto mma = [&](const double (&xa)[TI][4], const double (&xb)[TJ][4]) {
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int i = 0; i < TI; i++)
#pragma unroll
                for (int j = 0; j < TJ; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[i][s], xb[j][s], acc[i][j], 0, 0, 0);
    };
    // Full 16-deep blocks in a straight loop (each iteration fetches the next while multiplying the
    // current one), then the masked tail block: blocks are summed in k order either way.
    const int nfull = K >> 4;
    if (nfull == 0) {
        fetch_tail(0);
        mma(na, nb);
        return;
    }
    double ca[TI][4], cb[TJ][4];
    auto keep = [&]() {
#pragma unroll
        for (int s = 0; s < 4; s++) {
#pragma unroll
            for (int i = 0; i < TI; i++) ca[i][s] = na[i][s];
#pragma unroll
            for (int j = 0; j < TJ; j++) cb[j][s] = nb[j][s];
        }
    };
    fetch_full(0);
    for (int b = 1; b < nfull; b++) {
        keep();
        fetch_full(16 * b);
        mma(ca, cb);
    }
    if (K & 15) {
        keep();
        fetch_tail(16 * nfull);
        mma(ca, cb);
    }
    mma(na, nb);
}

// One whole tile of 16 TI x 16 TJ entries: C read (beta != 0) ahead of the product, the product (PRODUCT:
// alpha != 0 and K > 0), the store. Each layout path holds its own accumulators from start to store.
template <bool PRODUCT, bool A_ROW, bool B_ROW, int TI, int TJ>
__device__ __forceinline__ void tile(const GroupedDesc& d, int i0, int j0, int r16, int g) {
    const gptr c = (gptr)d.c;
    double cold[TI][TJ][4];
    if (d.beta != 0.0) {
#pragma unroll
        for (int i = 0; i < TI; i++)
#pragma unroll
            for (int j = 0; j < TJ; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = i0 + 16 * i + g + 4 * r, col = j0 + 16 * j + r16;
                    cold[i][j][r] = (row < d.m && col < d.n) ? c[(int64_t)row * d.ldc + col] : 0.0;
                }
    }
    d4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    if (PRODUCT) tile_product<A_ROW, B_ROW, TI, TJ>(d, i0, j0, r16, g, acc);
    // Accumulator element r of lane l holds C(l/16 + 4r, l%16) of its 16 x 16 block.
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++) {
            const int col = j0 + 16 * j + r16;
            if (col >= d.n) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = i0 + 16 * i + g + 4 * r;
                if (row < d.m) {
                    const double v = d.alpha * acc[i][j][r];
                    c[(int64_t)row * d.ldc + col] = d.beta == 0.0 ? v : v + d.beta * cold[i][j][r];
                }
            }
        }
}

// A tile of TI x TJ 16 x 16 blocks with the layout / alpha / k variant of its problem.
template <int TI, int TJ>
__device__ __forceinline__ void run_tile(const GroupedDesc& d, int i0, int j0, int r16, int g) {
    const bool a_row = d.layout_a == JLC_LAYOUT_ROW, b_row = d.layout_b == JLC_LAYOUT_ROW;
    if (d.alpha == 0.0 || d.k <= 0) tile<false, true, true, TI, TJ>(d, i0, j0, r16, g);
    else if (a_row && b_row) tile<true, true, true, TI, TJ>(d, i0, j0, r16, g);
    else if (a_row) tile<true, true, false, TI, TJ>(d, i0, j0, r16, g);
    else if (b_row) tile<true, false, true, TI, TJ>(d, i0, j0, r16, g);
    else tile<true, false, false, TI, TJ>(d, i0, j0, r16, g);
}

__global__ __launch_bounds__(64 * WAVES_PER_GROUP) __attribute__((amdgpu_waves_per_eu(2, 2))) void grouped_f64(
    const GroupedDesc* __restrict__ descs,
                                                                    const uint64_t* __restrict__ tiles, int ntiles) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES_PER_GROUP + (threadIdx.x >> 6));
    const int nwaves = gridDim.x * WAVES_PER_GROUP;
    const int r16 = lane & 15;
    const int g = lane >> 4;
    // The list entry two rounds ahead and the descriptor one round ahead are loaded while the current
    // tile runs, and C (when beta != 0) is read before the K loop instead of after it, so a tile starts
    // without a serial load. All wavefronts run the same number of rounds, with a workgroup barrier at
    // the start of each: the four wavefronts of a workgroup, which the host gives neighbouring tiles of
    // one problem where it can, then read their shared A and B blocks at about the same time instead of
    // drifting apart.
    uint64_t e = wave < ntiles ? tiles[wave] : GROUPED_SKIP;
    uint64_t e_next = wave + nwaves < ntiles ? tiles[wave + nwaves] : GROUPED_SKIP;
    GroupedDesc d = descs[e == GROUPED_SKIP ? 0 : __builtin_amdgcn_readfirstlane((int)(e >> 32))];
    const int rounds = (ntiles + nwaves - 1) / nwaves;
    for (int round = 0; round < rounds; round++) {
        const int t = wave + round * nwaves;
        __syncthreads();
        const GroupedDesc d_next = descs[e_next == GROUPED_SKIP ? 0 : __builtin_amdgcn_readfirstlane((int)(e_next >> 32))];
        const uint64_t e_next2 = t + 2 * nwaves < ntiles ? tiles[t + 2 * nwaves] : GROUPED_SKIP;
        if (e != GROUPED_SKIP) {
            const int shape = (int)((e >> GROUPED_SHAPE_SHIFT) & 3);
            const int tm = (int)((e >> 15) & 0x7fff), tn = (int)(e & 0x7fff);
            const int i0 = tm * (shape & 2 ? 16 : 32), j0 = tn * (shape & 1 ? 16 : 32);
            switch (shape) {
                case 0: run_tile<2, 2>(d, i0, j0, r16, g); break;
                case 1: run_tile<2, 1>(d, i0, j0, r16, g); break;
                case 2: run_tile<1, 2>(d, i0, j0, r16, g); break;
                default: run_tile<1, 1>(d, i0, j0, r16, g); break;
            }
        }
        e = e_next;
        e_next = e_next2;
        d = d_next;
    }
}

// ---------------------------------------------------------------------------
// Split-INT8 form of the grouped GEMM: every problem computed as the
// single-GEMM INT8 slice form computes it (split_int8.hip; the same exponents,
// digits, diagonal sums, K chunks, recombination and alpha / beta steps), one
// wavefront per 32 x 32 tile and no workspace, so a plan stays one graph node.
//  Pass 1 reads the tile's 32 rows of op(A) and 32 columns of op(B) over all K
//  for their exponents (frexp of max |x|; Inf / NaN marks the row or column).
//  Pass 2 slices each 32-deep block on the fly: lane l holds row / column l%32
//  and k in [16 (l/32), +16) of the block, the MFMA fragment order of
//  v_mfma_i32_32x32x32_i8, each value cut into S signed 7-bit digits packed 16
//  to four registers; the S(S+1)/2 slice products accumulate per diagonal.
constexpr int SPLIT_NONFINITE = -100000;

template <int S>
__device__ __forceinline__ void slice16(gcptr p, int64_t ks, int k0, int K, int e, v4i (&out)[S]) {
    unsigned q[S][4];
#pragma unroll
    for (int t = 0; t < S; t++) q[t][0] = q[t][1] = q[t][2] = q[t][3] = 0u;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int k = k0 + j;
        const double x = p[(int64_t)min(k, K - 1) * ks];
        double v = (k < K && e != SPLIT_NONFINITE) ? ldexp(x, -e) : 0.0;
#pragma unroll
        for (int t = 0; t < S; t++) {
            v *= 128.0;
            const double dg = trunc(v);
            v -= dg;
            q[t][j >> 2] |= ((unsigned)(int)dg & 255u) << (8 * (j & 3));
        }
    }
#pragma unroll
    for (int t = 0; t < S; t++) out[t] = v4i{(int)q[t][0], (int)q[t][1], (int)q[t][2], (int)q[t][3]};
}

// Frexp exponent of max |x| over the lane's 16-value share of every 32-deep block, combined with the
// other half-wave's share: the exponent of the whole row / column, or SPLIT_NONFINITE.
__device__ __forceinline__ int split_exponent(gcptr p, int64_t ks, int K, int h) {
    double mx = 0.0;
    int bad = 0;
    for (int k0 = 16 * h; k0 < K; k0 += 32) {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int k = k0 + j;
            const double v = fabs(p[(int64_t)min(k, K - 1) * ks]);
            if (k < K) {
                bad |= !(v <= 1.7976931348623157e308);
                mx = fmax(mx, v);
            }
        }
    }
    mx = fmax(mx, __shfl_xor(mx, 32, 64));
    bad |= __shfl_xor(bad, 32, 64);
    int e = 0;
    if (mx > 0.0) frexp(mx, &e);
    return bad ? SPLIT_NONFINITE : e;
}

template <int S>
__device__ __forceinline__ void split_tile(const GroupedDesc& d, int i0, int j0, int lane) {
    const int K = d.k;
    const int r32 = lane & 31, h = lane >> 5;
    const bool a_row = d.layout_a == JLC_LAYOUT_ROW, b_row = d.layout_b == JLC_LAYOUT_ROW;
    // Rows past m and columns past n read the last real one (their outputs are never stored).
    const int arow = min(i0 + r32, d.m - 1), bcol = min(j0 + r32, d.n - 1);
    const gcptr pa = (gcptr)d.a + (a_row ? (int64_t)arow * d.lda : (int64_t)arow);
    const gcptr pb = (gcptr)d.b + (b_row ? (int64_t)bcol : (int64_t)bcol * d.ldb);
    const int64_t ask = a_row ? 1 : d.lda, bsk = b_row ? d.ldb : 1;
    const int ea = split_exponent(pa, ask, K, h);
    const int eb = split_exponent(pb, bsk, K, h);
    // INT32 diagonal sums stay exact for `chunk` 32-deep blocks: S * chunk * 32 * 127^2 < 2^31.
    const int chunk = (int)(2147483647LL / ((long long)S * 32 * 127 * 127));
    const int kblocks = (K + 31) / 32;
    const gptr c = (gptr)d.c;
    const int col = j0 + r32;
    for (int c0 = 0; c0 < kblocks; c0 += chunk) {
        v16i acc[S];
#pragma unroll
        for (int dd = 0; dd < S; dd++) acc[dd] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int c1 = min(kblocks, c0 + chunk);
        for (int kb = c0; kb < c1; kb++) {
            v4i af[S], bf[S];
            slice16<S>(pa, ask, 32 * kb + 16 * h, K, ea, af);
            slice16<S>(pb, bsk, 32 * kb + 16 * h, K, eb, bf);
#pragma unroll
            for (int t = 0; t < S; t++)
#pragma unroll
                for (int u = S - 1 - t; u >= 0; u--)
                    acc[t + u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[t], bf[u], acc[t + u], 0, 0, 0);
        }
        // Accumulator element r of lane l holds C(8*(r/4) + 4*(l/32) + r%4, l%32) of the tile.
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int ri = 8 * (r >> 2) + 4 * h + (r & 3);
            const int ei = __shfl(ea, ri, 64);  // lane ri holds row i0 + ri's exponent
            const int row = i0 + ri;
            if (row < d.m && col < d.n) {
                double sum = 0.0;
#pragma unroll
                for (int dd = S - 1; dd >= 0; dd--) sum += (double)acc[dd][r] * ldexp(1.0, -7 * (dd + 2));
                const double v = (ei == SPLIT_NONFINITE || eb == SPLIT_NONFINITE) ? __builtin_nan("")
                                                                                  : ldexp(sum, ei + eb);
                const gptr pc = c + (int64_t)row * d.ldc + col;
                if (c0 == 0) *pc = (d.beta == 0.0) ? d.alpha * v : fma(d.alpha, v, d.beta * *pc);
                else *pc = fma(d.alpha, v, *pc);
            }
        }
    }
}

// C = beta * C for a tile whose problem has alpha == 0 or K == 0 (op(A), op(B) unread), as the
// single-GEMM entry points scale C without reaching the split kernels.
__device__ __forceinline__ void scale_tile(const GroupedDesc& d, int i0, int j0, int lane) {
    const gptr c = (gptr)d.c;
    const int col = j0 + (lane & 31);
    if (col >= d.n) return;
    for (int r = lane >> 5; r < 32; r += 2) {
        const int row = i0 + r;
        if (row >= d.m) break;
        const gptr pc = c + (int64_t)row * d.ldc + col;
        *pc = d.beta == 0.0 ? 0.0 : d.beta * *pc;
    }
}

template <int S>
__global__ __launch_bounds__(64 * WAVES_PER_GROUP) void grouped_split_i8(const GroupedDesc* __restrict__ descs,
                                                                         const uint64_t* __restrict__ tiles,
                                                                         int ntiles) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES_PER_GROUP + (threadIdx.x >> 6));
    const int nwaves = gridDim.x * WAVES_PER_GROUP;
    const int rounds = (ntiles + nwaves - 1) / nwaves;
    for (int round = 0; round < rounds; round++) {
        const int t = wave + round * nwaves;
        __syncthreads();
        const uint64_t e = t < ntiles ? tiles[t] : GROUPED_SKIP;
        if (e == GROUPED_SKIP) continue;
        const int p = __builtin_amdgcn_readfirstlane((int)(e >> 32));
        const int tm = (int)((e >> 15) & 0x7fff), tn = (int)(e & 0x7fff);
        const GroupedDesc d = descs[p];
        const int i0 = tm * 32, j0 = tn * 32;
        if (d.alpha == 0.0 || d.k <= 0) scale_tile(d, i0, j0, lane);
        else split_tile<S>(d, i0, j0, lane);
    }
}

template <int S>
int split_grid(int device_cus, int waves_per_simd) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_split_i8<S>, 64 * WAVES_PER_GROUP, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (waves_per_simd > 0 && waves_per_simd < per_cu) per_cu = waves_per_simd;
    return device_cus * per_cu;
}

}  // namespace

int grouped_grid(int device_cus, int waves_per_simd) {
    // Exactly the resident capacity: with static round-robin dealing, a
    // workgroup that only starts after another finishes would run its whole
    // share late.
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_f64, 64 * WAVES_PER_GROUP, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 2;
    // A tuned width below the occupancy limit (waves per SIMD == workgroups per CU here).
    if (waves_per_simd > 0 && waves_per_simd < per_cu) per_cu = waves_per_simd;
    return device_cus * per_cu;
}

int grouped_split_grid(int slices, int device_cus, int waves_per_simd) {
    switch (slices) {
        case 3: return split_grid<3>(device_cus, waves_per_simd);
        case 4: return split_grid<4>(device_cus, waves_per_simd);
        case 5: return split_grid<5>(device_cus, waves_per_simd);
        case 6: return split_grid<6>(device_cus, waves_per_simd);
        case 7: return split_grid<7>(device_cus, waves_per_simd);
        case 8: return split_grid<8>(device_cus, waves_per_simd);
        case 9: return split_grid<9>(device_cus, waves_per_simd);
        default: return 0;
    }
}

hipError_t launch_grouped_split_i8(int slices, const GroupedDesc* descs, const uint64_t* tiles, int ntiles, int grid,
                                   hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    const dim3 g(grid), b(64 * WAVES_PER_GROUP);
    switch (slices) {
        case 3: hipLaunchKernelGGL(grouped_split_i8<3>, g, b, 0, stream, descs, tiles, ntiles); break;
        case 4: hipLaunchKernelGGL(grouped_split_i8<4>, g, b, 0, stream, descs, tiles, ntiles); break;
        case 5: hipLaunchKernelGGL(grouped_split_i8<5>, g, b, 0, stream, descs, tiles, ntiles); break;
        case 6: hipLaunchKernelGGL(grouped_split_i8<6>, g, b, 0, stream, descs, tiles, ntiles); break;
        case 7: hipLaunchKernelGGL(grouped_split_i8<7>, g, b, 0, stream, descs, tiles, ntiles); break;
        case 8: hipLaunchKernelGGL(grouped_split_i8<8>, g, b, 0, stream, descs, tiles, ntiles); break;
        case 9: hipLaunchKernelGGL(grouped_split_i8<9>, g, b, 0, stream, descs, tiles, ntiles); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_grouped_f64(const GroupedDesc* descs, const uint64_t* tiles, int ntiles, int grid,
                              hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(grouped_f64, dim3(grid), dim3(64 * WAVES_PER_GROUP), 0, stream, descs, tiles, ntiles);
    return hipGetLastError();
}

}  // namespace jlc
