// Grouped (batched, uneven) FP64 GEMM: many independent problems in one
// persistent launch.
//
// This serves the reference's TINY / SMALL regimes in bulk
// (SpecializedKernels.tinyGemm / smallGemm, src/main/java/net/faulj/compute/
// SpecializedKernels.java:79-158, and the GemmDispatch thresholds that route
// 512-flop and 50k-flop problems there, GemmDispatch.java:30-33, :76-106)
// and the many small trailing updates of batched QR factorisations.
//
// Work decomposition. Every problem is cut into 32 x 32 tiles of C; the host
// builds one 64-bit entry per tile (problem index, tile row, tile column),
// longest K first. A persistent grid of wavefronts deals the list round-robin
// (wavefront w takes entries w, w + W, w + 2W, ...): no atomics, no counters
// to clear, so the launch is safe to capture in a HIP graph and to replay. The
// host places the entries so that the tiles of one problem (in chunks of whole
// tile rows) fall to the wavefronts of one XCD, whose L2 then serves the A row
// blocks and B column blocks the tiles share; slots left over by that placement
// hold a skip entry.
//
// Tile kernel. One wavefront owns a 32 x 32 tile as 2 x 2 v_mfma_f64_16x16x4
// blocks and reads its operands straight from global memory (L2): per 16-deep
// k block each lane fetches 4 consecutive k of its row of op(A) and column of
// op(B), lane group g taking k = k0 + 4g + j in MFMA step j (the same map for A
// and B, so the sums are the right ones). The next k block is fetched while
// the current one is multiplied. No LDS and no barriers: small tiles are
// latency-bound and the kernel relies on many resident wavefronts instead.
// Edges are masked; alpha == 0 or k == 0 reads neither A nor B, beta == 0
// never reads C (BLAS semantics).
#include "internal.h"

namespace jlc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
// Global-address-space views: the descriptor's pointers would otherwise compile
// to flat accesses, which also occupy the LDS counter and path.
typedef const __attribute__((address_space(1))) double* gcptr;
typedef __attribute__((address_space(1))) double* gptr;

constexpr int WAVES_PER_GROUP = GROUPED_WAVES_PER_GROUP;

// acc += op(A)[i0:+32, :] * op(B)[:, j0:+32] for one tile. Every lane keeps one
// base pointer per 16-row block of op(A) and per 16-column block of op(B),
// clamped to the last row / column: rows past m and columns past n then read
// real entries, which only reach C rows and columns that are never stored, so
// they need no mask. k past K does need one (both operands read as zero, so
// that 0 * Inf cannot turn into NaN), and only the last k block has such k.
template <bool A_ROW, bool B_ROW>
__device__ __forceinline__ void tile_product(const GroupedDesc& d, int i0, int j0, int r16, int g,
                                             d4 (&acc)[2][2]) {
    const int K = d.k;
    const int64_t ask = A_ROW ? 1 : d.lda;  // k stride of op(A)
    const int64_t bsk = B_ROW ? d.ldb : 1;  // k stride of op(B)
    gcptr pa[2], pb[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int row = min(i0 + 16 * i + r16, d.m - 1);
        pa[i] = (gcptr)d.a + (A_ROW ? (int64_t)row * d.lda : (int64_t)row);
        const int col = min(j0 + 16 * i + r16, d.n - 1);
        pb[i] = (gcptr)d.b + (B_ROW ? (int64_t)col : (int64_t)col * d.ldb);
    }
    // Lane group g takes k = k0 + 4g + s in MFMA step s, for A and B alike.
    double na[2][4], nb[2][4];
    auto fetch_full = [&](int k0) {
        const int kb = k0 + 4 * g;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const gcptr qa = pa[i] + kb * ask;
            const gcptr qb = pb[i] + kb * bsk;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                na[i][s] = qa[s * ask];
                nb[i][s] = qb[s * bsk];
            }
        }
    };
    auto fetch_tail = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = k0 + 4 * g + s;
            const int kc = min(k, K - 1);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const double va = pa[i][kc * ask], vb = pb[i][kc * bsk];
                na[i][s] = k < K ? va : 0.0;
                nb[i][s] = k < K ? vb : 0.0;
            }
        }
    };
    auto mma = [&](const double (&xa)[2][4], const double (&xb)[2][4]) {
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
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
    double ca[2][4], cb[2][4];
    fetch_full(0);
    for (int b = 1; b < nfull; b++) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                ca[i][s] = na[i][s];
                cb[i][s] = nb[i][s];
            }
        fetch_full(16 * b);
        mma(ca, cb);
    }
    if (K & 15) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                ca[i][s] = na[i][s];
                cb[i][s] = nb[i][s];
            }
        fetch_tail(16 * nfull);
        mma(ca, cb);
    }
    mma(na, nb);
}

// One whole 32 x 32 tile: C read (beta != 0) ahead of the product, the product (PRODUCT: alpha != 0 and
// K > 0), the store. Each layout path holds its own accumulators from start to store.
template <bool PRODUCT, bool A_ROW, bool B_ROW>
__device__ __forceinline__ void tile(const GroupedDesc& d, int i0, int j0, int r16, int g) {
    const gptr c = (gptr)d.c;
    double cold[2][2][4];
    if (d.beta != 0.0) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = i0 + 16 * i + g + 4 * r, col = j0 + 16 * j + r16;
                    cold[i][j][r] = (row < d.m && col < d.n) ? c[(int64_t)row * d.ldc + col] : 0.0;
                }
    }
    d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    if (PRODUCT) tile_product<A_ROW, B_ROW>(d, i0, j0, r16, g, acc);
    // Accumulator element r of lane l holds C(l/16 + 4r, l%16) of its 16 x 16 block.
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
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

__global__ __launch_bounds__(64 * WAVES_PER_GROUP) __attribute__((amdgpu_waves_per_eu(2, 2))) void grouped_f64(
    const GroupedDesc* __restrict__ descs,
                                                                    const uint64_t* __restrict__ tiles, int ntiles) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES_PER_GROUP + (threadIdx.x >> 6));
    const int nwaves = gridDim.x * WAVES_PER_GROUP;
    const int r16 = lane & 15;
    const int g = lane >> 4;
    // The next tile's list entry is fetched while the current tile runs, and C (when beta != 0) is read
    // before the K loop instead of after it: per tile, only the descriptor load stays serial.
    // All wavefronts run the same number of rounds, with a workgroup barrier at the start of each: the
    // four wavefronts of a workgroup, which the host gives neighbouring tiles of one problem where it
    // can, then read their shared A and B blocks at about the same time instead of drifting apart.
    uint64_t e_next = wave < ntiles ? tiles[wave] : GROUPED_SKIP;
    const int rounds = (ntiles + nwaves - 1) / nwaves;
    for (int round = 0; round < rounds; round++) {
        const int t = wave + round * nwaves;
        __syncthreads();
        const uint64_t e = e_next;
        e_next = t + nwaves < ntiles ? tiles[t + nwaves] : GROUPED_SKIP;
        if (e == GROUPED_SKIP) continue;
        const int p = __builtin_amdgcn_readfirstlane((int)(e >> 32));
        const int tm = (int)((e >> 16) & 0xffff), tn = (int)(e & 0xffff);
        const GroupedDesc d = descs[p];
        const int i0 = tm * 32, j0 = tn * 32;
        const bool a_row = d.layout_a == JLC_LAYOUT_ROW, b_row = d.layout_b == JLC_LAYOUT_ROW;
        if (d.alpha == 0.0 || d.k <= 0) tile<false, true, true>(d, i0, j0, r16, g);
        else if (a_row && b_row) tile<true, true, true>(d, i0, j0, r16, g);
        else if (a_row) tile<true, true, false>(d, i0, j0, r16, g);
        else if (b_row) tile<true, false, true>(d, i0, j0, r16, g);
        else tile<true, false, false>(d, i0, j0, r16, g);
    }
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

hipError_t launch_grouped_f64(const GroupedDesc* descs, const uint64_t* tiles, int ntiles, int grid,
                              hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(grouped_f64, dim3(grid), dim3(64 * WAVES_PER_GROUP), 0, stream, descs, tiles, ntiles);
    return hipGetLastError();
}

}  // namespace jlc
