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
// sorted by decreasing K so the longest tiles start first. A persistent grid of
// wavefronts deals the list round-robin (wavefront w takes entries w, w + W,
// w + 2W, ...): no atomics, no counters to clear, so the launch is safe to
// capture in a HIP graph and to replay.
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

constexpr int WAVES_PER_GROUP = 4;

__device__ __forceinline__ double load_a(const GroupedDesc& d, int row, int k) {
    if (row >= d.m || k >= d.k) return 0.0;
    return d.layout_a == JLC_LAYOUT_ROW ? d.a[(int64_t)row * d.lda + k] : d.a[(int64_t)k * d.lda + row];
}

__device__ __forceinline__ double load_b(const GroupedDesc& d, int k, int col) {
    if (col >= d.n || k >= d.k) return 0.0;
    return d.layout_b == JLC_LAYOUT_ROW ? d.b[(int64_t)k * d.ldb + col] : d.b[(int64_t)col * d.ldb + k];
}

__global__ __launch_bounds__(64 * WAVES_PER_GROUP) void grouped_f64(const GroupedDesc* __restrict__ descs,
                                                                    const uint64_t* __restrict__ tiles, int ntiles) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES_PER_GROUP + (threadIdx.x >> 6));
    const int nwaves = gridDim.x * WAVES_PER_GROUP;
    const int r16 = lane & 15;
    const int g = lane >> 4;
    // The next tile's list entry is fetched while the current tile runs, and C (when beta != 0) is read
    // before the K loop instead of after it: per tile, only the descriptor load stays serial.
    uint64_t e_next = wave < ntiles ? tiles[wave] : 0;
    for (int t = wave; t < ntiles; t += nwaves) {
        const uint64_t e = e_next;
        if (t + nwaves < ntiles) e_next = tiles[t + nwaves];
        const int p = __builtin_amdgcn_readfirstlane((int)(e >> 32));
        const int tm = (int)((e >> 16) & 0xffff), tn = (int)(e & 0xffff);
        const GroupedDesc d = descs[p];
        const int i0 = tm * 32, j0 = tn * 32;
        double cold[2][2][4];
        if (d.beta != 0.0) {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = i0 + 16 * i + g + 4 * r, col = j0 + 16 * j + r16;
                        cold[i][j][r] = (row < d.m && col < d.n) ? d.c[(int64_t)row * d.ldc + col] : 0.0;
                    }
        }
        d4 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
        if (d.alpha != 0.0 && d.k > 0) {
            double na[2][4], nb[2][4];
            auto fetch = [&](int k0) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int k = k0 + 4 * g + s;
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        na[i][s] = load_a(d, i0 + 16 * i + r16, k);
                        nb[i][s] = load_b(d, k, j0 + 16 * i + r16);
                    }
                }
            };
            fetch(0);
            for (int k0 = 0; k0 < d.k; k0 += 16) {
                double ca[2][4], cb[2][4];
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        ca[i][s] = na[i][s];
                        cb[i][s] = nb[i][s];
                    }
                if (k0 + 16 < d.k) fetch(k0 + 16);
#pragma unroll
                for (int s = 0; s < 4; s++)
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < 2; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[i][s], cb[j][s], acc[i][j], 0, 0, 0);
            }
        }
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
                        d.c[(int64_t)row * d.ldc + col] = d.beta == 0.0 ? v : v + d.beta * cold[i][j][r];
                    }
                }
            }
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
