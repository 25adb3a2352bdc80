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
// digits, diagonal sums, K chunks, recombination and alpha / beta steps), in
// two launches over a plan-owned workspace.
//  Pack, in two launches: one wavefront per 32-row block of op(A) or 32-column
//  block of op(B) of one problem (a job of the plan's job list) reads the block
//  over all K for the row / column exponents (frexp of max |x|; Inf / NaN marks
//  the row or column); then one wavefront per (job, 32-deep k block) slices the
//  block: lane l holds row / column l%32 and k in [16 (l/32), +16), the MFMA
//  fragment order of v_mfma_i32_32x32x32_i8, each value cut into S signed
//  7-bit digits packed 16 to four registers, and the S fragments (1 KiB each)
//  of the block are stored side by side. Rows / columns past m / n are zeros.
//  Product: one wavefront per 32 x 32 tile (the persistent tile list); per k
//  block it loads the S fragments of its A and B blocks (the next block's
//  while the current one is multiplied) and accumulates the S(S+1)/2 slice
//  products per diagonal. Each element is sliced once, not once per tile.
constexpr int SPLIT_NONFINITE = -100000;

// Lane's 16 values of one 32-deep k block: k in [k0, k0 + 16) of its row / column (clamped to K - 1;
// the slicing masks k >= K).
__device__ __forceinline__ void load16(gcptr p, int64_t ks, int k0, int K, double (&x)[16]) {
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = p[(int64_t)min(k0 + j, K - 1) * ks];
}

// S signed 7-bit digits of each of the 16 values scaled by 2^-e, 16 to four registers per slice:
// v <- v * 128, digit = trunc(v), v <- v - digit, S times, as split_int8.hip's pack does. Digit t is
// floor(|v| 2^(7t+7)) mod 128 with the sign of v, taken here from the 28-bit integers
// trunc(|v| 2^28), trunc(frac(|v| 2^28) 2^28), ... (every step exact in FP64).
template <int S>
__device__ __forceinline__ void slice16(const double (&x)[16], int k0, int K, int e, bool live, v4i (&out)[S]) {
    constexpr int NCH = (S + 3) / 4;  // 28-bit chunks of four digits
#pragma unroll
    for (int w = 0; w < 4; w++) {  // four values per register
        int dg[S][4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int j = 4 * w + i;
            const int k = k0 + j;
            const double v = (live && k < K && e != SPLIT_NONFINITE) ? ldexp(x[j], -e) : 0.0;
            const int neg = v < 0.0 ? -1 : 0;
            double y = fabs(v);
            int ch[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                y *= 268435456.0;  // 2^28
                const double hc = trunc(y);
                y -= hc;
                ch[c] = (int)hc;
            }
#pragma unroll
            for (int t = 0; t < S; t++) {
                const int d = (int)__builtin_amdgcn_ubfe((unsigned)ch[t >> 2], 21 - 7 * (t & 3), 7);
                dg[t][i] = (d ^ neg) - neg;
            }
        }
#pragma unroll
        for (int t = 0; t < S; t++) {
            const unsigned lo = __builtin_amdgcn_perm((unsigned)dg[t][1], (unsigned)dg[t][0], 0x0c0c0400u);
            const unsigned hi = __builtin_amdgcn_perm((unsigned)dg[t][3], (unsigned)dg[t][2], 0x0c0c0400u);
            out[t][w] = (int)__builtin_amdgcn_perm(hi, lo, 0x05040100u);
        }
    }
}

__device__ __forceinline__ void stat16(const double (&x)[16], int k0, int K, double& mx, int& bad) {
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const double v = fabs(x[j]);
        if (k0 + j < K) {
            bad |= !(v <= 1.7976931348623157e308);
            mx = fmax(mx, v);
        }
    }
}

// Row r of op(A) (is_b false) or column r of op(B) of a problem as a pointer and k stride; r is
// clamped to the last real one.
__device__ __forceinline__ void operand_line(const GroupedDesc& d, bool is_b, int r, gcptr& ptr, int64_t& ks) {
    if (!is_b) {
        const bool row = d.layout_a == JLC_LAYOUT_ROW;
        ptr = (gcptr)d.a + (row ? (int64_t)r * d.lda : (int64_t)r);
        ks = row ? 1 : d.lda;
    } else {
        const bool row = d.layout_b == JLC_LAYOUT_ROW;
        ptr = (gcptr)d.b + (row ? (int64_t)r : (int64_t)r * d.ldb);
        ks = row ? d.ldb : 1;
    }
}

// Pack, launch 1: one wavefront per job (a 32-row block of op(A) or 32-column block of op(B)) writes
// the exponents of its 32 rows / columns. Two register sets in turn keep the next k block's loads in
// flight while the current one is scanned (every load unconditional, the block index clamped).
__global__ __launch_bounds__(64 * WAVES_PER_GROUP) void grouped_split_expo(const GroupedDesc* __restrict__ descs,
                                                                           const GroupedPack* __restrict__ packs,
                                                                           const uint64_t* __restrict__ jobs,
                                                                           int64_t njobs, int* __restrict__ expo) {
    const int lane = threadIdx.x & 63;
    const int64_t job = (int64_t)blockIdx.x * WAVES_PER_GROUP + (threadIdx.x >> 6);
    if (job >= njobs) return;
    const uint64_t e = jobs[job];
    const int p = __builtin_amdgcn_readfirstlane((int)(e >> 32));
    const bool is_b = (e >> 31) & 1u;
    const int blk = (int)(e & 0x7fffffffu);
    const GroupedDesc d = descs[p];
    const int K = d.k, kblocks = (K + 31) / 32;
    const int r32 = lane & 31, h = lane >> 5;
    const int R = is_b ? d.n : d.m;
    const int r = blk * 32 + r32;
    gcptr ptr;
    int64_t ks;
    operand_line(d, is_b, min(r, R - 1), ptr, ks);
    double mx = 0.0;
    int bad = 0;
    if (ks == 1) {
        // k-contiguous lines: load j of a k block covers lines 2j and 2j + 1 of the block (lane l: line
        // 2j + l / 32, k = l % 32), two 256-byte runs, two blocks in flight; each lane keeps a maximum
        // per line (Inf / NaN count as +Inf), then a butterfly over the 32 lanes of each half leaves
        // lane l with line 2 (8 b4 + 4 b3 + 2 b2 + b1) + l / 32 (b: the bits of l % 32) reduced.
        const int kl = lane & 31;
        const gcptr base = (gcptr)(is_b ? d.b : d.a);
        const int64_t ld = is_b ? d.ldb : d.lda;  // line stride of a k-contiguous operand
        double m16[16], v0[16], v1[16];
#pragma unroll
        for (int j = 0; j < 16; j++) m16[j] = 0.0;
        auto fetch = [&](double (&v)[16], int k0) {
            const int k = min(k0 + kl, K - 1);
#pragma unroll
            for (int j = 0; j < 16; j++) v[j] = base[(int64_t)min(blk * 32 + 2 * j + h, R - 1) * ld + k];
        };
        auto scan = [&](const double (&v)[16], int k0) {
            const bool in = k0 + kl < K;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                double a = fabs(v[j]);
                a = a <= 1.7976931348623157e308 ? a : __builtin_inf();
                m16[j] = in ? fmax(m16[j], a) : m16[j];
            }
        };
        fetch(v0, 0);
        for (int kb = 0; kb < kblocks; kb += 2) {
            fetch(v1, 32 * min(kb + 1, kblocks - 1));
            scan(v0, 32 * kb);
            if (kb + 1 >= kblocks) break;
            fetch(v0, 32 * min(kb + 2, kblocks - 1));
            scan(v1, 32 * (kb + 1));
        }
        double m8[8], m4[4], m2[2];
        const bool b4 = kl & 16, b3 = kl & 8, b2 = kl & 4, b1 = kl & 2;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const double send = b4 ? m16[i] : m16[i + 8];
            m8[i] = fmax(b4 ? m16[i + 8] : m16[i], __shfl_xor(send, 16, 64));
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double send = b3 ? m8[i] : m8[i + 4];
            m4[i] = fmax(b3 ? m8[i + 4] : m8[i], __shfl_xor(send, 8, 64));
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const double send = b2 ? m4[i] : m4[i + 2];
            m2[i] = fmax(b2 ? m4[i + 2] : m4[i], __shfl_xor(send, 4, 64));
        }
        double m1 = fmax(b1 ? m2[1] : m2[0], __shfl_xor(b1 ? m2[0] : m2[1], 2, 64));
        m1 = fmax(m1, __shfl_xor(m1, 1, 64));
        const int line = 2 * ((b4 ? 8 : 0) + (b3 ? 4 : 0) + (b2 ? 2 : 0) + (b1 ? 1 : 0)) + h;
        int ex = 0;
        if (m1 > 0.0) frexp(m1, &ex);
        if (!(m1 <= 1.7976931348623157e308)) ex = SPLIT_NONFINITE;
        if (blk * 32 + line >= R) ex = 0;
        const GroupedPack pk = packs[p];
        if ((kl & 1) == 0) expo[(is_b ? pk.eb_off : pk.ea_off) + blk * 32 + line] = ex;
        return;
    } else {
        double x0[16], x1[16];
        load16(ptr, ks, 16 * h, K, x0);
        for (int kb = 0; kb < kblocks; kb += 2) {
            load16(ptr, ks, 32 * min(kb + 1, kblocks - 1) + 16 * h, K, x1);
            stat16(x0, 32 * kb + 16 * h, K, mx, bad);
            if (kb + 1 >= kblocks) break;
            load16(ptr, ks, 32 * min(kb + 2, kblocks - 1) + 16 * h, K, x0);
            stat16(x1, 32 * (kb + 1) + 16 * h, K, mx, bad);
        }
        mx = fmax(mx, __shfl_xor(mx, 32, 64));
        bad |= __shfl_xor(bad, 32, 64);
    }
    int ex = 0;
    if (mx > 0.0) frexp(mx, &ex);
    if (bad) ex = SPLIT_NONFINITE;
    if (r >= R) ex = 0;
    const GroupedPack pk = packs[p];
    if (h == 0) expo[(is_b ? pk.eb_off : pk.ea_off) + r] = ex;
}

// Pack, launch 2: one wavefront per (job, 32-deep k block), listed as (job << 24) | k block, slices
// the block with the exponents launch 1 wrote and stores its S fragments side by side. Rows /
// columns past m / n are stored as zeros.
template <int S>
__global__ __launch_bounds__(64 * WAVES_PER_GROUP) void grouped_split_slice(const GroupedDesc* __restrict__ descs,
                                                                            const GroupedPack* __restrict__ packs,
                                                                            const uint64_t* __restrict__ jobs,
                                                                            const uint64_t* __restrict__ sjobs,
                                                                            int64_t nsjobs, v4i* __restrict__ frag,
                                                                            const int* __restrict__ expo) {
    const int lane = threadIdx.x & 63;
    const int64_t sj = (int64_t)blockIdx.x * WAVES_PER_GROUP + (threadIdx.x >> 6);
    if (sj >= nsjobs) return;
    const uint64_t se = sjobs[sj];
    const int kb = (int)(se & 0xffffffu);
    const uint64_t e = jobs[se >> 24];
    const int p = __builtin_amdgcn_readfirstlane((int)(e >> 32));
    const bool is_b = (e >> 31) & 1u;
    const int blk = (int)(e & 0x7fffffffu);
    const GroupedDesc d = descs[p];
    const GroupedPack pk = packs[p];
    const int K = d.k, kblocks = (K + 31) / 32;
    const int r32 = lane & 31, h = lane >> 5;
    const int R = is_b ? d.n : d.m;
    const int r = blk * 32 + r32;
    gcptr ptr;
    int64_t ks;
    operand_line(d, is_b, min(r, R - 1), ptr, ks);
    double x[16];
    if (ks == 1) {
        // k-contiguous lines: load j covers lines 2j and 2j + 1 of the block as two 256-byte runs
        // (lane l: line 2j + l / 32, k = l % 32); the block is turned to fragment order through LDS.
        __shared__ double tile[WAVES_PER_GROUP][32][33];
        double (*tw)[33] = tile[threadIdx.x >> 6];
        const int kl = lane & 31;
        const int k = min(32 * kb + kl, K - 1);
        const gcptr base = (gcptr)(is_b ? d.b : d.a);
        const int64_t ld = is_b ? d.ldb : d.lda;  // line stride of a k-contiguous operand
        double v[16];
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = base[(int64_t)min(blk * 32 + 2 * j + h, R - 1) * ld + k];
#pragma unroll
        for (int j = 0; j < 16; j++) tw[2 * j + h][kl] = v[j];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int i = 0; i < 16; i++) x[i] = tw[r32][16 * h + i];
    } else {
        load16(ptr, ks, 32 * kb + 16 * h, K, x);
    }
    const int ex = expo[(is_b ? pk.eb_off : pk.ea_off) + r];
    v4i f[S];
    slice16<S>(x, 32 * kb + 16 * h, K, ex, r < R, f);
    v4i* out = frag + ((is_b ? pk.b_off : pk.a_off) + ((int64_t)blk * kblocks + kb) * S) * 64 + lane;
#pragma unroll
    for (int t = 0; t < S; t++) out[t * 64] = f[t];
}

template <int S>
__device__ __forceinline__ void split_tile(const GroupedDesc& d, const GroupedPack& pk, int tm, int tn, int lane,
                                           const v4i* __restrict__ frag, const int* __restrict__ expo) {
    const int K = d.k, kblocks = (K + 31) / 32;
    const int r32 = lane & 31, h = lane >> 5;
    const int i0 = tm * 32, j0 = tn * 32;
    const int ea = expo[pk.ea_off + i0 + r32];
    const int eb = expo[pk.eb_off + j0 + r32];
    const v4i* fa = frag + (pk.a_off + (int64_t)tm * kblocks * S) * 64 + lane;
    const v4i* fb = frag + (pk.b_off + (int64_t)tn * kblocks * S) * 64 + lane;
    // INT32 diagonal sums stay exact for `chunk` 32-deep blocks: S * chunk * 32 * 127^2 < 2^31.
    const int chunk = (int)(2147483647LL / ((long long)S * 32 * 127 * 127));
    const gptr c = (gptr)d.c;
    const int col = j0 + r32;
    // C (beta != 0) is read ahead of the products, all 16 entries of the lane at once, instead of one
    // dependent load per stored entry.
    double cold[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = i0 + 8 * (r >> 2) + 4 * h + (r & 3);
        cold[r] = (d.beta != 0.0 && row < d.m && col < d.n) ? c[(int64_t)row * d.ldc + col] : 0.0;
    }
    v4i a0[S], b0[S], a1[S], b1[S];
    auto load = [&](v4i (&xa)[S], v4i (&xb)[S], int kb) {
#pragma unroll
        for (int t = 0; t < S; t++) {
            xa[t] = fa[((int64_t)kb * S + t) * 64];
            xb[t] = fb[((int64_t)kb * S + t) * 64];
        }
    };
    for (int c0 = 0; c0 < kblocks; c0 += chunk) {
        v16i acc[S];
#pragma unroll
        for (int dd = 0; dd < S; dd++) acc[dd] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        auto mma = [&](const v4i (&af)[S], const v4i (&bf)[S]) {
#pragma unroll
            for (int t = 0; t < S; t++)
#pragma unroll
                for (int u = S - 1 - t; u >= 0; u--)
                    acc[t + u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[t], bf[u], acc[t + u], 0, 0, 0);
        };
        const int c1 = min(kblocks, c0 + chunk);
        // Two register sets in turn (one for S >= 7, whose accumulators leave no room: the next block
        // is then loaded into the set as the multiplies release it); every load is unconditional (the
        // block index clamped to the chunk's last), so the wait before each multiply covers only that
        // set's loads.
        load(a0, b0, c0);
        if (S >= 7) {
            for (int kb = c0; kb < c1; kb++) {
                mma(a0, b0);
                if (kb + 1 < c1) load(a0, b0, kb + 1);
            }
        } else {
            for (int kb = c0; kb < c1; kb += 2) {
                load(a1, b1, min(kb + 1, c1 - 1));
                mma(a0, b0);
                if (kb + 1 >= c1) break;
                load(a0, b0, min(kb + 2, c1 - 1));
                mma(a1, b1);
            }
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
                if (c0 == 0) *pc = (d.beta == 0.0) ? d.alpha * v : fma(d.alpha, v, d.beta * cold[r]);
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
                                                                         const GroupedPack* __restrict__ packs,
                                                                         const uint64_t* __restrict__ tiles,
                                                                         int ntiles, const v4i* __restrict__ frag,
                                                                         const int* __restrict__ expo) {
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
        if (d.alpha == 0.0 || d.k <= 0) scale_tile(d, tm * 32, tn * 32, lane);
        else split_tile<S>(d, packs[p], tm, tn, lane, frag, expo);
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

hipError_t launch_grouped_split_i8(int slices, const GroupedDesc* descs, const GroupedPack* packs,
                                   const uint64_t* jobs, int64_t njobs, const uint64_t* sjobs, int64_t nsjobs,
                                   const uint64_t* tiles, int ntiles, int grid, void* frag, int* expo,
                                   hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    const dim3 b(64 * WAVES_PER_GROUP);
    const dim3 eg((unsigned)((njobs + WAVES_PER_GROUP - 1) / WAVES_PER_GROUP));
    const dim3 sg((unsigned)((nsjobs + WAVES_PER_GROUP - 1) / WAVES_PER_GROUP)), g(grid);
    v4i* fr = static_cast<v4i*>(frag);
    if (njobs > 0) {
        hipLaunchKernelGGL(grouped_split_expo, eg, b, 0, stream, descs, packs, jobs, njobs, expo);
        const hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    switch (slices) {
#define JLC_GROUPED_SPLIT_CASE(S_)                                                                                \
    case S_:                                                                                                      \
        if (nsjobs > 0)                                                                                           \
            hipLaunchKernelGGL(grouped_split_slice<S_>, sg, b, 0, stream, descs, packs, jobs, sjobs, nsjobs, fr,  \
                               expo);                                                                             \
        hipLaunchKernelGGL(grouped_split_i8<S_>, g, b, 0, stream, descs, packs, tiles, ntiles, fr, expo);         \
        break;
        JLC_GROUPED_SPLIT_CASE(3) JLC_GROUPED_SPLIT_CASE(4) JLC_GROUPED_SPLIT_CASE(5) JLC_GROUPED_SPLIT_CASE(6)
        JLC_GROUPED_SPLIT_CASE(7) JLC_GROUPED_SPLIT_CASE(8) JLC_GROUPED_SPLIT_CASE(9)
#undef JLC_GROUPED_SPLIT_CASE
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
