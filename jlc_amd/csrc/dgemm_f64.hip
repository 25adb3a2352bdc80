// FP64 MFMA DGEMM for gfx950 (CDNA4).
//
// C = alpha * op(A) * op(B) + beta * C, C row-major, op(A) / op(B) row- or
// column-major (see include/jlc_hip.h). This is the device body of the HIP
// kernel choice behind the reference's GEMM API; the layout variants fuse the
// reference's pack routines (PackingUtils.packA alpha fold, java:37-56;
// BLAS3Kernels.packAStridedTransposed / packAStridedColMajor /
// packBStridedColMajor, src/main/java/net/faulj/compute/BLAS3Kernels.java:702-787)
// into the global->LDS staging of each tile, so no separate pack pass runs.
//
// Macro tile 128 x 128 x 16 per 256-thread workgroup, four wavefronts in a
// 2 x 2 grid, each wavefront owning a 64 x 64 block of C as 4 x 4
// v_mfma_f64_16x16x4 accumulators; two workgroups resident per CU. Operand tiles live in LDS as [k][m] and
// [k][n] with k in pairs (16-byte entries, see slot()), two buffers: while
// the MFMAs consume one buffer, the next K tile is fetched into registers and
// written to the other, one barrier per K tile (64 KiB per group, two groups
// per CU). Each ds_read_b128 fragment load feeds two MFMA steps.
//
// Edges are masked: loads outside [0,M) x [0,K) or [0,K) x [0,N) read as
// zero and stores outside C are skipped, so any M, N, K >= 1 is valid.
#include "internal.h"

namespace jlc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int BM = 128;   // rows of C per workgroup
constexpr int BN = 128;   // cols of C per workgroup
constexpr int BK = 16;    // K per stage
constexpr int NT = 256;   // threads per workgroup
constexpr int SROW = 128; // doubles per k row of a staged tile (BM == BN)

// Staged tiles hold k in pairs: entry (kp, r) is the 16-byte pair
// (x(r, 2kp), x(r, 2kp+1)) at slot kp*128 + (r ^ (kp & 7)). One ds_read_b128
// then feeds two MFMA steps, and the XOR spreads each 8-lane phase of a
// transposing store or fragment read over 8 distinct 16-byte bank groups.
__device__ __forceinline__ int slot(int kp, int r) { return kp * SROW + (r ^ (kp & 7)); }

// Loads one 128 x 16 operand tile as 4 pair entries per thread.
// KCONTIG: element (r, k) at base[r*ld + k]: thread reads the pair (r, 2kp..2kp+1) directly.
// !KCONTIG: element (r, k) at base[k*ld + r]: thread reads rows k = 2kp, 2kp+1 at r, r+1 and
//           regroups them into the pairs of r and r+1.
// MODE: LD_INTERIOR, every element in range; LD_EDGE, every k in range but r may pass rdim: r is
// clamped to rdim - 1 (those rows / columns only reach outputs that are never stored) and every load
// stays unconditional; LD_MASK, k may pass kdim too: out-of-range elements read as zero.
constexpr int LD_INTERIOR = 0, LD_EDGE = 1, LD_MASK = 2;

template <bool KCONTIG, bool VEC, int MODE>
__device__ __forceinline__ void load_tile(const double* __restrict__ base, int64_t ld, int rdim, int kdim, int r0,
                                          int k0, int t, d2 (&regs)[4]) {
    if (MODE == LD_EDGE) {
        if (KCONTIG) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = min(r0 + (t >> 3) + 32 * i, rdim - 1);
                const double* p = base + (int64_t)r * ld + k0 + (t & 7) * 2;
                regs[i] = VEC ? *reinterpret_cast<const d2*>(p) : d2{p[0], p[1]};
            }
        } else {
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = r0 + (t & 63) * 2;
                const double* p = base + (int64_t)(k0 + 2 * ((t >> 6) + 4 * i)) * ld;
                const int64_t c0 = min(r, rdim - 1), c1 = min(r + 1, rdim - 1);
                regs[2 * i] = d2{p[c0], p[c0 + ld]};
                regs[2 * i + 1] = d2{p[c1], p[c1 + ld]};
            }
        }
        return;
    }
    if (MODE == LD_INTERIOR) {
        // Interior tile: every element is in range.
        if (KCONTIG) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const double* p = base + (int64_t)(r0 + (t >> 3) + 32 * i) * ld + k0 + (t & 7) * 2;
                regs[i] = VEC ? *reinterpret_cast<const d2*>(p) : d2{p[0], p[1]};
            }
        } else {
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const double* p = base + (int64_t)(k0 + 2 * ((t >> 6) + 4 * i)) * ld + r0 + (t & 63) * 2;
                const d2 lo = VEC ? *reinterpret_cast<const d2*>(p) : d2{p[0], p[1]};
                const d2 hi = VEC ? *reinterpret_cast<const d2*>(p + ld) : d2{p[ld], p[ld + 1]};
                regs[2 * i] = d2{lo.x, hi.x};
                regs[2 * i + 1] = d2{lo.y, hi.y};
            }
        }
        return;
    }
    if (KCONTIG) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = r0 + (t >> 3) + 32 * i;
            const int kk = k0 + (t & 7) * 2;
            d2 v = {0.0, 0.0};
            if (r < rdim) {
                const double* p = base + (int64_t)r * ld + kk;
                if (VEC && kk + 1 < kdim) {
                    v = *reinterpret_cast<const d2*>(p);
                } else {
                    if (kk < kdim) v.x = p[0];
                    if (kk + 1 < kdim) v.y = p[1];
                }
            }
            regs[i] = v;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int kk = k0 + 2 * ((t >> 6) + 4 * i);
            const int r = r0 + (t & 63) * 2;
            d2 lo = {0.0, 0.0}, hi = {0.0, 0.0};  // rows kk and kk+1, columns r and r+1
            const double* p = base + (int64_t)kk * ld + r;
            if (VEC && r + 1 < rdim) {
                if (kk < kdim) lo = *reinterpret_cast<const d2*>(p);
                if (kk + 1 < kdim) hi = *reinterpret_cast<const d2*>(p + ld);
            } else {
                if (kk < kdim && r < rdim) lo.x = p[0];
                if (kk < kdim && r + 1 < rdim) lo.y = p[1];
                if (kk + 1 < kdim && r < rdim) hi.x = p[ld];
                if (kk + 1 < kdim && r + 1 < rdim) hi.y = p[ld + 1];
            }
            regs[2 * i] = d2{lo.x, hi.x};
            regs[2 * i + 1] = d2{lo.y, hi.y};
        }
    }
}

// Writes the pair entries of load_tile into a staged tile.
template <bool KCONTIG>
__device__ __forceinline__ void store_tile(d2* __restrict__ s, int t, const d2 (&regs)[4]) {
    if (KCONTIG) {
#pragma unroll
        for (int i = 0; i < 4; i++) s[slot(t & 7, (t >> 3) + 32 * i)] = regs[i];
    } else {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int kp = (t >> 6) + 4 * i;
            const int r = (t & 63) * 2;
            s[slot(kp, r)] = regs[2 * i];
            s[slot(kp, r + 1)] = regs[2 * i + 1];
        }
    }
}

// Runs k pairs [kp0, kp1) (each two MFMA steps) of one staged K tile on a
// wavefront's 4 x 4 accumulator blocks. Lane group g of step j takes
// k = 8b + 2g + j, the same for A and B, so the products sum the right terms.
__device__ __forceinline__ void mfma_steps(const d2* __restrict__ as, const d2* __restrict__ bs, int b0, int b1,
                                           int wm, int wn, int g, int c16, d4 (&acc)[4][4]) {
#pragma unroll
    for (int b = b0; b < b1; b++) {
        const int kp = 4 * b + g;
        d2 af[4], bf[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            af[i] = as[slot(kp, wm * 64 + i * 16 + c16)];
            bf[i] = bs[slot(kp, wn * 64 + i * 16 + c16)];
        }
#pragma unroll
        for (int j2 = 0; j2 < 2; j2++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(j2 ? af[i].y : af[i].x, j2 ? bf[j].y : bf[j].x,
                                                                     acc[i][j], 0, 0, 0);
    }
}

// Maps a linear workgroup id to a C tile: tiles are walked in column strips of
// GROUP tile rows. Workgroups are dealt to the eight XCDs round-robin, so the
// ones resident on one XCD at a time hold GROUP/8 tile rows of a strip and
// share those A panels in its L2. Measured at 16384^3 this beats regrouping
// ids into per-XCD contiguous runs (fewer L2 misses, 2.9% shorter run).
__device__ __forceinline__ void tile_coords(int bid, int nblocks, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int GROUP = 16;
    (void)nblocks;
    const int lid = bid;
    int per_group = GROUP * tiles_n;
    int g = lid / per_group;
    int first_m = g * GROUP;
    int gsize = min(tiles_m - first_m, GROUP);
    int in = lid % per_group;
    tm = first_m + in % gsize;
    tn = in / gsize;
}

// Main loop and epilogue of one C tile. MODE (see load_tile): LD_INTERIOR for interior tiles with
// K a multiple of BK; LD_EDGE for the rest, which clamp rows / columns on every K tile, zero-fill the
// last K tile when K is not a multiple of BK (LD_MASK loads), and skip stores outside C.
template <bool A_COL, bool B_COL, bool VEC, int MODE>
__device__ __forceinline__ void gemm_tile(int M, int N, int K, double alpha, const double* __restrict__ A,
                                          int64_t lda, const double* __restrict__ B, int64_t ldb, double beta,
                                          double* __restrict__ C, int64_t ldc, int m0, int n0, d2 (*As)[BK / 2 * SROW],
                                          d2 (*Bs)[BK / 2 * SROW]) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wid = t >> 6;
    const int wm = wid >> 1;
    const int wn = wid & 1;

    // op(A) row-major => k contiguous; op(B) column-major => k contiguous.
    constexpr bool A_KC = !A_COL;
    constexpr bool B_KC = B_COL;

    d4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    constexpr bool MASKED = MODE != LD_INTERIOR;
    // Edge tiles clamp rows / columns on every K tile and zero-fill only the last K tile, when K is
    // not a multiple of BK (tail is uniform per workgroup).
    const int nk = (K + BK - 1) / BK;
    const bool tail = MASKED && K % BK != 0;
    d2 ra[4], rb[4];
    if (MASKED && tail && nk == 1) {
        load_tile<A_KC, VEC, LD_MASK>(A, lda, M, K, m0, 0, t, ra);
        load_tile<B_KC, VEC, LD_MASK>(B, ldb, N, K, n0, 0, t, rb);
    } else {
        load_tile<A_KC, VEC, MODE>(A, lda, M, K, m0, 0, t, ra);
        load_tile<B_KC, VEC, MODE>(B, ldb, N, K, n0, 0, t, rb);
    }
    store_tile<A_KC>(As[0], t, ra);
    store_tile<B_KC>(Bs[0], t, rb);
    __syncthreads();

    const int g = lane >> 4;   // k group within an MFMA step
    const int c16 = lane & 15; // column within a 16-wide fragment
    // Branch-free steady state: every iteration but the last prefetches the
    // next K tile into registers, runs half of the current tile's MFMAs,
    // writes the prefetched tile to the idle LDS buffer, then runs the other
    // half, so the LDS writes sit between independent MFMAs.
    int kt = 0;
    for (; kt < nk - 1; kt++) {
        const int cur = kt & 1;
        if (MASKED && tail && kt + 1 == nk - 1) {
            load_tile<A_KC, VEC, LD_MASK>(A, lda, M, K, m0, (kt + 1) * BK, t, ra);
            load_tile<B_KC, VEC, LD_MASK>(B, ldb, N, K, n0, (kt + 1) * BK, t, rb);
        } else {
            load_tile<A_KC, VEC, MODE>(A, lda, M, K, m0, (kt + 1) * BK, t, ra);
            load_tile<B_KC, VEC, MODE>(B, ldb, N, K, n0, (kt + 1) * BK, t, rb);
        }
        mfma_steps(As[cur], Bs[cur], 0, 1, wm, wn, g, c16, acc);
        store_tile<A_KC>(As[cur ^ 1], t, ra);
        store_tile<B_KC>(Bs[cur ^ 1], t, rb);
        mfma_steps(As[cur], Bs[cur], 1, 2, wm, wn, g, c16, acc);
        __syncthreads();
    }
    mfma_steps(As[kt & 1], Bs[kt & 1], 0, 2, wm, wn, g, c16, acc);

    // v_mfma_f64_16x16x4 accumulator layout: element r of lane l holds C(l/16 + 4*r, l%16)
    // of its 16 x 16 block.
    if (beta == 0.0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int col = n0 + wn * 64 + j * 16 + c16;
                if (MASKED && col >= N) continue;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = m0 + wm * 64 + i * 16 + g + 4 * r;
                    if (!MASKED || row < M) C[(int64_t)row * ldc + col] = alpha * acc[i][j][r];
                }
            }
        }
        return;
    }
    // beta != 0: C is read eight entries (two 16 x 16 fragments' share of the lane) at a time, the next
    // group's loads issued before the current group's stores, so the reads overlap one another instead of
    // each waiting behind the previous entry's store (a short K loop would otherwise spend most of its
    // time here).
    double cv[2][2][4];
    auto load_c = [&](int q, double (&x)[2][4]) {
        const int i = q >> 1;
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int col = n0 + wn * 64 + ((q & 1) * 2 + jj) * 16 + c16;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 64 + i * 16 + g + 4 * r;
                x[jj][r] = (!MASKED || (row < M && col < N)) ? C[(int64_t)row * ldc + col] : 0.0;
            }
        }
    };
    load_c(0, cv[0]);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (q < 7) load_c(q + 1, cv[(q + 1) & 1]);
        const int i = q >> 1;
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int j = (q & 1) * 2 + jj;
            const int col = n0 + wn * 64 + j * 16 + c16;
            if (MASKED && col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 64 + i * 16 + g + 4 * r;
                if (!MASKED || row < M) {
                    double v = alpha * acc[i][j][r];
                    v += beta * cv[q & 1][jj][r];
                    C[(int64_t)row * ldc + col] = v;
                }
            }
        }
    }
}

// Two wavefronts per SIMD (two workgroups per CU) hide the barrier and LDS
// latency of one group behind the other's MFMAs; the register budget is 256.
template <bool A_COL, bool B_COL, bool VEC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2))) void dgemm_f64_mfma(int M, int N, int K, double alpha,
                                                     const double* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ B, int64_t ldb, double beta,
                                                     double* __restrict__ C, int64_t ldc, int tiles_m,
                                                     int tiles_n, int kslice, int64_t c_slice) {
    __shared__ __attribute__((aligned(16))) d2 As[2][BK / 2 * SROW];
    __shared__ __attribute__((aligned(16))) d2 Bs[2][BK / 2 * SROW];
    // Split K: slice blockIdx.y covers k in [y*kslice, min(K, (y+1)*kslice)) and
    // writes its partial product to C + y*c_slice (kslice is a multiple of BK).
    if (gridDim.y > 1) {
        const int64_t koff = (int64_t)blockIdx.y * kslice;
        K = (int)min((int64_t)kslice, (int64_t)K - koff);
        A += A_COL ? koff * lda : koff;
        B += B_COL ? koff : koff * ldb;
        C += blockIdx.y * c_slice;
    }
    int tm, tn;
    tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM;
    const int n0 = tn * BN;
    // Uniform per workgroup: interior tiles take the unchecked path.
    if (m0 + BM <= M && n0 + BN <= N && K % BK == 0)
        gemm_tile<A_COL, B_COL, VEC, LD_INTERIOR>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0, As, Bs);
    else
        gemm_tile<A_COL, B_COL, VEC, LD_EDGE>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0, As, Bs);
}

__global__ __launch_bounds__(256) void scale_kernel(int M, int N, double beta, double* __restrict__ C, int64_t ldc) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)M * N;
    for (; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = idx / N, j = idx % N;
        double* p = C + i * ldc + j;
        *p = (beta == 0.0) ? 0.0 : beta * *p;
    }
}

template <bool A_COL, bool B_COL, bool VEC>
hipError_t launch_variant(int M, int N, int K, double alpha, const double* A, int64_t lda, const double* B,
                          int64_t ldb, double beta, double* C, int64_t ldc, int slices, int kslice, int64_t c_slice,
                          hipStream_t stream) {
    int tiles_m = (M + BM - 1) / BM;
    int tiles_n = (N + BN - 1) / BN;
    int64_t nblocks = (int64_t)tiles_m * tiles_n;
    hipLaunchKernelGGL((dgemm_f64_mfma<A_COL, B_COL, VEC>), dim3((unsigned)nblocks, (unsigned)slices), dim3(NT), 0,
                       stream, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n, kslice, c_slice);
    return hipGetLastError();
}

// C = alpha * sum_s P_s + beta * C, slices summed in order (deterministic).
__global__ __launch_bounds__(256) void splitk_reduce(int M, int N, int slices, const double* __restrict__ P,
                                                     int64_t p_slice, double alpha, double beta,
                                                     double* __restrict__ C, int64_t ldc) {
    const int64_t total = (int64_t)M * N;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < slices; q++) s += P[q * p_slice + idx];
        double* c = C + (idx / N) * ldc + idx % N;
        *c = beta == 0.0 ? alpha * s : alpha * s + beta * *c;
    }
}

}  // namespace

hipError_t launch_scale(int64_t m, int64_t n, double beta, double* c, int64_t ldc, hipStream_t stream,
                        TileInfo* info) {
    if (m <= 0 || n <= 0 || beta == 1.0) return hipSuccess;
    int64_t total = m * n;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, (int)m, (int)n, beta, c, ldc);
    if (info) info->launches++;
    return hipGetLastError();
}

int dgemm_splitk_slices(int64_t m, int64_t n, int64_t k, int device_cus, int tiles_per_cu, int min_kslice,
                        bool quant) {
    // Few output tiles and a long K: give every CU work by cutting K, up to tiles_per_cu workgroups per CU
    // (and slices of at least min_kslice). Within that bound the count is the one with the least modelled
    // time, so the last round of resident workgroups (two per CU) is not left mostly empty: in units of one
    // tile's 16-deep step, ceil(tiles * s / slots) rounds of ceil(k / (16 s)) + 2 steps (the +2: a
    // workgroup's prologue and epilogue), plus (s + 2) passes over the m x n partials at 5 TB/s (4e-7 steps
    // per element and pass). E.g. 60 tiles, k = 32768 on 256 CUs: 18 slices (1080 workgroups, three rounds,
    // the third 11% full) -> 17 (1020, two rounds).
    if (tiles_per_cu <= 0) return 1;
    const int64_t tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
    const int64_t want = ((int64_t)tiles_per_cu * device_cus + tiles - 1) / tiles;
    int64_t smax = want;
    if (min_kslice < BK) min_kslice = BK;
    if (smax > k / min_kslice) smax = k / min_kslice;
    if (smax > 64) smax = 64;
    if (smax < 2) return 1;
    if (!quant) return (int)smax;  // the tuning table's splitk_quant 0: the full count, no round model
    const int64_t slots = 2 * (int64_t)(device_cus > 0 ? device_cus : 1);
    auto cost = [&](int64_t s) {
        const double rounds = (double)((tiles * s + slots - 1) / slots);
        const double steps = (double)((k + 16 * s - 1) / (16 * s)) + 2.0;
        return rounds * steps + (s > 1 ? (double)(s + 2) * (double)m * (double)n * 4e-7 : 0.0);
    };
    int64_t best = smax;
    for (int64_t s = smax - 1; s >= 1; s--)
        if (cost(s) < cost(best)) best = s;
    return best < 2 ? 1 : (int)best;
}

int dgemm_quant_slices(int64_t m, int64_t n, int64_t k, int device_cus, int64_t max_partial_bytes) {
    // Many tiles, but a last round of resident tiles (two workgroups per CU) that is mostly empty: cutting
    // K into s <= 4 slices fills the rounds at the cost of writing and re-reading s partial products. The
    // time in units of one tile's 16-deep step (2 * 128 * 128 * 16 flops at 1/512 of 67 TFLOP/s, 4.0 us):
    //   ceil(tiles * s / slots) * ceil(k / (16 s))  +  (s + 2) * m * n * 8 bytes at 5 TB/s (4e-7 steps per
    //   element and pass).
    const int64_t tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
    const int64_t slots = 2 * (int64_t)device_cus;
    if (device_cus <= 0 || tiles <= slots || k < 512) return 1;
    auto cost = [&](int s) {
        const double rounds = (double)((tiles * s + slots - 1) / slots);
        const double steps = (double)((k + 16 * s - 1) / (16 * s));
        return rounds * steps + (s > 1 ? (s + 2) * (double)m * (double)n * 4e-7 : 0.0);
    };
    int best = 1;
    for (int s = 2; s <= 4; s++)
        if ((int64_t)s * m * n * (int64_t)sizeof(double) <= max_partial_bytes && cost(s) < 0.97 * cost(best)) best = s;
    return best;
}

int64_t dgemm_splitk_kslice(int64_t k, int slices) {
    int64_t ks = (k + slices - 1) / slices;
    return (ks + BK - 1) / BK * BK;
}

hipError_t launch_dgemm_f64_splitk(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha,
                                   const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                                   double* c, int64_t ldc, int slices, double* work, hipStream_t stream,
                                   TileInfo* info) {
    const bool vec = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) % 16 == 0) &&
                     (lda % 2 == 0) && (ldb % 2 == 0);
    const int M = (int)m, N = (int)n, K = (int)k;
    const int kslice = (int)dgemm_splitk_kslice(k, slices);
    slices = (int)((k + kslice - 1) / kslice);
    const int64_t ps = m * n;
    if (info) {
        info->tile_m = BM;
        info->tile_n = BN;
        info->tile_k = BK;
        info->launches += 2;
    }
    const int code = (layout_a ? 4 : 0) | (layout_b ? 2 : 0) | (vec ? 1 : 0);
    hipError_t e;
    mark_main(info, true, stream);
#define JLC_SPLITK_CASE(c_, a_, b_, v_)                                                                         \
    case c_:                                                                                                  \
        e = launch_variant<a_, b_, v_>(M, N, K, 1.0, a, lda, b, ldb, 0.0, work, N, slices, kslice, ps, stream); \
        break;
    switch (code) {
        JLC_SPLITK_CASE(0, false, false, false)
        JLC_SPLITK_CASE(1, false, false, true)
        JLC_SPLITK_CASE(2, false, true, false)
        JLC_SPLITK_CASE(3, false, true, true)
        JLC_SPLITK_CASE(4, true, false, false)
        JLC_SPLITK_CASE(5, true, false, true)
        JLC_SPLITK_CASE(6, true, true, false)
        default: JLC_SPLITK_CASE(7, true, true, true)
    }
#undef JLC_SPLITK_CASE
    mark_main(info, false, stream);
    if (e != hipSuccess) return e;
    int64_t blocks = (m * n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(splitk_reduce, dim3((unsigned)blocks), dim3(256), 0, stream, M, N, slices, work, ps, alpha,
                       beta, c, ldc);
    return hipGetLastError();
}

hipError_t launch_dgemm_f64(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha,
                            const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                            int64_t ldc, hipStream_t stream, TileInfo* info) {
    const bool vec = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) % 16 == 0) &&
                     (lda % 2 == 0) && (ldb % 2 == 0);
    const int M = (int)m, N = (int)n, K = (int)k;
    if (info) {
        info->tile_m = BM;
        info->tile_n = BN;
        info->tile_k = BK;
        info->launches++;
    }
    const int code = (layout_a ? 4 : 0) | (layout_b ? 2 : 0) | (vec ? 1 : 0);
    hipError_t e;
    mark_main(info, true, stream);
    switch (code) {
        case 0: e = launch_variant<false, false, false>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
        case 1: e = launch_variant<false, false, true>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
        case 2: e = launch_variant<false, true, false>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
        case 3: e = launch_variant<false, true, true>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
        case 4: e = launch_variant<true, false, false>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
        case 5: e = launch_variant<true, false, true>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
        case 6: e = launch_variant<true, true, false>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
        default: e = launch_variant<true, true, true>(M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1, K, 0, stream); break;
    }
    mark_main(info, false, stream);
    return e;
}

}  // namespace jlc
