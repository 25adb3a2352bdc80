// Compact-WY block update kernels for panels of at most 64 reflectors (gfx950, FP64).
//
// jlc_wy_apply_dev's update of a row-major m x n block,
//   left:  C <- C - V op(T) (V^T C)     V m x kb,
//   right: C <- C - ((C V) op(T)) V^T   V n x kb,
// has one dimension of kb in every product. As macro-tile GEMMs the 128-wide tiles would be mostly
// padding (kb = 32 fills a quarter of one) and the K = kb update loop too short to hide a K tile's
// latency, so the update runs as three kernels that stream C instead:
//   1. partial W: each workgroup reduces a slice of C's rows (left, W = V^T C, KB x n) or columns
//      (right, W = C V, m x KB) into its own slot of a partial buffer; C is read once, coalesced;
//   2. W2: one thread per column (left) or row (right) and quarter of W2's KB entries sums the slots
//      in slot order and applies op(T) from LDS, writing W2 with k contiguous (left: W2^T, n x KB;
//      right: W2, m x KB);
//   3. update: C -= A B^T with A and B k-contiguous (left: A = V, B = W2^T; right: A = W2, B = V).
//      Each 64 x 32 tile of C is loaded into the MFMA accumulators, updated over K = KB and stored:
//      C is read and written once.
// KB = 16 * ceil(kb / 16). Every product runs on v_mfma_f64_16x16x4. Reflector columns past kb and
// rows / columns past the block are selected to zero after a clamped (in-range) load, so any m, n >= 1
// and 1 <= kb <= 64 are valid and no value outside the operation (an Inf past kb) reaches a sum.
//
// v_mfma_f64_16x16x4 lane layout (lane l, g = l / 16, c16 = l % 16): lane l supplies A(c16, g) and
// B(g, c16) of a 16 x 4 by 4 x 16 step; element q of its accumulator holds D(g + 4q, c16).
#include <algorithm>

#include "internal.h"
#include "jlc_hip.h"

namespace jlc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 4-row k steps per unrolled group of the left partial kernel: two groups of operands stay in registers
template <int KF>
constexpr int wy_u() { return KF <= 2 ? 8 : 4; }
constexpr int WY_ROWS = 32;  // row-slice granularity: a multiple of every group (4 * wy_u rows)

// ---- 1L: P[s][i][col] = sum over rows r of slice s of V[r][i] C[r][col] (i < KB, zero for i >= kb).
// Workgroup: 4 wavefronts, 32 columns each (two 16-column fragments), all over the slice's rows; the
// four read the same V rows, which the L1 serves after the first. Rows stream in groups of WY_U k steps,
// the next group's loads issued before the current group's MFMAs.
// Loads are left raw and V's columns past kb are selected to zero where the MFMAs use them: a select
// at load time would make the prefetch of the next group wait for its loads before the current group's
// MFMAs could start.
template <int KF, int WY_U>
__device__ __forceinline__ void wy_left_load(const double* __restrict__ V, int64_t ldv, const double* __restrict__ C,
                                             int64_t ldc, int row0, int g, const int (&vc)[KF], int ca, int cb,
                                             double (&a)[WY_U][KF], double (&b)[WY_U][2]) {
#pragma unroll
    for (int u = 0; u < WY_U; u++) {
        const int64_t row = row0 + 4 * u + g;
        const double* vr = V + row * ldv;
        const double* cr = C + row * ldc;
#pragma unroll
        for (int i = 0; i < KF; i++) a[u][i] = vr[vc[i]];
        b[u][0] = cr[ca];
        b[u][1] = cr[cb];
    }
}

template <int KF, int WY_U>
__device__ __forceinline__ void wy_left_mfma(const double (&a)[WY_U][KF], const double (&b)[WY_U][2],
                                             const bool (&vok)[KF], d4 (&acc)[KF][2]) {
#pragma unroll
    for (int u = 0; u < WY_U; u++)
#pragma unroll
        for (int i = 0; i < KF; i++) {
            const double av = vok[i] ? a[u][i] : 0.0;
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = mfma(av, b[u][j], acc[i][j]);
        }
}

template <int KF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void wy_partial_left(int m, int n, int kb, const double* __restrict__ V,
                                                       int64_t ldv, const double* __restrict__ C, int64_t ldc,
                                                       double* __restrict__ P, int rows_per_slice, int cshift) {
    constexpr int KB = 16 * KF, WY_U = wy_u<KF>();
    const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
    const int col0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 32 - cshift;
    const int r0 = blockIdx.y * rows_per_slice;
    const int r1 = min(m, r0 + rows_per_slice);
    // columns outside [0, n) are clamped into it: their sums are never stored, and columns do not mix
    const int ca = max(0, min(col0 + c16, n - 1)), cb = max(0, min(col0 + 16 + c16, n - 1));
    int vc[KF];
    bool vok[KF];
#pragma unroll
    for (int i = 0; i < KF; i++) {
        vc[i] = min(16 * i + c16, kb - 1);
        vok[i] = 16 * i + c16 < kb;
    }
    d4 acc[KF][2];
#pragma unroll
    for (int i = 0; i < KF; i++) acc[i][0] = acc[i][1] = d4{0.0, 0.0, 0.0, 0.0};

    // Steady state in one basic block: the next group's loads, this group's MFMAs, then the prefetched
    // operands become the current ones. (Conditional prefetches let the compiler hoist a group's MFMAs
    // above the loads meant to overlap them.)
    const int groups = (r1 - r0) / (4 * WY_U);
    double a0[WY_U][KF], b0[WY_U][2], a1[WY_U][KF], b1[WY_U][2];
    if (groups > 0) {
        wy_left_load<KF, WY_U>(V, ldv, C, ldc, r0, g, vc, ca, cb, a0, b0);
        for (int q = 1; q < groups; q++) {
            wy_left_load<KF, WY_U>(V, ldv, C, ldc, r0 + q * 4 * WY_U, g, vc, ca, cb, a1, b1);
            wy_left_mfma<KF, WY_U>(a0, b0, vok, acc);
#pragma unroll
            for (int u = 0; u < WY_U; u++) {
#pragma unroll
                for (int i = 0; i < KF; i++) a0[u][i] = a1[u][i];
                b0[u][0] = b1[u][0];
                b0[u][1] = b1[u][1];
            }
        }
        wy_left_mfma<KF, WY_U>(a0, b0, vok, acc);
    }
    // tail: fewer than 4 * WY_U rows as one group whose loads all go out together (row r1 - 1 is loaded
    // for the rows past r1, which are selected to zero)
    const int rt = r0 + groups * 4 * WY_U;
    if (rt < r1) {
        double a[WY_U][KF], b[WY_U][2];
#pragma unroll
        for (int u = 0; u < WY_U; u++) {
            const int64_t rr = min(rt + 4 * u + g, r1 - 1);
#pragma unroll
            for (int i = 0; i < KF; i++) a[u][i] = V[rr * ldv + vc[i]];
            b[u][0] = C[rr * ldc + ca];
            b[u][1] = C[rr * ldc + cb];
        }
#pragma unroll
        for (int u = 0; u < WY_U; u++) {
            const bool ok = rt + 4 * u + g < r1;
#pragma unroll
            for (int i = 0; i < KF; i++) {
                const double av = ok && vok[i] ? a[u][i] : 0.0;
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = mfma(av, ok ? b[u][j] : 0.0, acc[i][j]);
            }
        }
    }
    double* ps = P + (int64_t)blockIdx.y * KB * n;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int col = col0 + 16 * j + c16;
        if (col < 0 || col >= n) continue;
#pragma unroll
        for (int i = 0; i < KF; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) ps[(int64_t)(16 * i + g + 4 * q) * n + col] = acc[i][j][q];
    }
}

// ---- 1R: P[s][row][i] = sum over columns c of slice s of C[row][c] V[c][i].
// Workgroup: 64 rows. Each 64-column chunk of those rows is staged in LDS with coalesced loads (one row
// of the chunk per wavefront load); wavefront w takes the chunk's k steps 4w..4w+3 for all four row
// fragments, reading V's rows for its own k directly (each V element read by one wavefront). At the end
// the four wavefronts' sums are added in wavefront order through LDS.
constexpr int WY_CH = 64;          // columns per staged chunk
constexpr int WY_SR = WY_CH + 4;   // LDS row stride (doubles): fragment reads of 16 rows hit distinct banks

template <int KF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void wy_partial_right(int m, int n, int kb, const double* __restrict__ C,
                                                        int64_t ldc, const double* __restrict__ V, int64_t ldv,
                                                        double* __restrict__ P, int cols_per_slice, int cshift) {
    constexpr int KB = 16 * KF;
    __shared__ double cs[64 * WY_SR];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, g = lane >> 4, c16 = lane & 15;
    const int row0 = blockIdx.x * 64;
    // the slice's chunks start at kbase (shifted so that chunk rows begin on 128-byte lines); the
    // columns it owns are [k0, k1)
    const int kbase = blockIdx.y * cols_per_slice - cshift;
    const int k0 = max(0, kbase);
    const int k1 = min(n, kbase + cols_per_slice);
    int vc[KF];
    bool vok[KF];
#pragma unroll
    for (int i = 0; i < KF; i++) {
        vc[i] = min(16 * i + c16, kb - 1);
        vok[i] = 16 * i + c16 < kb;
    }
    d4 acc[4][KF];
#pragma unroll
    for (int f = 0; f < 4; f++)
#pragma unroll
        for (int i = 0; i < KF; i++) acc[f][i] = d4{0.0, 0.0, 0.0, 0.0};

    // staging: thread t loads column (t & 63) of the chunk for rows (t >> 6) + 4 e, e < 16; rows past m are
    // clamped (their sums are never stored, rows do not mix). Loads stay raw; columns outside [k0, k1) are
    // selected to zero when the chunk is written to LDS, and V's rows outside [k0, k1) or columns past kb
    // at the MFMA, so that no select waits on a prefetch.
    double x[16];
    auto load_chunk = [&](int kc) {
        const int64_t cc = max(k0, min(kc + lane, k1 - 1));
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int64_t row = min(row0 + w + 4 * e, m - 1);
            x[e] = C[row * ldc + cc];
        }
    };
    // V rows of this wavefront's k steps (16w .. 16w + 15 of the chunk)
    auto load_v = [&](int kc, double (&b)[4][KF]) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int64_t vrr = max(k0, min(kc + 16 * w + 4 * s + g, k1 - 1));
#pragma unroll
            for (int i = 0; i < KF; i++) b[s][i] = V[vrr * ldv + vc[i]];
        }
    };
    auto stage_chunk = [&](int kc) {
        __syncthreads();  // the previous chunk's fragment reads are done
        const bool cok = kc + lane >= k0 && kc + lane < k1;
#pragma unroll
        for (int e = 0; e < 16; e++) cs[(w + 4 * e) * WY_SR + lane] = cok ? x[e] : 0.0;
        __syncthreads();
    };
    auto mfma_chunk = [&](int kc, const double (&bc)[4][KF]) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int kk = 16 * w + 4 * s;  // this wavefront's k step within the chunk
            const bool rok = kc + kk + g >= k0 && kc + kk + g < k1;
            double b[KF];
#pragma unroll
            for (int i = 0; i < KF; i++) b[i] = rok && vok[i] ? bc[s][i] : 0.0;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const double a = cs[(16 * f + c16) * WY_SR + kk + g];
#pragma unroll
                for (int i = 0; i < KF; i++) acc[f][i] = mfma(a, b[i], acc[f][i]);
            }
        }
    };
    // Loads complete in issue order, so each chunk's V rows are issued before the next chunk's C
    // prefetch: the MFMAs of a chunk then never wait on that prefetch.
    if (KF <= 2) {
        // V rows one chunk ahead as well, in two buffers used alternately by name
        double vb0[4][KF], vb1[4][KF];
        if (k0 < k1) {
            load_v(kbase, vb0);
            load_chunk(kbase);
        }
        for (int kc = kbase; kc < k1; kc += 2 * WY_CH) {
            stage_chunk(kc);
            if (kc + WY_CH < k1) {
                load_v(kc + WY_CH, vb1);
                load_chunk(kc + WY_CH);
            }
            mfma_chunk(kc, vb0);
            if (kc + WY_CH < k1) {
                stage_chunk(kc + WY_CH);
                if (kc + 2 * WY_CH < k1) {
                    load_v(kc + 2 * WY_CH, vb0);
                    load_chunk(kc + 2 * WY_CH);
                }
                mfma_chunk(kc + WY_CH, vb1);
            }
        }
    } else {
        // wider panels: the accumulators take 4 * KF registers of every lane, so V rows are loaded for the
        // current chunk only (after its staging, before the next chunk's C) and two workgroups fit a CU
        if (k0 < k1) load_chunk(kbase);
        for (int kc = kbase; kc < k1; kc += WY_CH) {
            stage_chunk(kc);
            double vb[4][KF];
            load_v(kc, vb);
            if (kc + WY_CH < k1) load_chunk(kc + WY_CH);
            mfma_chunk(kc, vb);
        }
    }
    // wavefront-ordered sum of the four partial 64 x KB blocks through LDS (64 * KB <= 64 * WY_SR doubles
    // for KB <= 64), then the slice's block to P
    __syncthreads();
    for (int src = 0; src < 4; src++) {
        if (w == src) {
#pragma unroll
            for (int f = 0; f < 4; f++)
#pragma unroll
                for (int i = 0; i < KF; i++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        double* p = &cs[(16 * f + g + 4 * q) * KB + 16 * i + c16];
                        *p = src == 0 ? acc[f][i][q] : *p + acc[f][i][q];
                    }
        }
        __syncthreads();
    }
    double* ps = P + ((int64_t)blockIdx.y * m + row0) * KB;
    for (int e = t; e < 64 * KB; e += 256) {
        if (row0 + e / KB < m) ps[e] = cs[e];
    }
}

// ---- 2: W2 from the partial slots. A workgroup takes 16 columns (left) or rows (right) of W: its 256
// threads first sum the slots in slot order for 16 * KB entries (coalesced: 16 columns of a slot row, or
// KB entries of a slot's row), park the sums in LDS, then compute the 16 * KB entries of W2 against
// op(T) staged in LDS (row stride KB + 1: threads of one wavefront read distinct banks).
constexpr int WY_W2 = 16;

// w[u] = p[u][0] + p[u][stride] + ... over `slices` slots, added in slot order, for NE entries at once:
// the loads of all entries go out eight slots at a time (slots past the last read the last one and add
// zero), so the sums wait for one load latency per eight slots.
template <int NE>
__device__ __forceinline__ void wy_slot_sums(const double* const (&p)[NE], int64_t stride, int slices,
                                             double (&w)[NE]) {
#pragma unroll
    for (int u = 0; u < NE; u++) w[u] = 0.0;
    for (int s0 = 0; s0 < slices; s0 += 8) {
        double x[NE][8];
#pragma unroll
        for (int u = 0; u < NE; u++)
#pragma unroll
            for (int j = 0; j < 8; j++) x[u][j] = p[u][(int64_t)min(s0 + j, slices - 1) * stride];
#pragma unroll
        for (int u = 0; u < NE; u++)
#pragma unroll
            for (int j = 0; j < 8; j++) w[u] += s0 + j < slices ? x[u][j] : 0.0;
    }
}

// 2L: W2T[col][r] = sum_i opT(r, i) W[i][col], W[i][col] = sum_s P[s][i][col];
// opT(r, i) = T[r][i] (trans 0) or T[i][r] (trans 1), zero past kb.
template <int KF>
__global__ __launch_bounds__(256) void wy_w2_left(int n, int kb, int slices, const double* __restrict__ P,
                                                  const double* __restrict__ T, int64_t ldt, int trans,
                                                  double* __restrict__ W2T) {
    constexpr int KB = 16 * KF, ST = KB + 1;
    __shared__ double ts[KB * ST];      // ts[r * ST + i] = opT(r, i)
    __shared__ double ws[KB * WY_W2];   // ws[i * 16 + cl] = W[i][c0 + cl]
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * WY_W2;
    // op(T) staged with every load issued before the first LDS write (clamped, then selected)
    constexpr int TPER = (KB * KB + 255) / 256;
    double tv[TPER];
#pragma unroll
    for (int u = 0; u < TPER; u++) {
        const int e = min(t + 256 * u, KB * KB - 1), r = min(e / KB, kb - 1), i = min(e % KB, kb - 1);
        tv[u] = trans ? T[(int64_t)i * ldt + r] : T[(int64_t)r * ldt + i];
    }
    // the slot sums next (their loads overlap op(T)'s), then both go to LDS
    constexpr int WPER = KB * WY_W2 / 256;
    double wsum[WPER];
    const double* pe[WPER];
#pragma unroll
    for (int u = 0; u < WPER; u++) {
        const int e = t + 256 * u;
        pe[u] = P + (int64_t)(e / WY_W2) * n + min(c0 + e % WY_W2, n - 1);
    }
    wy_slot_sums<WPER>(pe, (int64_t)KB * n, slices, wsum);
#pragma unroll
    for (int u = 0; u < TPER; u++) {
        const int e = t + 256 * u, r = e / KB, i = e % KB;
        if (e < KB * KB) ts[r * ST + i] = (r < kb && i < kb) ? tv[u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < WPER; u++) ws[t + 256 * u] = wsum[u];
    __syncthreads();
    for (int e = t; e < KB * WY_W2; e += 256) {
        const int cl = e / KB, r = e % KB;
        if (c0 + cl >= n) break;  // e grows with cl: every later entry is past n too
        double acc = 0.0;
#pragma unroll 16
        for (int i = 0; i < KB; i++) acc = fma(ts[r * ST + i], ws[i * WY_W2 + cl], acc);
        W2T[(int64_t)(c0 + cl) * KB + r] = acc;
    }
}

// 2R: W2[row][c] = sum_i W[row][i] opT(i, c), W[row][i] = sum_s P[s][row][i];
// opT(i, c) = T[i][c] (trans 0) or T[c][i] (trans 1), zero past kb.
template <int KF>
__global__ __launch_bounds__(256) void wy_w2_right(int m, int kb, int slices, const double* __restrict__ P,
                                                   const double* __restrict__ T, int64_t ldt, int trans,
                                                   double* __restrict__ W2) {
    constexpr int KB = 16 * KF, ST = KB + 1;
    __shared__ double ts[KB * ST];      // ts[c * ST + i] = opT(i, c)
    __shared__ double ws[WY_W2 * KB];   // ws[rl * KB + i] = W[r0 + rl][i]
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * WY_W2;
    constexpr int TPER = (KB * KB + 255) / 256;
    double tv[TPER];
#pragma unroll
    for (int u = 0; u < TPER; u++) {
        const int e = min(t + 256 * u, KB * KB - 1), c = min(e / KB, kb - 1), i = min(e % KB, kb - 1);
        tv[u] = trans ? T[(int64_t)c * ldt + i] : T[(int64_t)i * ldt + c];
    }
    constexpr int WPER = KB * WY_W2 / 256;
    double wsum[WPER];
    const double* pe[WPER];
#pragma unroll
    for (int u = 0; u < WPER; u++) {
        const int e = t + 256 * u;
        pe[u] = P + (int64_t)min(r0 + e / KB, m - 1) * KB + e % KB;
    }
    wy_slot_sums<WPER>(pe, (int64_t)m * KB, slices, wsum);
#pragma unroll
    for (int u = 0; u < TPER; u++) {
        const int e = t + 256 * u, c = e / KB, i = e % KB;
        if (e < KB * KB) ts[c * ST + i] = (c < kb && i < kb) ? tv[u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < WPER; u++) ws[t + 256 * u] = wsum[u];
    __syncthreads();
    for (int e = t; e < KB * WY_W2; e += 256) {
        const int rl = e / KB, c = e % KB;
        if (r0 + rl >= m) break;  // e grows with rl: every later entry is past m too
        double acc = 0.0;
#pragma unroll 16
        for (int i = 0; i < KB; i++) acc = fma(ws[rl * KB + i], ts[c * ST + i], acc);
        W2[(int64_t)(r0 + rl) * KB + c] = acc;
    }
}

// ---- 3: C[r][c] -= sum_{k < kb} A[r][k] B[c][k] (A m x kb, B n x kb, both k-contiguous).
// Workgroup: one WY_TM x WY_TN tile, four wavefronts in 2 x 2 (each two 16 x 16 fragments at 64 x 32). K
// runs in chunks of KCH <= 32 staged in LDS ((WY_TM + WY_TN) x (KCH + 4) doubles). The first chunk's
// operand rows (-A and B, k-contiguous, L2 hits) are loaded before the tile's C entries go into the
// accumulators, so their LDS writes wait for the operands alone while C's HBM reads are in flight; each
// later chunk's operands are loaded during the previous chunk's MFMAs. Stores last.
template <int KF>
constexpr int wy_kch() { return KF <= 2 ? 16 * KF : (KF == 4 ? 32 : 16); }

// Rows and columns of C per update workgroup. Measured per kb = 32 / 64 update (average over the n = 4096
// sweeps): 64 x 128 28.2 / 33.2 us, 64 x 64 24.5 / 31.1, 64 x 32 24.1 / 29.8, 32 x 32 24.1 / 32.3.
constexpr int WY_TM = 64, WY_TN = 32;

template <int KF>
__global__ __launch_bounds__(256) void wy_update(int m, int n, int kb, const double* __restrict__ A, int64_t lda,
                                                 const double* __restrict__ B, int64_t ldb, double* __restrict__ C,
                                                 int64_t ldc, int tiles_n, int cshift) {
    constexpr int KB = 16 * KF, KCH = wy_kch<KF>(), NCH = KB / KCH, SK = KCH + 4;
    constexpr int PA = WY_TM * KCH / 256, PB = WY_TN * KCH / 256, WM = WY_TM / 2, WN = WY_TN / 2;
    constexpr int NI = WM / 16, NJ = WN / 16;
    __shared__ double as[WY_TM * SK];
    __shared__ double bs[WY_TN * SK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1, g = lane >> 4, c16 = lane & 15;
    const int r0 = (blockIdx.x / tiles_n) * WY_TM, c0 = (blockIdx.x % tiles_n) * WY_TN - cshift;
    const bool interior = r0 + WY_TM <= m && c0 >= 0 && c0 + WY_TN <= n;
    double xa[PA], xb[PB];
    auto load_ops = [&](int k0) {
#pragma unroll
        for (int u = 0; u < PA; u++) {
            const int e = t + 256 * u;
            const int64_t kc = min(k0 + e % KCH, kb - 1);
            xa[u] = A[(int64_t)min(r0 + e / KCH, m - 1) * lda + kc];
        }
#pragma unroll
        for (int u = 0; u < PB; u++) {
            const int e = t + 256 * u;
            const int64_t kc = min(k0 + e % KCH, kb - 1);
            xb[u] = B[(int64_t)max(0, min(c0 + e / KCH, n - 1)) * ldb + kc];
        }
    };
    auto store_ops = [&](int k0) {
#pragma unroll
        for (int u = 0; u < PA; u++) {
            const int e = t + 256 * u;
            const int rr = e / KCH, k = e % KCH;
            as[rr * SK + k] = k0 + k < kb ? -xa[u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PB; u++) {
            const int e = t + 256 * u;
            const int rr = e / KCH, k = e % KCH;
            bs[rr * SK + k] = k0 + k < kb ? xb[u] : 0.0;
        }
    };
    load_ops(0);
    d4 acc[NI][NJ];
#pragma unroll
    for (int i = 0; i < NI; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int64_t col = max(0, min(c0 + wn * WN + 16 * j + c16, n - 1));
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t row = min(r0 + wm * WM + 16 * i + g + 4 * q, m - 1);
                acc[i][j][q] = C[row * ldc + col];
            }
        }
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
        if (ch) __syncthreads();
        store_ops(ch * KCH);
        __syncthreads();
        if (ch + 1 < NCH) load_ops((ch + 1) * KCH);
#pragma unroll
        for (int kk = 0; kk < KCH; kk += 4) {
            double a[NI], b[NJ];
#pragma unroll
            for (int i = 0; i < NI; i++) a[i] = as[(wm * WM + 16 * i + c16) * SK + kk + g];
#pragma unroll
            for (int j = 0; j < NJ; j++) b[j] = bs[(wn * WN + 16 * j + c16) * SK + kk + g];
#pragma unroll
            for (int i = 0; i < NI; i++)
#pragma unroll
                for (int j = 0; j < NJ; j++) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < NI; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int col = c0 + wn * WN + 16 * j + c16;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = r0 + wm * WM + 16 * i + g + 4 * q;
                if (interior || (row < m && col >= 0 && col < n)) C[(int64_t)row * ldc + col] = acc[i][j][q];
            }
        }
}

template <int KF>
hipError_t wy_launch(int side, int trans_t, int m, int n, int kb, const double* v, int64_t ldv, const double* t,
                     int64_t ldt, double* c, int64_t ldc, const WyPlan& p, double* part, double* w2,
                     hipStream_t stream) {
    constexpr int KB = 16 * KF;
    const int tiles_m = (m + WY_TM - 1) / WY_TM, tiles_n = (n + p.cshift + WY_TN - 1) / WY_TN;
    if (side == JLC_LEFT) {
        hipLaunchKernelGGL(wy_partial_left<KF>, dim3((unsigned)((n + p.cshift + 127) / 128), (unsigned)p.slices),
                           dim3(256), 0, stream, m, n, kb, v, ldv, c, ldc, part, p.span, p.cshift);
        hipLaunchKernelGGL(wy_w2_left<KF>, dim3((unsigned)((n + WY_W2 - 1) / WY_W2)), dim3(256), 0, stream, n, kb,
                           p.slices, part, t, ldt, trans_t, w2);
        hipLaunchKernelGGL(wy_update<KF>, dim3((unsigned)(tiles_m * tiles_n)), dim3(256), 0, stream, m, n, kb, v,
                           ldv, w2, (int64_t)KB, c, ldc, tiles_n, p.cshift);
    } else {
        hipLaunchKernelGGL(wy_partial_right<KF>, dim3((unsigned)((m + 63) / 64), (unsigned)p.slices), dim3(256), 0, stream,
                           m, n, kb, c, ldc, v, ldv, part, p.span, p.cshift);
        hipLaunchKernelGGL(wy_w2_right<KF>, dim3((unsigned)((m + WY_W2 - 1) / WY_W2)), dim3(256), 0, stream, m, kb,
                           p.slices, part, t, ldt, trans_t, w2);
        hipLaunchKernelGGL(wy_update<KF>, dim3((unsigned)(tiles_m * tiles_n)), dim3(256), 0, stream, m, n, kb, w2,
                           (int64_t)KB, v, ldv, c, ldc, tiles_n, p.cshift);
    }
    return hipGetLastError();
}

}  // namespace

WyPlan wy_plan(int side, int64_t m, int64_t n, int64_t kb, const double* c, int64_t ldc, int device_cus,
               int wgs_per_cu) {
    WyPlan p;
    p.kb_pad = (int)(16 * ((kb + 15) / 16));
    // C's column tiling starts cshift columns before the block when that puts every row's tiles on
    // 128-byte lines (ldc a multiple of 16 doubles): a block at an odd column (H[:, k+1:]) would otherwise
    // split each 128-byte row segment over two lines
    const uintptr_t addr = reinterpret_cast<uintptr_t>(c);
    p.cshift = (ldc % 16 == 0 && addr % 8 == 0) ? (int)((addr / 8) % 16) : 0;
    const int64_t target = (int64_t)std::max(1, device_cus) * std::max(1, wgs_per_cu);
    if (side == JLC_LEFT) {
        // slices of rows, a multiple of WY_ROWS each and at least 2 * WY_ROWS (a short slice costs the
        // W2 kernel one more slot to sum and hides no latency of its own)
        const int64_t blocks = (n + p.cshift + 127) / 128;
        int64_t s = std::max<int64_t>(1, (target + blocks - 1) / blocks);
        s = std::max<int64_t>(1, std::min<int64_t>(s, m / (2 * WY_ROWS)));
        int64_t span = (m + s - 1) / s;
        span = (span + WY_ROWS - 1) / WY_ROWS * WY_ROWS;
        p.span = (int)span;
        p.slices = (int)((m + span - 1) / span);
        p.part_elems = (int64_t)p.slices * p.kb_pad * n;
        p.w2_elems = n * p.kb_pad;
    } else {
        // slices of shifted columns, a multiple of one staged chunk each (every slice holds a column:
        // span >= 64 > cshift)
        const int64_t blocks = (m + 63) / 64;
        const int64_t cols = n + p.cshift;
        int64_t s = std::max<int64_t>(1, (target + blocks - 1) / blocks);
        s = std::min<int64_t>(s, (cols + WY_CH - 1) / WY_CH);
        int64_t span = (cols + s - 1) / s;
        span = (span + WY_CH - 1) / WY_CH * WY_CH;
        p.span = (int)span;
        p.slices = (int)((cols + span - 1) / span);
        p.part_elems = (int64_t)p.slices * m * p.kb_pad;
        p.w2_elems = m * p.kb_pad;
    }
    return p;
}

hipError_t launch_wy_update(int side, int trans_t, int m, int n, int kb, const double* v, int64_t ldv,
                            const double* t, int64_t ldt, double* c, int64_t ldc, const WyPlan& plan, double* part,
                            double* w2, hipStream_t stream) {
    if (m < 1 || n < 1 || kb < 1 || kb > 64 || plan.kb_pad != 16 * ((kb + 15) / 16) || plan.slices < 1)
        return hipErrorInvalidValue;
    switch (plan.kb_pad / 16) {
        case 1: return wy_launch<1>(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, plan, part, w2, stream);
        case 2: return wy_launch<2>(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, plan, part, w2, stream);
        case 3: return wy_launch<3>(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, plan, part, w2, stream);
        default: return wy_launch<4>(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, plan, part, w2, stream);
    }
}

}  // namespace jlc
