// Internal declarations shared by the libjlc_hip translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jlc {

// Tile shape of a launch, reported back through jlc_stats.
struct TileInfo {
    int tile_m = 0, tile_n = 0, tile_k = 0;
    int launches = 0;
    int slices = 0;
    int bits_a = 0, bits_b = 0;
    // When set, recorded on the launch stream right before / after the dominant kernel.
    hipEvent_t main_begin = nullptr, main_end = nullptr;
    double main_ms = 0.0;
};

inline void mark_main(const TileInfo* info, bool begin, hipStream_t stream) {
    if (!info) return;
    hipEvent_t e = begin ? info->main_begin : info->main_end;
    if (e) (void)hipEventRecord(e, stream);
}

// C = beta*C on an m x n row-major block (beta == 0 writes zeros, C unread).
hipError_t launch_scale(int64_t m, int64_t n, double beta, double* c, int64_t ldc, hipStream_t stream,
                        TileInfo* info);

// Tuned parameters (read by dispatch; written by the autotuner's table).
struct TuneTable {
    double hip_min_flops = 2.0e8;     // host-side offload gate (GemmDispatch reads it)
    int splitk_tiles_per_cu = 2;      // split K until slices * tiles >= this * CUs (0 = never split)
    int splitk_min_kslice = 1024;     // smallest K per slice
    int grouped_waves_per_simd = 0;   // persistent grouped grid (0 = occupancy limit)
    int grouped_xcds = 8;             // grouped tile list dealt per XCD (workgroup b on XCD b % this; 1 = off)
    double crt_bits_offset = 49.4;    // CRT form: norm bits = ceil(offset + log2(k)/2 + slope * max sig)
    double crt_bits_slope = 0.43;
    int crt_waves = 8;                // CRT product kernel: 4 (128 x 128 per wavefront) or 8 (128 x 64)
    int crt_group = 16;               // CRT product kernel: tile rows per column strip of the tile order
    int crt_xcd = 0;                  // CRT product kernel: 1 = keep neighbouring tiles on one XCD
    int splitk_quant = 1;             // split-K counts chosen against wave quantization: the count of a
                                      // few-tile split and a 2-4 split of many-tile grids (0 = off)
    int grouped_small_tile = 16;      // grouped FP64: 16-row / 16-column tiles for m / n up to this (0: off)
    int grouped_chunk_tiles = 64;     // grouped: tiles of one problem dealt to one XCD together (whole tile rows)
    int qr_wave_max_lds = 2560;       // batched QR: size classes up to this many LDS doubles run one
                                      // wavefront per matrix, larger ones qr_block_threads threads
    int qr_block_threads = 256;       // batched QR: threads per matrix above qr_wave_max_lds (256/512/1024)
    int qr_reg_max_m = 192;           // batched QR: matrices with m <= this (at most 192), n <= 128 held
                                      // in registers; larger ones in LDS
    int qr_cqr_min_rows = 256;        // blocked QR: panels of at least this many rows (and 2 * width)
                                      // by CholeskyQR2 + reconstruction (0 = column kernels only)
    int wy_left_wgs_per_cu = 2;       // compact-WY updates (kb <= 64, FP64): partial-W workgroups per CU,
    int wy_right_wgs_per_cu = 2;      // left and right side (0 = the three-GEMM sequence for that side)
};
TuneTable tune_snapshot();

// FP64 MFMA GEMM, device pointers, layouts as in jlc_hip.h. Requires m, n, k > 0 and alpha != 0.
hipError_t launch_dgemm_f64(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha,
                            const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                            int64_t ldc, hipStream_t stream, TileInfo* info);

// Split-K variant: K cut into `slices` pieces whose partial products go to
// work (slices * m * n doubles) and are summed in order into C.
// quant: pick, within that bound, the count whose resident rounds are fullest (see dgemm_f64.hip).
int dgemm_splitk_slices(int64_t m, int64_t n, int64_t k, int device_cus, int tiles_per_cu, int min_kslice,
                        bool quant);
// Split factor (1..4) against wave quantization for grids of more tiles than are resident at once, with
// the partial products bounded by max_partial_bytes; 1 when the split does not pay.
int dgemm_quant_slices(int64_t m, int64_t n, int64_t k, int device_cus, int64_t max_partial_bytes);
hipError_t launch_dgemm_f64_splitk(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha,
                                   const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                                   double* c, int64_t ldc, int slices, double* work, hipStream_t stream,
                                   TileInfo* info);

// y = alpha*A*x + beta*y, A row-major. Requires m > 0; k == 0 or alpha == 0 handled inside.
hipError_t launch_dgemv(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x,
                        int64_t incx, double beta, double* y, int64_t incy, hipStream_t stream, TileInfo* info);

constexpr int JLC_LAYOUT_ROW = 0;
constexpr int JLC_LAYOUT_COL = 1;

// Geometry and workspace of one split-precision (INT8 slice) GEMM.
struct SplitPlan {
    int64_t m = 0, n = 0, k = 0;
    int64_t mpad = 0, npad = 0, kpad = 0;  // padded to the C tile (128 x 64) and the MFMA depth (32)
    int slices = 0;                        // S, 3..9
    int64_t a_slice_bytes = 0;             // bytes of one packed slice of op(A) (mpad * kpad)
    int64_t b_slice_bytes = 0;             // bytes of one packed slice of op(B) (npad * kpad)
    void* a_pack = nullptr;                // S * a_slice_bytes
    void* b_pack = nullptr;                // S * b_slice_bytes
    int* ea = nullptr;                     // mpad row exponents of op(A)
    int* eb = nullptr;                     // npad column exponents of op(B)
    float* spread_a = nullptr;             // mpad
    float* spread_b = nullptr;             // npad
    double* stats_part = nullptr;          // 4 * split_stats_part_elems doubles (column statistics partials)
};

// K ranges of the two-pass column statistics, and the partial-buffer size
// (in 32-byte records) that covers both operands of an m x n x k product.
int split_stats_kparts(int64_t rows, int64_t k);
int64_t split_stats_part_elems(int64_t m, int64_t n, int64_t k);

constexpr int SPLIT_MIN_SLICES = 3;
constexpr int SPLIT_MAX_SLICES = 9;

void split_plan_sizes(int64_t m, int64_t n, int64_t k, int slices, SplitPlan* p);
// Row exponents / spreads of op(A) and column exponents / spreads of op(B).
// Statistics of one operand's rows (kcontig: element (r, k) at x[r*stride + k], else x[k*stride + r]).
hipError_t launch_operand_stats(bool kcontig, const double* x, int64_t stride, int R, int K, double* part, int* e,
                                float* spread, hipStream_t stream);
hipError_t launch_split_stats(int layout_a, int layout_b, const double* a, int64_t lda, const double* b, int64_t ldb,
                              SplitPlan* p, hipStream_t stream);
// Slices both operands and runs the slice-product GEMM with FP64 recombination into C.
hipError_t launch_split_pack_and_gemm(int layout_a, int layout_b, const double* a, int64_t lda, const double* b,
                                      int64_t ldb, const SplitPlan& p, double alpha, double beta, double* c,
                                      int64_t ldc, hipStream_t stream, TileInfo* info);

// One problem of a grouped launch; the same layout as jlc_gemm_desc.
struct GroupedDesc {
    int32_t layout_a, layout_b;
    int32_t m, n, k;
    int32_t flags;
    double alpha, beta;
    const double* a;
    const double* b;
    double* c;
    int64_t lda, ldb, ldc;
};

// Persistent grid size (workgroups) for the grouped kernel on a device with device_cus CUs.
int grouped_grid(int device_cus, int waves_per_simd);
constexpr int GROUPED_WAVES_PER_GROUP = 4;  // wavefronts per grouped workgroup
constexpr uint64_t GROUPED_SKIP = ~0ull;     // tile-list entry that names no tile
constexpr int GROUPED_SHAPE_SHIFT = 30;
// tiles: (problem << 32) | (shape << 30) | (tile_row << 15) | tile_col, where shape bit 1 selects 16-row
// and bit 0 16-column tiles (else 32; the split form always uses 32 x 32); wavefront w of the
// grid takes entries w, w + W, w + 2W, ... (W = grid * GROUPED_WAVES_PER_GROUP) and skips
// GROUPED_SKIP entries.
hipError_t launch_grouped_f64(const GroupedDesc* descs, const uint64_t* tiles, int ntiles, int grid,
                              hipStream_t stream);
// Split-INT8 form of the grouped launch (slices in [SPLIT_MIN_SLICES, SPLIT_MAX_SLICES]): the grid that
// fills the device for that kernel, and the launch over the same tile list.
int grouped_split_grid(int slices, int device_cus, int waves_per_simd);
// Where the INT8 slice form keeps one problem's packed operands in the plan's workspace: fragments of
// 32-row block rb of op(A), k block kb, slice t at 1 KiB block a_off + (rb * kblocks + kb) * S + t
// (kblocks = ceil(k / 32)), of op(B)'s 32-column blocks likewise from b_off; row / column exponents
// at expo[ea_off + i] / expo[eb_off + j] (rows and columns padded to 32).
struct GroupedPack {
    int64_t a_off, b_off, ea_off, eb_off;
};
// Pack jobs: (problem << 32) | (operand << 31) | block, operand 0 = a 32-row block of op(A), 1 = a
// 32-column block of op(B); slice jobs: (job index << 24) | 32-deep k block of that job. The exponent
// launch (one wavefront per job) and the slicing launch (one per slice job) precede the product.
hipError_t launch_grouped_split_i8(int slices, const GroupedDesc* descs, const GroupedPack* packs,
                                   const uint64_t* jobs, int64_t njobs, const uint64_t* sjobs, int64_t nsjobs,
                                   const uint64_t* tiles, int ntiles, int grid, void* frag, int* expo,
                                   hipStream_t stream);

// ---- QR
// Largest m * n of one batched small QR: the matrix tile and the kernel's 2080 bytes of static LDS fit
// the 160 KiB a workgroup may hold.
constexpr int QR_BATCH_MAX_MN = (163840 - 2080) / 8;
struct QrBatchDesc {
    int32_t m, n;
    int64_t lda;
    double* a;    // column-major m x n, factored in place
    double* tau;  // min(m, n)
};
// One blocked-QR panel (column-major m x n at a, lda; m <= 192, n <= 128) factored in place on one
// workgroup by the register kernel of the batched QR (same reflectors and tau as the column kernels).
hipError_t launch_qr_panel_reg(double* a, int64_t lda, int m, int n, double* tau, hipStream_t stream);
// out[0..1] (device memory) = max(0, a[0..m)), max(0, b[0..n)): one workgroup (the count choosers' inputs).
hipError_t launch_max2_nonneg(const float* a, int64_t m, const float* b, int64_t n, float* out, hipStream_t stream);
// Compact-WY update kernels for 1 <= kb <= 64 (wy.hip): side 0 (left) C <- C - V op(T) V^T C,
// side 1 (right) C <- C - C V op(T) V^T. wy_plan sizes the partial-W slices for a grid of about
// device_cus * wgs_per_cu workgroups (c, ldc: the block, for its line alignment); part holds part_elems
// doubles, w2 w2_elems.
struct WyPlan {
    int kb_pad = 0;          // 16 * ceil(kb / 16)
    int slices = 0;          // partial-W slots
    int span = 0;            // rows (left) or columns (right) per slot
    int cshift = 0;          // C's column tiles start this many columns before the block (line alignment)
    int64_t part_elems = 0;
    int64_t w2_elems = 0;
};
WyPlan wy_plan(int side, int64_t m, int64_t n, int64_t kb, const double* c, int64_t ldc, int device_cus,
               int wgs_per_cu);
hipError_t launch_wy_update(int side, int trans_t, int m, int n, int kb, const double* v, int64_t ldv,
                            const double* t, int64_t ldt, double* c, int64_t ldc, const WyPlan& plan, double* part,
                            double* w2, hipStream_t stream);
int qr_panel_chunk();
// Factors panel columns [k0, k1) of the column-major m-row matrix at (in place, tau[k0..k1)), one
// launch per column; part holds 2 * (2G + 1 + 128) doubles, dots 2 * G * 128, xbuf 2 * m
// (G = ceil((m - k0) / qr_panel_chunk())); k1 - k0 <= 128.
hipError_t launch_qr_panel_fused(double* at, int64_t lda, int m, int k0, int k1, double* tau, double* part,
                                 double* dots, double* xbuf, hipStream_t stream);
// CholeskyQR2 panel with Householder reconstruction (qr_cqr.hip), b <= 128:
// pass 1: R1 = chol(G), X1 = R1^-1, *flag = 1 on breakdown (else 0);
// pass 2: R2 = chol(G), r_out = R2 * r1, X2 = R2^-1, *flag |= 2 on breakdown or max|G - I| > 2^-8.
hipError_t launch_cqr_chol(const double* g, int b, int pass, const double* r1, double* r_out, double* x_out,
                           int* flag, hipStream_t stream);
// G = X^T X (b x b row-major, both triangles) for the column-major m x b matrix X (ld ldx), b <= 128:
// partial 16 x 16 tiles per row range in part (gram_work_doubles(m) doubles), then a fixed-order sum.
int gram_parts(int64_t m);
size_t gram_work_doubles(int64_t m);
hipError_t launch_gram(const double* x, int64_t ldx, int m, int b, double* part, double* g, hipStream_t stream);
// T (b x b row-major with row stride ldt, upper) of the compact-WY product of b reflectors from G = V^T V
// and tau, as the inverse of diag(1/tau) + strictly-upper(G) (reflectors with tau = 0 take no part),
// b <= 128.
hipError_t launch_wy_t(const double* g, int b, const double* tau, double* T, int ldt, hipStream_t stream);
// G = X^T Y (bx x by row-major, row stride ldg) for column-major m x bx and m x by matrices X, Y (ld ldx,
// ldy), bx, by <= 128; part holds gram_cross_work_doubles(m) doubles.
size_t gram_cross_work_doubles(int64_t m);
hipError_t launch_gram_cross(const double* x, int64_t ldx, int bx, const double* y, int64_t ldy, int by, int m,
                             double* part, double* g, int ldg, hipStream_t stream);
// Modified LU of Q's top b x b block (column-major, ldq): Y1 and D R into the panel's top block ap (lda),
// tau, ui = U^-1 (row-major b x b), spec (b ints) = candidates for the tau = 0 case.
hipError_t launch_cqr_lu(const double* q, int64_t ldq, int b, const double* r, double* ap, int64_t lda, double* tau,
                         double* ui, int* spec, hipStream_t stream);
// Columns flagged in spec whose Y2 part (rows b..mp-1 of ap) is exactly zero: tau = 0, R row negated.
hipError_t launch_cqr_special(double* ap, int64_t lda, int mp, int b, double* tau, const int* spec,
                              hipStream_t stream);
hipError_t launch_build_vc(const double* at, int64_t lda, int m, int k0, int nb, double* vc, hipStream_t stream);
// The same into rows of stride ldv, each starting with pad zeros (V of a later panel inside a combined
// V of several panels).
hipError_t launch_build_vc(const double* at, int64_t lda, int m, int k0, int nb, double* vc, int64_t ldv, int pad,
                           hipStream_t stream);
hipError_t launch_build_t(const double* g, int nb, const double* tau, double* T, hipStream_t stream);
hipError_t launch_extract_r(const double* at, int64_t lda, int n, double* r, hipStream_t stream);
hipError_t launch_fill_eye(double* x, int64_t ld, int rows, int cols, double diag, hipStream_t stream);
hipError_t launch_sumsq_partial(const double* x, int64_t ld, int rows, int cols, double* part, int nparts,
                                hipStream_t stream);
// threads = 64 (one wavefront per matrix), 256, 512 or 1024; max_lds = the largest
// qr_batched_lds_entries of the launch (doubles).
hipError_t launch_qr_batched_small(const QrBatchDesc* descs, int count, int max_lds, int threads,
                                   hipStream_t stream);
// Batched QR with each matrix in registers: m <= 8 * rows_per_thread (4, 8, ..., 24), n <= max_n <= 128
// for every descriptor of the launch.
hipError_t launch_qr_batched_reg(const QrBatchDesc* descs, int count, int rows_per_thread, int max_n,
                                 hipStream_t stream);
// LDS doubles one matrix of the batched QR takes: leading dimension m | 1 (odd: conflict-free
// column-parallel reads) while that fits QR_ODD_LD_MAX entries, else m; times n.
int qr_batched_lds_entries(int m, int n);
constexpr int QR_ODD_LD_MAX = (163840 - 2048 - 16 * 8) / 8;  // 160 KiB less the largest kernel's static LDS
// Upper trapezoids of stacked row blocks into a column-major matrix (TSQR's stack of R factors, or R).
hipError_t launch_copy_upper_blocks(const double* src, int64_t lds, int64_t stride, int per, int rows_total, int cols,
                                    double* dst, int64_t ldd, hipStream_t stream);
// Rows of a stacked matrix spread back to row blocks of `stride` rows (per rows each, the rest zero).
hipError_t launch_spread_blocks(const double* src, int64_t lds, int per, int stride, int rows_total, int cols,
                                int rows_src, double* dst, int64_t ldd, hipStream_t stream);
// X <- Q X for many small factored blocks (m <= 256, nx <= 256, m * nx <= QR_APPLY_MAX_MNX).
struct QrApplyDesc {
    int32_t m, nx, k, pad;
    int64_t lda, ldx;
    const double* a;
    const double* tau;
    double* x;
};
constexpr int QR_APPLY_MAX_MNX = (163840 - 4096 - 256) / 8;
hipError_t launch_qr_batched_apply_q(const QrApplyDesc* descs, int count, int max_mnx, hipStream_t stream);

// BF16-slice variant (8-bit digits, slices 3..8).
// CRT split form: N pairwise coprime moduli <= 256, residues in int8.
constexpr int CRT_MAX_MODULI = 20;
constexpr int CRT_MIN_MODULI = 2;
constexpr int CRT_MAX_BITS = 65;  // norm bits per operand (|a'| < 2^66: three signed 22-bit chunks)
extern const int CRT_MODULI[CRT_MAX_MODULI];

// Per-modulus constants and the reconstruction tables (passed by value to kernels).
struct CrtTables {
    int nmod = 0;
    int pad_ = 0;
    double mod[CRT_MAX_MODULI];
    double inv[CRT_MAX_MODULI];     // 1/m rounded
    int kdig[CRT_MAX_MODULI][3];    // 2^(22c+8i) mod m (c, i < 3), symmetric, as packed int8 (see crt_pack)
    float invf[CRT_MAX_MODULI];     // 1/m in FP32
    int pad2_ = 0;
    double hi_lim[CRT_MAX_MODULI];  // symmetric residue range [lo_lim, hi_lim]
    double lo_lim[CRT_MAX_MODULI];
    double w[CRT_MAX_MODULI][4];    // 40-bit limbs of W_t = (M/m_t) * ((M/m_t)^-1 mod m_t)
    double mlimb[4];                // 40-bit limbs of M
    double finv[4];                 // 2^(40 j) / M
};

struct CrtPlan {
    int64_t m = 0, n = 0, k = 0;
    int64_t mpad = 0, npad = 0, kpad = 0;  // 256 x 256 C tile, 64-deep K stage
    int nmod = 0, pa = 0, pb = 0;          // moduli, bits of op(A) rows / op(B) columns
    int64_t a_plane_bytes = 0, b_plane_bytes = 0, c_plane_bytes = 0;
    void* a_pack = nullptr;   // nmod planes of op(A) residues (fragment order)
    void* b_pack = nullptr;   // nmod planes of op(B) residues
    void* c_res = nullptr;    // nmod planes of C' residues (accumulator order)
    int* ea = nullptr;        // mpad row norm exponents of op(A)
    int* eb = nullptr;        // npad column norm exponents of op(B)
};

double crt_log2_product(int nmod);
// Split statistics (sp.ea/eb, sp.spread_a/b), norm exponents ta/tb (2^t at or
// above each row / column 2-norm) and the chooser statistic sig per row /
// column; uses sp.stats_part as scratch.
hipError_t launch_crt_stats_norms(int layout_a, int layout_b, const double* a, int64_t lda, const double* b,
                                  int64_t ldb, const SplitPlan& sp, int* ta, int* tb, float* sig_a, float* sig_b,
                                  hipStream_t stream);
void crt_tables(int nmod, CrtTables* T);
void crt_plan_sizes(int64_t m, int64_t n, int64_t k, int nmod, CrtPlan* p);
hipError_t launch_crt(int layout_a, int layout_b, const double* a, int64_t lda, const double* b, int64_t ldb,
                      const CrtPlan& p, double alpha, double beta, double* c, int64_t ldc, hipStream_t stream,
                      TileInfo* info);

constexpr int SPLIT_BF16_MAX_SLICES = 8;
void split_bf16_plan_sizes(int64_t m, int64_t n, int64_t k, int slices, SplitPlan* p);
int64_t split_bf16_chunk_k(int slices, int64_t k);
hipError_t launch_split_bf16(int layout_a, int layout_b, const double* a, int64_t lda, const double* b, int64_t ldb,
                             const SplitPlan& p, double alpha, double beta, double* c, int64_t ldc,
                             hipStream_t stream, TileInfo* info);

// Device-to-device copy kernel (bytes may be any size).
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

}  // namespace jlc
