/*
 * libjlc_hip: MI355X (gfx950) DGEMM engine behind a plain C ABI.
 *
 * This is the native side of the "HIP" kernel choice that plugs into the
 * reference's GEMM dispatch. Every entry point returns an int status
 * (JLC_OK == 0), never throws and never aborts; jlc_last_error() returns a
 * thread-local message for the last failing call on the calling thread.
 * All entry points are thread-safe: each host thread gets its own HIP stream
 * and its own staging buffers, so concurrent callers (the reference calls
 * its GEMM from many ForkJoin threads) never share mutable state. One thread's
 * device-pointer calls may run on different streams: the split-form, split-K
 * and compact-WY workspace the thread's calls share is ordered across streams
 * (a call on another stream than the workspace's last user waits for that
 * user's launches on the device). A stream passed to the library must stay
 * alive until the work queued on it has finished.
 *
 * Index conventions (shared by every GEMM entry point). C is always row-major:
 * element (i,j) of C is C[i*ldc + j]. For op(A) (m x k) and op(B) (k x n):
 *   layout_a == JLC_ROW : op(A)(i,p) = A[i*lda + p]
 *   layout_a == JLC_COL : op(A)(i,p) = A[p*lda + i]
 *   layout_b == JLC_ROW : op(B)(p,j) = B[p*ldb + j]
 *   layout_b == JLC_COL : op(B)(p,j) = B[j*ldb + p]
 * Semantics are BLAS semantics: C = alpha*op(A)*op(B) + beta*C where beta == 0
 * never reads C (NaN in C does not survive) and alpha == 0 or k == 0 never
 * reads A or B (C = beta*C). This is the behaviour of the reference's
 * MICROKERNEL / SMALL / MATVEC regimes (PackingUtils.scaleCPanel,
 * src/main/java/net/faulj/compute/PackingUtils.java:159-183;
 * OptimizedBLAS3.java:190-192; SpecializedKernels.java:32-49). It differs from
 * the reference's OptimizedBLAS3 TINY CPU path, which keeps a NaN-seeded C at
 * beta == 0 (src/test/java/net/faulj/compute/GemmNumericEdgeTest.java:44-47):
 * the host-side dispatch only reaches this library above its offload gate
 * (hip_min_flops, 14,155,776 flops in the shipped table), far above the TINY
 * regime (under 512 flops, GemmDispatch.java:30-33), so calls routed by the
 * default gate keep the reference's TINY semantics on the CPU.
 */
#ifndef JLC_HIP_H
#define JLC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define JLC_OK 0
#define JLC_ERR_INVALID_ARG 1   /* null pointer, negative size, bad ld, bad layout/mode */
#define JLC_ERR_NO_DEVICE 2     /* no usable gfx950 device */
#define JLC_ERR_DISABLED 3      /* JLC_HIP_ENABLED=0 in the environment */
#define JLC_ERR_HIP 4           /* a HIP runtime call failed */
#define JLC_ERR_ALLOC 5         /* device allocation failed */
#define JLC_ERR_UNSUPPORTED 6   /* valid request this build does not serve */
#define JLC_ERR_ACCURACY 7      /* split precision cannot meet the requested error bound */

/* Operand layouts. */
#define JLC_ROW 0
#define JLC_COL 1

/* Arithmetic modes. The low byte selects the form; for the split forms bits
 * 8..15 carry the slice count, or for the CRT form the moduli count (0 = choose
 * from the inputs' exponent spread). JLC_MODE_SPLIT_CRT: rows of op(A) and
 * columns of op(B) scaled by powers of two at their 2-norms, modular INT8
 * products recombined by the Chinese remainder theorem (2..20 moduli, k below
 * 2^22). In automatic mode a spread that needs more bits than 20 moduli hold
 * (CRT), or more slices than the form has (9 INT8, 8 BF16), runs the FP64 MFMA
 * kernel instead (jlc_stats.kernel says which ran). */
#define JLC_MODE_FP64 0
#define JLC_MODE_SPLIT_INT8 1
#define JLC_MODE_SPLIT_BF16 2
#define JLC_MODE_SPLIT_CRT 3
#define JLC_MODE_SLICES(mode_, s_) ((mode_) | ((s_) << 8))

/* Kernel identifiers reported in jlc_stats.kernel. */
#define JLC_KERNEL_NONE 0
#define JLC_KERNEL_SCALE 1      /* alpha == 0 or k == 0: C = beta*C */
#define JLC_KERNEL_MFMA_F64 2   /* FP64 MFMA macro-tile kernel */
#define JLC_KERNEL_GEMV 3       /* wavefront-reduction matvec */
#define JLC_KERNEL_SPLIT_INT8 4 /* INT8 slice products, FP64 recombination */
#define JLC_KERNEL_SPLIT_BF16 5 /* BF16 slice products, FP64 recombination */
#define JLC_KERNEL_GROUPED 6    /* persistent grouped kernel */
#define JLC_KERNEL_QR 7         /* Householder QR (panel kernels + compact-WY GEMMs) */
#define JLC_KERNEL_SPLIT_CRT 8  /* modular INT8 products, CRT reconstruction in FP64 */
#define JLC_KERNEL_WY 9         /* compact-WY kernels (partial W, W2, update), kb <= 64 */

typedef struct jlc_stats {
    double kernel_ms;   /* device time of the call's kernels (HIP events) */
    double total_ms;    /* host wall time of the call, transfers included */
    int32_t kernel;     /* JLC_KERNEL_* that produced C */
    int32_t tile_m;     /* macro tile chosen */
    int32_t tile_n;
    int32_t tile_k;
    int32_t slices;     /* split-precision slice count / CRT moduli count (0 for FP64) */
    int32_t launches;   /* kernel launches issued */
    int32_t bits_a;     /* CRT form: norm bits of op(A) rows, ||a'|| <= 2^bits_a + sqrt(k)/2 (0 otherwise) */
    int32_t bits_b;     /* CRT form: norm bits of op(B) columns */
    double main_ms;     /* device time of the dominant kernel alone (the GEMM / GEMV launch; for
                           jlc_dgeqrf_dev the trailing-update products C -= W2^T V^T, summed over
                           the panels) */
} jlc_stats;

/* ---- device detection and lifetime (analogue of CudaSupport / CudaContext,
 * src/main/java/net/faulj/compute/CudaSupport.java:34-123, CudaContext.java:11-55).
 * Honours JLC_HIP_ENABLED (0 disables), JLC_HIP_FORCE (1: skip the gfx950
 * check) and HIP_VISIBLE_DEVICES. */
int jlc_hip_probe(int* device_count, char* name, size_t name_len);
int jlc_hip_init(int device);
int jlc_hip_shutdown(void);
const char* jlc_last_error(void);
const char* jlc_version(void);

/* ---- GEMM on host memory. Replaces the body of CudaGemm.gemm
 * (src/main/java/net/faulj/compute/CudaGemm.java:53-135) and serves every
 * layout of BLAS3Kernels.gemmStrided / gemmStridedTransA / gemmStridedColMajorA /
 * gemmStridedColMajorB (BLAS3Kernels.java:244, :272, :284, :296) and
 * OptimizedBLAS3.gemm / gemmStrided (OptimizedBLAS3.java:61, :135).
 * Operands are staged into per-thread device buffers that persist across
 * calls; C is updated in place. stats may be NULL. */
int jlc_dgemm(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
              int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int mode,
              jlc_stats* stats);

/* ---- GEMM on device memory (device residency twin of jlc_dgemm).
 * stream is a hipStream_t (NULL = the calling thread's library stream). The
 * call is asynchronous unless stats is non-NULL (then it waits for the
 * kernels to finish to report their time). */
int jlc_dgemm_dev(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                  int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int mode,
                  void* stream, jlc_stats* stats);

/* ---- FP64 GEMM on device memory that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as jlc_dgemm_dev with mode JLC_MODE_FP64 (the
 * same kernels, bit for bit, with one exception: a shape with more output
 * tiles than run at once whose last round jlc_dgemm_dev fills by splitting K,
 * splitk_quant, is only split here when the partials fit the array below, and
 * otherwise runs unsplit, its sums then in another order); stream NULL is the
 * null stream; always asynchronous. The split-K partial products that
 * jlc_dgemm_dev keeps in a
 * per-thread buffer go to a 256 MiB device array the library holds on each
 * device for the life of the process (its address is resolved once per device
 * by jlc_hip_init, which this call runs for device 0 if nothing has). A shape
 * whose partials exceed it fails with JLC_ERR_UNSUPPORTED before anything is
 * launched (the shipped tuning never makes one). Every call on a device
 * shares that array: calls that overlap on one device (other threads or
 * streams) race on it. */
int jlc_dgemm_dev_noalloc(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha,
                          const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                          int64_t ldc, void* stream);

/* ---- compact-WY block update of a strided sub-matrix: the three-GEMM
 * sequences of the reference's blocked reductions in one call. C is a
 * row-major m x n block (ldc) of a larger matrix; T is kb x kb upper
 * triangular, row-major (ldt); op(T) = T (trans_t == 0) or T^T (trans_t == 1).
 *   side == JLC_LEFT  (V row-major m x kb, ldv):  C <- C - V op(T) (V^T C)
 *     replaces the body of BlockedHessenbergQR.applyBlockLeft
 *     (src/main/java/net/faulj/eigen/qr/BlockedHessenbergQR.java:277-306, trans_t = 1) and the left
 *     block update of BlockedBidiagonalization
 *     (src/main/java/net/faulj/decomposition/bidiagonal/BlockedBidiagonalization.java:379-401, trans_t = 0);
 *   side == JLC_RIGHT (V row-major n x kb, ldv):  C <- C - ((C V) op(T)) V^T
 *     replaces applyBlockRight (BlockedHessenbergQR.java:308-345, the colStart = k + 1 form its
 *     callers use, trans_t = 0) and the right block update of BlockedBidiagonalization (:581-603).
 * In JLC_MODE_FP64 with kb <= 64 the update runs three kernels (partial sums of V^T C or C V over
 * slices of C, W2 = op(T) W, then C -= A B^T tile by tile; stats->kernel = JLC_KERNEL_WY); otherwise
 * the two large products run in `mode` (as jlc_dgemm_dev) and the kb x kb product in FP64. The W and
 * W2 intermediates live in the calling thread's workspace. kb == 0 leaves C unchanged. The device
 * form is asynchronous unless stats is non-NULL; the host form stages only V, T and C's block. */
#define JLC_LEFT 0
#define JLC_RIGHT 1
int jlc_wy_apply_dev(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                     const double* t, int64_t ldt, double* c, int64_t ldc, int mode, void* stream, jlc_stats* stats);
int jlc_wy_apply(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                 const double* t, int64_t ldt, double* c, int64_t ldc, int mode, jlc_stats* stats);

/* ---- MATVEC: y = alpha*A*x + beta*y, A row-major m x k (lda). Replaces
 * SpecializedKernels.matvec (src/main/java/net/faulj/compute/SpecializedKernels.java:27-71),
 * the n == 1 regime of GemmDispatch.selectKernel (GemmDispatch.java:80-82).
 * incx / incy are element strides of x and y (the B and C columns of an
 * n == 1 GEMM have strides ldb and ldc). */
int jlc_dgemv(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x, int64_t incx,
              double beta, double* y, int64_t incy, jlc_stats* stats);
int jlc_dgemv_dev(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x,
                  int64_t incx, double beta, double* y, int64_t incy, void* stream, jlc_stats* stats);

/* ---- grouped (batched, uneven) GEMM: thousands of independent problems in
 * one persistent launch. Serves the reference's TINY / SMALL regimes in bulk
 * (SpecializedKernels.tinyGemm / smallGemm,
 * src/main/java/net/faulj/compute/SpecializedKernels.java:79-158) and batched
 * QR trailing updates. All pointers in a descriptor are device pointers. */
typedef struct jlc_gemm_desc {
    int32_t layout_a, layout_b; /* JLC_ROW / JLC_COL */
    int32_t m, n, k;
    int32_t flags;              /* reserved, 0 */
    double alpha, beta;
    const double* a;
    const double* b;
    double* c;                  /* row-major, ldc >= n */
    int64_t lda, ldb, ldc;
} jlc_gemm_desc;

typedef struct jlc_grouped_plan jlc_grouped_plan;
/* Validates the descriptors, builds the tile list (longest K first, each
 * problem's tiles dealt to one XCD) and uploads both to the device. mode is
 * JLC_MODE_FP64, or JLC_MODE_SPLIT_INT8 with an explicit slice count in [3, 9]
 * (JLC_MODE_SLICES): every problem then computed exactly as jlc_dgemm's INT8
 * slice form computes it with that count, its operands sliced once per run into
 * a plan-owned workspace by a pack launch ahead of the product launch (S KiB per
 * 32 x 32 block of op(A) and op(B); no-counterpart extension: the
 * reference has no batched GEMM entry point; its TINY / SMALL kernels,
 * SpecializedKernels.java:79-158, serve one problem per call). */
int jlc_grouped_plan_create(int64_t count, const jlc_gemm_desc* descs, int mode, jlc_grouped_plan** plan);
/* One persistent launch over every tile of the plan (the slice form: the pack
 * launch, then the product launch). */
int jlc_grouped_plan_run(jlc_grouped_plan* plan, void* stream, jlc_stats* stats);
/* Captures the launch into a HIP graph (once) and replays it on stream. */
int jlc_grouped_plan_replay(jlc_grouped_plan* plan, void* stream, jlc_stats* stats);
int jlc_grouped_plan_destroy(jlc_grouped_plan* plan);
/* Create + run + destroy. */
int jlc_dgemm_grouped_dev(int64_t count, const jlc_gemm_desc* descs, int mode, void* stream, jlc_stats* stats);

/* ---- blocked Householder QR (HouseholderQR.decompose / factorizeBlocked,
 * src/main/java/net/faulj/decomposition/qr/HouseholderQR.java:52-209).
 * a is column-major m x n (element (i,j) at a[j*lda + i], the reference's AT
 * layout); on return R is on and above the diagonal and the Householder vectors
 * (v0 = 1 implicit) below it, tau[min(m,n)] their scalars, as computeHouseholder
 * (:219-291) stores them. Panels of nb <= 128 columns are factored in FP64 on
 * the GPU; the compact-WY trailing updates (:167-209) run on the GEMM entry
 * points in `mode` (JLC_MODE_FP64, JLC_MODE_SPLIT_INT8[, slices] or
 * JLC_MODE_SPLIT_CRT[, moduli]); T is FP64. A panel of at least the tuning
 * table's qr_cqr_min_rows rows and twice its width is factored by CholeskyQR2
 * with Householder reconstruction (the same reflectors, tau and R in exact
 * arithmetic; a panel whose Cholesky breaks down or loses orthogonality runs
 * the column kernels instead); a panel of at most qr_reg_max_m rows and 128
 * columns that does not take it is factored column by column on one workgroup,
 * others by the column kernels; stats->tile_n counts the CholeskyQR2 panels,
 * stats->tile_k the panels factored column by column. */
int jlc_dgeqrf_dev(int64_t m, int64_t n, double* a, int64_t lda, double* tau, int64_t nb, int mode, void* stream,
                   jlc_stats* stats);
/* Thin Q (column-major m x n, ldq) from the factors (buildQ, HouseholderQR.java:390-444), n <= m. */
int jlc_dorgqr_dev(int64_t m, int64_t n, const double* a, int64_t lda, const double* tau, int64_t nb, double* q,
                   int64_t ldq, void* stream);
/* ||A - QR||_F / ||A||_F and ||Q^T Q - I||_F for a factorisation of a_orig (the
 * reference's QR acceptance metrics, src/test/java/net/faulj/stress/HouseholderQRTest.java:26-121). */
int jlc_dqr_errors_dev(int64_t m, int64_t n, const double* a_orig, int64_t lda_orig, const double* qr, int64_t lda,
                       const double* tau, int64_t nb, double* recon_err, double* orth_err);

/* ---- tall-skinny QR (TallSkinnyQR.decompose, which CommunicationAvoidingQR.decompose delegates to;
 * src/main/java/net/faulj/decomposition/qr/TallSkinnyQR.java:22-90, CommunicationAvoidingQR.java:18-20).
 * k = min(m, n). m <= n: the thin Householder QR of A. m > n: row blocks of max(n, 64) rows, a thin
 * QR of each (one batched launch while a block fits in LDS), a QR of the R factors stacked in block
 * order, and each block's rows of Q as its block Q times its rows of the stack's Q (applied as
 * reflectors, one batched launch). Outputs the explicit thin Q (m x k) and R (k x n, upper
 * trapezoidal), with the reference's reflector convention.
 * jlc_dtsqr_dev: column-major device arrays (a is not modified); jlc_dtsqr: row-major host arrays,
 * the reference's Matrix storage (lda >= n, ldq >= k, ldr >= n). Synchronous. */
int jlc_dtsqr_dev(int64_t m, int64_t n, const double* a, int64_t lda, double* q, int64_t ldq, double* r, int64_t ldr,
                  void* stream, jlc_stats* stats);
int jlc_dtsqr(int64_t m, int64_t n, const double* a, int64_t lda, double* q, int64_t ldq, double* r, int64_t ldr,
              jlc_stats* stats);

/* Batched small QR: each matrix (m*n <= 20220, n <= 256) factored by one
 * wavefront in LDS, unblocked, same reflector convention (the HouseholderQR
 * column sequence, src/main/java/net/faulj/decomposition/qr/HouseholderQR.java:219-291).
 * No counterpart in the reference, which factors one matrix per call. */
typedef struct jlc_qr_desc {
    int32_t m, n;
    int64_t lda;
    double* a;   /* column-major, device */
    double* tau; /* min(m, n), device */
} jlc_qr_desc;
typedef struct jlc_qr_batched_plan jlc_qr_batched_plan;
int jlc_qr_batched_plan_create(int64_t count, const jlc_qr_desc* descs, jlc_qr_batched_plan** plan);
/* Captures the single launch into a HIP graph (once) and replays it. */
int jlc_qr_batched_plan_replay(jlc_qr_batched_plan* plan, void* stream, jlc_stats* stats);
int jlc_qr_batched_plan_destroy(jlc_qr_batched_plan* plan);
int jlc_dgeqrf_batched_dev(int64_t count, const jlc_qr_desc* descs, void* stream, jlc_stats* stats);

/* ---- tuning table: the autotuner's choices, read by dispatch. Keys:
 * hip_min_flops (offload gate, the HIP analogue of GemmDispatch.CUDA_THRESHOLD,
 * src/main/java/net/faulj/compute/GemmDispatch.java:32), splitk_tiles_per_cu,
 * splitk_min_kslice, grouped_waves_per_simd, grouped_xcds (tile lists dealt per XCD;
 * 1 = plain order), grouped_small_tile (grouped FP64 problems with m or n up to this run 16-row or
 * 16-column tiles; 0 = always 32 x 32), splitk_quant (1: split K by 2-4 when the
 * last round of a many-tile grid would run mostly empty), grouped_chunk_tiles (tiles of one problem dealt to one XCD
 * together), crt_bits_offset, crt_bits_slope, crt_waves, crt_group, crt_xcd,
 * qr_wave_max_lds (batched QR size classes up to this many LDS doubles run one wavefront per matrix).
 * Text file, one "key value" per line.
 * On first use the library loads $JLC_TUNE_TABLE, or else the table shipped
 * beside it (<libdir>/../tuning/mi355x.tune); jlc_tune_load replaces it later. */
int jlc_tune_load(const char* path);
int jlc_tune_write(const char* path);
int jlc_tune_get(const char* key, double* value);
int jlc_tune_set(const char* key, double value);

/* ---- device memory helpers for callers that keep operands resident. */
int jlc_dev_alloc(void** ptr, size_t bytes);
int jlc_dev_free(void* ptr);
int jlc_memcpy_h2d(void* dst, const void* src, size_t bytes);
int jlc_memcpy_d2h(void* dst, const void* src, size_t bytes);
int jlc_dev_synchronize(void);
/* Device-to-device copy by a kernel on stream (NULL = the thread's stream), asynchronous. */
int jlc_dev_copy(void* dst, const void* src, size_t bytes, void* stream);
/* The calling thread's library stream (what a NULL stream argument means). */
int jlc_stream_get(void** stream);
int jlc_stream_synchronize(void* stream);
/* A new non-blocking stream (default priority) on the library's device, for
 * callers that queue work on several streams; jlc_stream_destroy releases it
 * (the thread's library stream cannot be destroyed). No counterpart in the
 * reference, whose CudaGemm uses the default stream (CudaGemm.java:85-112). */
int jlc_stream_create(void** stream);
int jlc_stream_destroy(void* stream);

/* HIP-event timers for callers that time work on a stream (benchmarks). */
typedef struct jlc_timer jlc_timer;
int jlc_timer_create(jlc_timer** timer);
int jlc_timer_start(jlc_timer* timer, void* stream);
/* Records the end event on stream, waits for it and returns start->end in ms. */
int jlc_timer_stop(jlc_timer* timer, void* stream, double* elapsed_ms);
int jlc_timer_destroy(jlc_timer* timer);

#ifdef __cplusplus
}
#endif
#endif
