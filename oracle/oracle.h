/*
 * CPU oracle for the DGEMM path: TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker. Only tests, the smoke check and the benchmark's
 * CPU-baseline leg may load it; the product path (libjlc_hip.so) never does.
 *
 * Contents, each restating the reference algorithm cited beside it:
 *   - oracle_seeded_array      GemmReference.seededArray      (src/test/java/net/faulj/compute/GemmReference.java:14-25)
 *   - oracle_gemm_*            GemmReference.gemm* layouts    (GemmReference.java:31-112)
 *   - oracle_parity            GemmReference.assertParity     (GemmReference.java:148-173)
 *   - oracle_cpu_abs_tol /     cpuAbsTolerance / cudaAbsTolerance (GemmReference.java:175-181)
 *     oracle_gpu_abs_tol
 *   - oracle_parallel_micro    OptimizedBLAS3 PARALLEL_MICRO  (src/main/java/net/faulj/compute/OptimizedBLAS3.java:295-385,
 *                              GemmDispatch.java:118-149, :183-208, PackingUtils.java:37-56, :121-133, :159-183,
 *                              MicroKernel.java:61-271)
 *   - oracle_matvec            SpecializedKernels.matvec      (src/main/java/net/faulj/compute/SpecializedKernels.java:27-71)
 *   - oracle_dd_*              double-double dot products (Knuth TwoSum, FMA TwoProduct; Ogita-Rump-Oishi 2005).
 *                              No reference counterpart: it is the accuracy oracle for the split-precision form.
 *   - oracle_crt_*             CRT split form (modular INT8 products + Chinese remainder reconstruction).
 *                              No reference counterpart: the north star's split-precision component.
 *   - oracle_qr_*              HouseholderQR blocked restatement (src/main/java/net/faulj/decomposition/qr/HouseholderQR.java)
 *
 * Parity pins (the checks that make a plausible mistake fail) live in tests/test_oracle.py.
 * SplitMix64 bit-equality with java.util.SplittableRandom is "parity unpinned": no JDK exists
 * here to produce a vector; parity never depends on it because the oracle and the kernel read
 * the same arrays.
 */
#ifndef JLC_ORACLE_H
#define JLC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout codes shared with the product ABI: element (i,p) of op(A) is
 * a[i*lda+p] for ROW and a[p*lda+i] for COL; element (p,j) of op(B) is
 * b[p*ldb+j] for ROW and b[j*ldb+p] for COL. */
enum { ORACLE_ROW = 0, ORACLE_COL = 1 };

void oracle_seeded_array(double* out, int64_t len, int64_t seed);
void oracle_uniform_array(double* out, int64_t len, int64_t seed, double origin, double bound);
uint64_t oracle_splitmix_next(uint64_t* state);

/* General restatement: c (strided, row-major, ldc) <- alpha*op(A)op(B) + beta*c,
 * sum accumulated left-to-right with separate multiply and add (no FMA). */
void oracle_gemm_general(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                         double* c, int64_t ldc, int threads);

/* Reference-named entry points (same argument order and meaning as GemmReference). */
void oracle_gemm(const double* a, const double* b, double* c, int m, int k, int n, double alpha, double beta);
void oracle_gemm_strided(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                         double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta);
void oracle_gemm_strided_trans_a(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                                 double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta);
void oracle_gemm_strided_col_major_a(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                                     double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta);
void oracle_gemm_strided_col_major_b(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                                     double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta);

void oracle_parity(const double* expected, const double* actual, int64_t len, double* max_abs_err, double* rel_err);
double oracle_cpu_abs_tol(const double* expected, int64_t len, int64_t k);
double oracle_gpu_abs_tol(const double* expected, int64_t len, int64_t k);

/* PARALLEL_MICRO restatement (the timed CPU baseline). Row-major A (lda), B (ldb), C (ldc). */
void oracle_parallel_micro(int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                           const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int max_threads,
                           int* threads_used);
/* Block sizes the restatement uses: mc, kc, nc, mr, nr. */
void oracle_block_sizes(int* out5);

void oracle_matvec(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x,
                   double beta, double* y);

/* Double-double dot products for selected entries (i,j) of op(A)op(B): hi+lo ~ exact sum
 * with relative error ~ k*2^-104. abs_out (optional) receives sum |a||b| (rounded up slightly). */
void oracle_dd_entries(int a_layout, int b_layout, int64_t k, const double* a, int64_t lda, const double* b,
                       int64_t ldb, int64_t count, const int64_t* ii, const int64_t* jj, double* hi, double* lo,
                       double* abs_out, int threads);
/* Double-double full GEMM (small sizes): hi/lo arrays m*n row-major of alpha*AB + beta*C evaluated in dd. */
void oracle_dd_gemm(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                    int64_t lda, const double* b, int64_t ldb, double beta, const double* c, int64_t ldc,
                    double* hi, double* lo, int threads);

/* Split-precision (INT8 slice) GEMM restatement: the exact algorithm of the
 * split form (row/column power-of-two scaling, S truncated 7-bit slices,
 * triangular pair set t+u <= S-1, per-diagonal exact integer sums in K chunks of
 * chunk_k, FP64 recombination from the smallest diagonal up, alpha/beta via fma).
 * No reference counterpart: the north star's new component (Ozaki et al. 2012,
 * Ootomo-Ozaki-Yokota 2024). Returns the exponent marker used for Inf/NaN rows. */
int oracle_split_exponent(const double* x, int64_t stride, int64_t count);
void oracle_split_slices(const double* x, int64_t stride, int64_t count, int e, int slices, int8_t* out,
                         int64_t out_stride);
void oracle_split_gemm(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                       int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int slices,
                       int64_t chunk_k, int threads);
int64_t oracle_split_chunk_k(int slices, int64_t k);
/* Digit width 7 (INT8 form) or 8 (BF16 form); oracle_split_gemm is the 7-bit case. */
void oracle_split_gemm_digits(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha,
                              const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                              int64_t ldc, int slices, int bits, int64_t chunk_k, int threads);
int64_t oracle_split_bf16_chunk_k(int slices, int64_t k);

/* CRT split form (modular integer products recombined by the Chinese remainder
 * theorem; Ozaki-Uchino-Imamura 2025). Moduli: a fixed list of 20 pairwise
 * coprime integers <= 256. nmod moduli; row i of op(A) is scaled by
 * 2^(pa - t_i) with 2^t_i >= ||row||_2 (oracle_crt_norm_exponent), column j of
 * op(B) by 2^(pb - t_j); requires
 * log2(prod m) >= log2((2^pa + sqrt(k)/2)(2^pb + sqrt(k)/2)) + 2. */
int oracle_crt_norm_exponent(const double* x, int64_t stride, int64_t count, int e);
int oracle_crt_max_moduli(void);
int oracle_crt_modulus(int t);
double oracle_crt_log2_product(int nmod);
void oracle_crt_tables(int nmod, double* w, double* mlimb, double* finv);
void oracle_crt_residues(const double* x, int64_t stride, int64_t count, int e, int p, int nmod, int8_t* out,
                         int64_t out_stride);
double oracle_crt_reconstruct(const int* r, int nmod, const double* w, const double* mlimb, const double* finv);
void oracle_crt_gemm(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                     int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int nmod, int pa,
                     int pb, int threads);

/* Blocked Householder QR restatement (column-major in/out). */
/* Compact-WY block update of a row-major m x n block C (ldc) by V (row-major, ldv; m x kb for side 0,
 * n x kb for side 1) and the kb x kb upper-triangular T (row-major, ldt; trans_t selects T^T):
 *   side 0 (left):  W = V^T C, W2 = op(T) W, C = C - V W2
 *   side 1 (right): W = C V,   W2 = W op(T), C = C - W2 V^T
 * each product a GemmReference triple loop; see oracle.c for the reference routines restated. */
void oracle_wy_apply(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                     const double* t, int64_t ldt, double* c, int64_t ldc, int threads);

int oracle_qr_factor(int64_t m, int64_t n, double* a_colmajor, int64_t lda, double* tau, int64_t nb);
/* TallSkinnyQR.decompose restated (oracle.c): column-major A, Q (m x min(m,n)), R (min(m,n) x n). */
void oracle_tsqr(int64_t m, int64_t n, const double* a, int64_t lda, double* q, int64_t ldq, double* r, int64_t ldr);
void oracle_qr_form_q(int64_t m, int64_t n, const double* qr_colmajor, int64_t lda, const double* tau,
                      double* q_colmajor, int64_t ldq);

#ifdef __cplusplus
}
#endif
#endif
