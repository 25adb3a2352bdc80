/*
 * CPU oracle for the DGEMM path: TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Built with -ffp-contract=off so that "a*b + c" is two roundings, as in Java.
 * Every function cites the reference code it restates; paths are relative to
 * the reference repository root.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <immintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Seeded inputs: java.util.SplittableRandom (SplitMix64, Steele-Lea-Flood 2014)
 * as used by GemmReference.seededArray (src/test/java/net/faulj/compute/GemmReference.java:14-25).
 * nextLong = mix64(seed += GOLDEN_GAMMA); nextDouble = (nextLong >>> 11) * 2^-53;
 * nextDouble(origin, bound) = r*(bound-origin)+origin, clamped below bound. */
static const uint64_t GOLDEN_GAMMA = 0x9e3779b97f4a7c15ULL;

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t oracle_splitmix_next(uint64_t* state) {
    *state += GOLDEN_GAMMA;
    return mix64(*state);
}

static double next_double(uint64_t* state, double origin, double bound) {
    double r = (double)(oracle_splitmix_next(state) >> 11) * 0x1.0p-53;
    r = r * (bound - origin) + origin;
    if (r >= bound) r = nextafter(bound, -INFINITY);
    return r;
}

void oracle_uniform_array(double* out, int64_t len, int64_t seed, double origin, double bound) {
    uint64_t s = (uint64_t)seed;
    for (int64_t i = 0; i < len; i++) out[i] = next_double(&s, origin, bound);
}

void oracle_seeded_array(double* out, int64_t len, int64_t seed) {
    uint64_t s = (uint64_t)seed;
    for (int64_t i = 0; i < len; i++) {
        double v = next_double(&s, -1.0, 1.0);
        out[i] = (i % 11 == 0) ? 0.0 : v;
    }
}

/* ------------------------------------------------------------------------- */
/* GemmReference triple loops (GemmReference.java:31-112): for each (i,j),
 * sum += a*b over p in increasing order (separate rounding), then
 * out = alpha*sum + beta*out. Rows are independent, so splitting rows over
 * threads leaves every result bit-identical to the serial loop. */
void oracle_gemm_general(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                         double* c, int64_t ldc, int threads) {
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t i = 0; i < m; i++) {
        for (int64_t j = 0; j < n; j++) {
            double sum = 0.0;
            for (int64_t p = 0; p < k; p++) {
                double av = (a_layout == ORACLE_ROW) ? a[i * lda + p] : a[p * lda + i];
                double bv = (b_layout == ORACLE_ROW) ? b[p * ldb + j] : b[j * ldb + p];
                sum += av * bv;
            }
            double* out = &c[i * ldc + j];
            *out = alpha * sum + beta * *out;
        }
    }
}

/* The compact-WY updates of the reference's blocked reductions: the three strided GEMMs of
 * BlockedHessenbergQR.applyBlockLeft (src/main/java/net/faulj/eigen/qr/BlockedHessenbergQR.java:277-306:
 * W = V^T C by gemmStridedTransA, W2 = T^T W by gemmStridedTransA, C = C - V W2 by gemmStrided) and of
 * applyBlockRight (:308-345, called with colStart = k + 1: W = C V, W2 = W T by gemmStrided, C = C - W2 V^T by
 * gemmStridedColMajorB), and the left / right block updates of BlockedBidiagonalization
 * (src/main/java/net/faulj/decomposition/bidiagonal/BlockedBidiagonalization.java:379-401 with W2 = T W,
 * :581-603 with W2 = W T). W and W2 start as zero arrays, as the reference allocates them. */
void oracle_wy_apply(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                     const double* t, int64_t ldt, double* c, int64_t ldc, int threads) {
    if (m <= 0 || n <= 0 || kb <= 0) return;
    const int64_t wn = side == 0 ? kb * n : m * kb;
    double* w = (double*)calloc((size_t)wn, sizeof(double));
    double* w2 = (double*)calloc((size_t)wn, sizeof(double));
    if (!w || !w2) {
        free(w);
        free(w2);
        return;
    }
    const int lt = trans_t ? ORACLE_COL : ORACLE_ROW;
    if (side == 0) {
        oracle_gemm_general(ORACLE_COL, ORACLE_ROW, kb, n, m, 1.0, v, ldv, c, ldc, 0.0, w, n, threads);
        oracle_gemm_general(lt, ORACLE_ROW, kb, n, kb, 1.0, t, ldt, w, n, 0.0, w2, n, threads);
        oracle_gemm_general(ORACLE_ROW, ORACLE_ROW, m, n, kb, -1.0, v, ldv, w2, n, 1.0, c, ldc, threads);
    } else {
        oracle_gemm_general(ORACLE_ROW, ORACLE_ROW, m, kb, n, 1.0, c, ldc, v, ldv, 0.0, w, kb, threads);
        oracle_gemm_general(ORACLE_ROW, lt, m, kb, kb, 1.0, w, kb, t, ldt, 0.0, w2, kb, threads);
        oracle_gemm_general(ORACLE_ROW, ORACLE_COL, m, n, kb, -1.0, w2, kb, v, ldv, 1.0, c, ldc, threads);
    }
    free(w);
    free(w2);
}

/* GemmReference.gemm (GemmReference.java:31-44): contiguous row-major, C updated in place here. */
void oracle_gemm(const double* a, const double* b, double* c, int m, int k, int n, double alpha, double beta) {
    oracle_gemm_general(ORACLE_ROW, ORACLE_ROW, m, n, k, alpha, a, k, b, n, beta, c, n, 1);
}

/* GemmReference.gemmStrided (GemmReference.java:46-61): a[aOff+i*lda+p], b[bOff+p*ldb+j]. */
void oracle_gemm_strided(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                         double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta) {
    oracle_gemm_general(ORACLE_ROW, ORACLE_ROW, m, n, k, alpha, a + aOff, lda, b + bOff, ldb, beta, c + cOff, ldc, 1);
}

/* GemmReference.gemmStridedTransA (GemmReference.java:63-78): C is k x n, inner dimension m,
 * a[aOff+p*lda+i] — the same index map as a column-major op(A) of k rows. */
void oracle_gemm_strided_trans_a(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                                 double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta) {
    oracle_gemm_general(ORACLE_COL, ORACLE_ROW, k, n, m, alpha, a + aOff, lda, b + bOff, ldb, beta, c + cOff, ldc, 1);
}

/* GemmReference.gemmStridedColMajorA (GemmReference.java:80-95): a[aOff+p*lda+i]. */
void oracle_gemm_strided_col_major_a(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                                     double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta) {
    oracle_gemm_general(ORACLE_COL, ORACLE_ROW, m, n, k, alpha, a + aOff, lda, b + bOff, ldb, beta, c + cOff, ldc, 1);
}

/* GemmReference.gemmStridedColMajorB (GemmReference.java:97-112): b[bOff+j*ldb+p]. */
void oracle_gemm_strided_col_major_b(const double* a, int aOff, int lda, const double* b, int bOff, int ldb,
                                     double* c, int cOff, int ldc, int m, int k, int n, double alpha, double beta) {
    oracle_gemm_general(ORACLE_ROW, ORACLE_COL, m, n, k, alpha, a + aOff, lda, b + bOff, ldb, beta, c + cOff, ldc, 1);
}

/* GemmReference.assertParity metrics (GemmReference.java:148-173). */
void oracle_parity(const double* expected, const double* actual, int64_t len, double* max_abs_err, double* rel_err) {
    double maxe = 0.0, diff_sum = 0.0, ref_sum = 0.0;
    for (int64_t i = 0; i < len; i++) {
        double d = actual[i] - expected[i];
        double ad = fabs(d);
        if (ad > maxe || isnan(ad)) maxe = isnan(ad) ? NAN : (isnan(maxe) ? maxe : ad);
        diff_sum += d * d;
        ref_sum += expected[i] * expected[i];
    }
    double ref_norm = sqrt(ref_sum);
    *max_abs_err = maxe;
    *rel_err = ref_norm == 0.0 ? 0.0 : sqrt(diff_sum) / ref_norm;
}

static double max_abs(const double* v, int64_t len) {
    double m = 0.0;
    for (int64_t i = 0; i < len; i++) m = fmax(m, fabs(v[i]));
    return m;
}

/* cpuAbsTolerance (GemmReference.java:175-177). */
double oracle_cpu_abs_tol(const double* expected, int64_t len, int64_t k) {
    return fmax(1e-12, 1e-9 * fmax(1.0, max_abs(expected, len)) * fmax(1.0, (double)k / 16.0));
}

/* cudaAbsTolerance (GemmReference.java:179-181): the reference's GPU grade. */
double oracle_gpu_abs_tol(const double* expected, int64_t len, int64_t k) {
    return fmax(1e-12, 1e-8 * fmax(1.0, max_abs(expected, len)) * fmax(1.0, (double)k / 16.0));
}

/* ------------------------------------------------------------------------- */
/* PARALLEL_MICRO restatement, the timed CPU baseline.
 * Block sizes: GemmDispatch.computeBlockSizes (GemmDispatch.java:118-149) with
 * the 32 KiB / 256 KiB / 8 MiB cache constants (:25-27) gives KC=128, NC=128,
 * MC=2048 for both the AVX2 (MR=4, NR=4) and AVX-512 (MR=6, NR=8) cases
 * (MicroKernel.optimalMR, MicroKernel.java:29-37). */
#define PM_KC 128
#define PM_NC 128
#define PM_MC 2048


static int pm_use_avx512(void) {
#if defined(__x86_64__)
    return __builtin_cpu_supports("avx512f");
#else
    return 0;
#endif
}

void oracle_block_sizes(int* out5) {
    int wide = pm_use_avx512();
    out5[0] = PM_MC; out5[1] = PM_KC; out5[2] = PM_NC;
    out5[3] = wide ? 6 : 4; out5[4] = wide ? 8 : 4;
}

/* PackingUtils.scaleCPanel (PackingUtils.java:159-183): beta==0 fills zero, beta==1 no-op. */
static void pm_scale_c(double* c, int64_t m, int64_t n, int64_t ldc, double beta) {
    if (beta == 1.0) return;
    for (int64_t i = 0; i < m; i++) {
        double* row = c + i * ldc;
        if (beta == 0.0) memset(row, 0, (size_t)n * sizeof(double));
        else for (int64_t j = 0; j < n; j++) row[j] *= beta;
    }
}

/* MicroKernel.computeBlock (MicroKernel.java:94-271): C held in registers for
 * one NR-wide column strip while an FMA chain runs over the KC panel. */
__attribute__((target("avx2,fma")))
static void pm_micro_4x4(int mr, int kb, int packed_n, int actual_n, const double* ap, const double* bp,
                         double* c, int64_t ldc) {
    for (int j = 0; j < actual_n; j += 4) {
        int w = actual_n - j < 4 ? actual_n - j : 4;
        __m256d acc[4];
        for (int r = 0; r < mr; r++) {
            double t[4] = {0, 0, 0, 0};
            for (int q = 0; q < w; q++) t[q] = c[r * ldc + j + q];
            acc[r] = _mm256_loadu_pd(t);
        }
        for (int p = 0; p < kb; p++) {
            __m256d bv = _mm256_loadu_pd(bp + (int64_t)p * packed_n + j);
            for (int r = 0; r < mr; r++)
                acc[r] = _mm256_fmadd_pd(_mm256_set1_pd(ap[r * kb + p]), bv, acc[r]);
        }
        for (int r = 0; r < mr; r++) {
            double t[4];
            _mm256_storeu_pd(t, acc[r]);
            for (int q = 0; q < w; q++) c[r * ldc + j + q] = t[q];
        }
    }
}

__attribute__((target("avx512f")))
static void pm_micro_6x8(int mr, int kb, int packed_n, int actual_n, const double* ap, const double* bp,
                         double* c, int64_t ldc) {
    for (int j = 0; j < actual_n; j += 8) {
        int w = actual_n - j < 8 ? actual_n - j : 8;
        __m512d acc[6];
        for (int r = 0; r < mr; r++) {
            double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int q = 0; q < w; q++) t[q] = c[r * ldc + j + q];
            acc[r] = _mm512_loadu_pd(t);
        }
        for (int p = 0; p < kb; p++) {
            __m512d bv = _mm512_loadu_pd(bp + (int64_t)p * packed_n + j);
            for (int r = 0; r < mr; r++)
                acc[r] = _mm512_fmadd_pd(_mm512_set1_pd(ap[r * kb + p]), bv, acc[r]);
        }
        for (int r = 0; r < mr; r++) {
            double t[8];
            _mm512_storeu_pd(t, acc[r]);
            for (int q = 0; q < w; q++) c[r * ldc + j + q] = t[q];
        }
    }
}

/* OptimizedBLAS3.computeTile (OptimizedBLAS3.java:356-385): per KC block, packB
 * (PackingUtils.java:121-133, zero padded to NR), then per MR strip packA with
 * alpha folded (PackingUtils.java:37-56) and the microkernel. */
static void pm_tile(const double* a, int64_t lda, const double* b, int64_t ldb, double* c, int64_t ldc, int64_t k,
                    int64_t r0, int64_t r1, int64_t c0, int64_t c1, double alpha, int mr, int nr, int wide,
                    double* bpack, double* apack) {
    int nb = (int)(c1 - c0);
    int packed_n = (nb + nr - 1) / nr * nr;
    for (int64_t kk = 0; kk < k; kk += PM_KC) {
        int kb = (int)((kk + PM_KC < k ? kk + PM_KC : k) - kk);
        for (int p = 0; p < kb; p++) {
            memcpy(bpack + (int64_t)p * packed_n, b + (kk + p) * ldb + c0, (size_t)nb * sizeof(double));
            for (int q = nb; q < packed_n; q++) bpack[(int64_t)p * packed_n + q] = 0.0;
        }
        for (int64_t i = r0; i < r1; i += mr) {
            int mb = (int)(r1 - i < mr ? r1 - i : mr);
            for (int r = 0; r < mb; r++) {
                const double* src = a + (i + r) * lda + kk;
                if (alpha == 1.0) memcpy(apack + r * kb, src, (size_t)kb * sizeof(double));
                else for (int p = 0; p < kb; p++) apack[r * kb + p] = src[p] * alpha;
            }
            if (wide) pm_micro_6x8(mb, kb, packed_n, nb, apack, bpack, c + i * ldc + c0, ldc);
            else pm_micro_4x4(mb, kb, packed_n, nb, apack, bpack, c + i * ldc + c0, ldc);
        }
    }
}

/* GemmDispatch.optimalParallelism (GemmDispatch.java:183-208). */
static int pm_threads(int64_t m, int64_t n, int64_t k, int max_threads) {
    if (max_threads <= 1) return 1;
    int64_t flops = 2 * m * n * k;
    if (flops < 5000000) return 1;
    int64_t tiles = ((m + PM_MC - 1) / PM_MC) * ((n + PM_NC - 1) / PM_NC);
    int64_t useful = flops / 500000;
    if (useful > tiles) useful = tiles;
    int64_t t = max_threads < useful ? max_threads : useful;
    if (t < 1) t = 1;
    int64_t p = 1;
    while (p * 2 <= t) p *= 2;
    return (int)p;
}

/* OptimizedBLAS3.gemmParallelMicrokernel (OptimizedBLAS3.java:295-351): scale C
 * by beta, return when alpha==0 or k==0, then one task per MC x NC tile of C. */
void oracle_parallel_micro(int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                           const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int max_threads,
                           int* threads_used) {
    pm_scale_c(c, m, n, ldc, beta);
    if (threads_used) *threads_used = 1;
    if (alpha == 0.0 || k == 0 || m == 0 || n == 0) return;
    int wide = pm_use_avx512();
    int mr = wide ? 6 : 4, nr = wide ? 8 : 4;
    int threads = pm_threads(m, n, k, max_threads);
    if (threads_used) *threads_used = threads;
    int64_t brows = (m + PM_MC - 1) / PM_MC, bcols = (n + PM_NC - 1) / PM_NC;
    int64_t tiles = brows * bcols;
#pragma omp parallel num_threads(threads)
    {
        double* bpack = (double*)aligned_alloc(64, (size_t)PM_KC * (PM_NC + 8) * sizeof(double));
        double* apack = (double*)aligned_alloc(64, (size_t)8 * PM_KC * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < tiles; t++) {
            int64_t br = t / bcols, bc = t % bcols;
            int64_t r0 = br * PM_MC, c0 = bc * PM_NC;
            int64_t r1 = r0 + PM_MC < m ? r0 + PM_MC : m;
            int64_t c1 = c0 + PM_NC < n ? c0 + PM_NC : n;
            pm_tile(a, lda, b, ldb, c, ldc, k, r0, r1, c0, c1, alpha, mr, nr, wide, bpack, apack);
        }
        free(bpack);
        free(apack);
    }
}

/* SpecializedKernels.matvec (SpecializedKernels.java:27-71): beta-scale y
 * (beta==0 fills zero), return on alpha==0 or k==0, then y += alpha*dot(row, x). */
void oracle_matvec(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x,
                   double beta, double* y) {
    for (int64_t i = 0; i < m; i++) y[i] = (beta == 0.0) ? 0.0 : (beta == 1.0 ? y[i] : y[i] * beta);
    if (alpha == 0.0 || k == 0) return;
    for (int64_t i = 0; i < m; i++) {
        double s = 0.0;
        for (int64_t p = 0; p < k; p++) s = fma(a[i * lda + p], x[p], s);
        y[i] += alpha * s;
    }
}

/* ------------------------------------------------------------------------- */
/* Double-double arithmetic: TwoSum (Knuth), TwoProduct via FMA, and the
 * accurate dd+dd addition (Li et al., "Design, implementation and testing of
 * extended and mixed precision BLAS", 2002; QD library algorithm). */
static inline void two_sum(double a, double b, double* s, double* e) {
    double x = a + b;
    double bb = x - a;
    *e = (a - (x - bb)) + (b - bb);
    *s = x;
}

static inline void fast_two_sum(double a, double b, double* s, double* e) {
    double x = a + b;
    *e = b - (x - a);
    *s = x;
}

static inline void dd_add(double ah, double al, double bh, double bl, double* rh, double* rl) {
    double s1, s2, t1, t2;
    two_sum(ah, bh, &s1, &s2);
    two_sum(al, bl, &t1, &t2);
    s2 += t1;
    fast_two_sum(s1, s2, &s1, &s2);
    s2 += t2;
    fast_two_sum(s1, s2, rh, rl);
}

static inline void dd_mul_d(double ah, double al, double b, double* rh, double* rl) {
    double p = ah * b;
    double e = fma(ah, b, -p);
    e = fma(al, b, e);
    fast_two_sum(p, e, rh, rl);
}

static void dd_dot(int a_layout, int b_layout, int64_t k, const double* a, int64_t lda, const double* b,
                   int64_t ldb, int64_t i, int64_t j, double* hi, double* lo, double* absum) {
    double sh = 0.0, sl = 0.0, ab = 0.0;
    for (int64_t p = 0; p < k; p++) {
        double av = (a_layout == ORACLE_ROW) ? a[i * lda + p] : a[p * lda + i];
        double bv = (b_layout == ORACLE_ROW) ? b[p * ldb + j] : b[j * ldb + p];
        double ph = av * bv;
        double pl = fma(av, bv, -ph);
        dd_add(sh, sl, ph, pl, &sh, &sl);
        ab += fabs(ph);
    }
    *hi = sh;
    *lo = sl;
    if (absum) *absum = ab * (1.0 + 2.0 * (double)(k + 1) * 0x1.0p-53);
}

void oracle_dd_entries(int a_layout, int b_layout, int64_t k, const double* a, int64_t lda, const double* b,
                       int64_t ldb, int64_t count, const int64_t* ii, const int64_t* jj, double* hi, double* lo,
                       double* abs_out, int threads) {
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16)
    for (int64_t t = 0; t < count; t++)
        dd_dot(a_layout, b_layout, k, a, lda, b, ldb, ii[t], jj[t], &hi[t], &lo[t], abs_out ? &abs_out[t] : NULL);
}

void oracle_dd_gemm(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                    int64_t lda, const double* b, int64_t ldb, double beta, const double* c, int64_t ldc,
                    double* hi, double* lo, int threads) {
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t i = 0; i < m; i++) {
        for (int64_t j = 0; j < n; j++) {
            double sh, sl;
            dd_dot(a_layout, b_layout, k, a, lda, b, ldb, i, j, &sh, &sl, NULL);
            double xh, xl, yh = 0.0, yl = 0.0;
            dd_mul_d(sh, sl, alpha, &xh, &xl);
            if (beta != 0.0) {
                double cv = c[i * ldc + j];
                yh = cv * beta;
                yl = fma(cv, beta, -yh);
            }
            dd_add(xh, xl, yh, yl, &hi[i * n + j], &lo[i * n + j]);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Householder QR restatement (src/main/java/net/faulj/decomposition/qr/HouseholderQR.java).
 * Column-major storage, element (i,j) at a[j*lda+i] (the reference's AT layout, :98-127).
 * Reflector convention of computeHouseholder (:219-291): max-scaled norm,
 * beta = -sign(x0)*hypot(x0,||x1||), tau = (beta-x0)/beta, v = x/(x0-beta) with
 * v0 = 1 implicit, R diagonal = beta. The blocked driver (:133-209) and the
 * unblocked one give the same factors in exact arithmetic; this restatement
 * applies each reflector to the trailing columns directly (the panel apply of
 * applyHouseholderToColumn, :343-385), which is the numerically plain form. */
static double qr_reflector(int64_t m, double* col, int64_t start, double* tau_out) {
    int64_t len = m - start;
    if (len <= 1) { *tau_out = 0.0; return col[start]; }
    double maxv = fabs(col[start]);
    for (int64_t i = start + 1; i < m; i++) maxv = fmax(maxv, fabs(col[i]));
    if (maxv < 2.2250738585072014e-308) { *tau_out = 0.0; return col[start]; }
    double inv = 1.0 / maxv;
    double x0 = col[start] * inv;
    double sigma = 0.0;
    for (int64_t i = start + 1; i < m; i++) { double v = col[i] * inv; sigma += v * v; }
    double xnorm = sqrt(sigma);
    if (xnorm == 0.0) {
        if (x0 >= 0.0) { *tau_out = 0.0; return col[start]; }
        *tau_out = 2.0;
        col[start] = -col[start];
        return -col[start];
    }
    double nbeta = -copysign(hypot(x0, xnorm), x0);
    double beta = nbeta * maxv;
    double tau = (nbeta - x0) / nbeta;
    double v0inv = 1.0 / ((x0 - nbeta) * maxv);
    for (int64_t i = start + 1; i < m; i++) col[i] *= v0inv;
    *tau_out = tau;
    col[start] = beta;
    return beta;
}

int oracle_qr_factor(int64_t m, int64_t n, double* a, int64_t lda, double* tau, int64_t nb) {
    (void)nb;
    int64_t kmax = m < n ? m : n;
    for (int64_t k = 0; k < kmax; k++) {
        double* ck = a + k * lda;
        qr_reflector(m, ck, k, &tau[k]);
        if (tau[k] == 0.0) continue;
        /* H = I - tau v v^T with v = [1; ck[k+1:m]] applied to columns j > k. */
        for (int64_t j = k + 1; j < n; j++) {
            double* cj = a + j * lda;
            double dot = cj[k];
            for (int64_t i = k + 1; i < m; i++) dot += ck[i] * cj[i];
            double s = tau[k] * dot;
            cj[k] -= s;
            for (int64_t i = k + 1; i < m; i++) cj[i] -= s * ck[i];
        }
    }
    return 0;
}

/* Q (m x n, thin) by reverse reflector application, as buildQ (:390-444). */
void oracle_qr_form_q(int64_t m, int64_t n, const double* qr, int64_t lda, const double* tau, double* q,
                      int64_t ldq) {
    int64_t kmax = m < n ? m : n;
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < m; i++) q[j * ldq + i] = (i == j) ? 1.0 : 0.0;
    for (int64_t k = kmax - 1; k >= 0; k--) {
        const double* v = qr + k * lda;
        if (tau[k] == 0.0) continue;
        for (int64_t j = 0; j < n; j++) {
            double* qj = q + j * ldq;
            double dot = qj[k];
            for (int64_t i = k + 1; i < m; i++) dot += v[i] * qj[i];
            double s = tau[k] * dot;
            qj[k] -= s;
            for (int64_t i = k + 1; i < m; i++) qj[i] -= s * v[i];
        }
    }
}

/* Thin QR of the column-major m x n matrix a (lda): Q (m x k, ldq) and the upper-trapezoidal R
 * (k x n, ldr), k = min(m, n) (HouseholderQR.decomposeThin, HouseholderQR.java:27, :52-76). */
static void thin_qr(int64_t m, int64_t n, const double* a, int64_t lda, double* q, int64_t ldq, double* r,
                    int64_t ldr) {
    const int64_t k = m < n ? m : n;
    double* w = (double*)malloc(sizeof(double) * (size_t)(m * n + 1));
    double* tau = (double*)calloc((size_t)(k + 1), sizeof(double));
    for (int64_t j = 0; j < n; j++) memcpy(w + j * m, a + j * lda, sizeof(double) * (size_t)m);
    oracle_qr_factor(m, n, w, m, tau, 64);
    oracle_qr_form_q(m, k, w, m, tau, q, ldq);
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < k; i++) r[j * ldr + i] = i <= j ? w[j * m + i] : 0.0;
    free(w);
    free(tau);
}

/* TallSkinnyQR.decompose (src/main/java/net/faulj/decomposition/qr/TallSkinnyQR.java:22-90): m <= n is
 * a thin QR of A; otherwise row blocks of max(n, 64) rows, a thin QR of each, a thin QR of the R
 * factors stacked in block order, and each block's rows of Q as its Q times its rows of the stack's
 * Q (a left-to-right sum). Column-major A (lda), Q (m x k, ldq), R (k x n, ldr), k = min(m, n). */
void oracle_tsqr(int64_t m, int64_t n, const double* a, int64_t lda, double* q, int64_t ldq, double* r, int64_t ldr) {
    if (m <= 0 || n <= 0) return;
    if (m <= n) {
        thin_qr(m, n, a, lda, q, ldq, r, ldr);
        return;
    }
    const int64_t b = n > 64 ? n : 64;
    const int64_t nblk = (m + b - 1) / b;
    int64_t srows = 0;
    for (int64_t t = 0; t < nblk; t++) {
        const int64_t rows = m - t * b < b ? m - t * b : b;
        srows += rows < n ? rows : n;
    }
    double* stack = (double*)calloc((size_t)(srows * n + 1), sizeof(double));
    double* qblk = (double*)calloc((size_t)(m * n + 1), sizeof(double));  /* block t's Q at qblk + t*b*n */
    double* blk = (double*)malloc(sizeof(double) * (size_t)(b * n + 1));
    double* rblk = (double*)malloc(sizeof(double) * (size_t)(b * n + 1));
    for (int64_t t = 0, off = 0; t < nblk; t++) {
        const int64_t r0 = t * b, rows = m - r0 < b ? m - r0 : b, kt = rows < n ? rows : n;
        for (int64_t j = 0; j < n; j++) memcpy(blk + j * rows, a + j * lda + r0, sizeof(double) * (size_t)rows);
        thin_qr(rows, n, blk, rows, qblk + r0 * n, rows, rblk, kt);
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < kt; i++) stack[j * srows + off + i] = rblk[j * kt + i];
        off += kt;
    }
    double* q2 = (double*)malloc(sizeof(double) * (size_t)(srows * n + 1));
    thin_qr(srows, n, stack, srows, q2, srows, r, ldr);
    for (int64_t t = 0, off = 0; t < nblk; t++) {
        const int64_t r0 = t * b, rows = m - r0 < b ? m - r0 : b, kt = rows < n ? rows : n;
        const double* qt = qblk + r0 * n;  /* rows x kt, column-major, ld rows */
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < rows; i++) {
                double sum = 0.0;
                for (int64_t p = 0; p < kt; p++) sum += qt[p * rows + i] * q2[j * srows + off + p];
                q[j * ldq + r0 + i] = sum;
            }
        off += kt;
    }
    free(stack);
    free(qblk);
    free(blk);
    free(rblk);
    free(q2);
}

/* ------------------------------------------------------------------------- */
/* Split-precision restatement (see oracle.h). Every FP64 step mirrors the
 * device code: frexp of the max |x|, ldexp scaling, v*128 / trunc / subtract
 * digit extraction, exact integer diagonal sums, sequential recombination. */
#define SPLIT_NONFINITE (-100000)

int oracle_split_exponent(const double* x, int64_t stride, int64_t count) {
    double mx = 0.0;
    for (int64_t i = 0; i < count; i++) {
        double v = fabs(x[i * stride]);
        if (!(v <= 1.7976931348623157e308)) return SPLIT_NONFINITE;
        mx = fmax(mx, v);
    }
    int ex = 0;
    if (mx > 0.0) frexp(mx, &ex);
    return ex;
}

static void split_digits(const double* x, int64_t stride, int64_t count, int e, int slices, int bits, int16_t* out,
                         int64_t out_stride) {
    const double radix = ldexp(1.0, bits);
    for (int64_t i = 0; i < count; i++) {
        double v = (e == SPLIT_NONFINITE) ? 0.0 : ldexp(x[i * stride], -e);
        for (int t = 0; t < slices; t++) {
            v *= radix;
            double d = trunc(v);
            v -= d;
            out[t * out_stride + i] = (int16_t)(int)d;
        }
    }
}

void oracle_split_slices(const double* x, int64_t stride, int64_t count, int e, int slices, int8_t* out,
                         int64_t out_stride) {
    for (int64_t i = 0; i < count; i++) {
        double v = (e == SPLIT_NONFINITE) ? 0.0 : ldexp(x[i * stride], -e);
        for (int t = 0; t < slices; t++) {
            v *= 128.0;
            double d = trunc(v);
            v -= d;
            out[t * out_stride + i] = (int8_t)(int)d;
        }
    }
}

/* Diagonal-sum chunk: S * chunk * 127^2 < 2^31 with chunk a multiple of 32. */
int64_t oracle_split_chunk_k(int slices, int64_t k) {
    int64_t steps = 2147483647LL / ((int64_t)slices * 32 * 127 * 127);
    int64_t ksteps = (k + 31) / 32;
    if (steps > ksteps) steps = ksteps;
    return steps * 32;
}

void oracle_split_gemm(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                       int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int slices,
                       int64_t chunk_k, int threads) {
    oracle_split_gemm_digits(a_layout, b_layout, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, slices, 7, chunk_k,
                             threads);
}

/* BF16 variant: 8-bit digits, FP64 flush every chunk_k (oracle_split_bf16_chunk_k). */
int64_t oracle_split_bf16_chunk_k(int slices, int64_t k) {
    int64_t fsteps = (1LL << 24) / ((int64_t)slices * 255 * 255 * 16);
    if (fsteps < 1) fsteps = 1;
    int64_t chunk = 2147483647LL / ((int64_t)slices * 16 * 255 * 255);
    chunk = chunk / fsteps * fsteps;
    int64_t ksteps = (k + 15) / 16;
    if (chunk > ksteps) chunk = ksteps;
    return chunk * 16;
}

void oracle_split_gemm_digits(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha,
                              const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                              int64_t ldc, int slices, int bits, int64_t chunk_k, int threads) {
    const int S = slices;
    int16_t* as = (int16_t*)calloc((size_t)(S * m * k + 1), sizeof(int16_t));
    int16_t* bs = (int16_t*)calloc((size_t)(S * n * k + 1), sizeof(int16_t));
    int* ea = (int*)malloc(sizeof(int) * (size_t)(m + 1));
    int* eb = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    for (int64_t i = 0; i < m; i++) {
        const double* row = a_layout == ORACLE_ROW ? a + i * lda : a + i;
        int64_t st = a_layout == ORACLE_ROW ? 1 : lda;
        ea[i] = oracle_split_exponent(row, st, k);
        split_digits(row, st, k, ea[i], S, bits, as + i * k, m * k);
    }
    for (int64_t j = 0; j < n; j++) {
        const double* col = b_layout == ORACLE_ROW ? b + j : b + j * ldb;
        int64_t st = b_layout == ORACLE_ROW ? ldb : 1;
        eb[j] = oracle_split_exponent(col, st, k);
        split_digits(col, st, k, eb[j], S, bits, bs + j * k, n * k);
    }
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t i = 0; i < m; i++) {
        for (int64_t j = 0; j < n; j++) {
            double* p = &c[i * ldc + j];
            for (int64_t k0 = 0, first = 1; k0 < k || first; k0 += chunk_k, first = 0) {
                int64_t k1 = k0 + chunk_k < k ? k0 + chunk_k : k;
                int64_t acc[16] = {0};
                for (int t = 0; t < S; t++)
                    for (int u = 0; u + t <= S - 1; u++) {
                        const int16_t* ar = as + (int64_t)t * m * k + i * k;
                        const int16_t* br = bs + (int64_t)u * n * k + j * k;
                        int64_t sdot = 0;
                        for (int64_t q = k0; q < k1; q++) sdot += (int64_t)ar[q] * br[q];
                        acc[t + u] += sdot;
                    }
                double sum = 0.0;
                for (int d = S - 1; d >= 0; d--) sum += (double)acc[d] * ldexp(1.0, -bits * (d + 2));
                double v = (ea[i] == SPLIT_NONFINITE || eb[j] == SPLIT_NONFINITE) ? NAN : ldexp(sum, ea[i] + eb[j]);
                if (first) *p = (beta == 0.0) ? alpha * v : fma(alpha, v, beta * *p);
                else *p = fma(alpha, v, *p);
                if (k1 >= k) break;
            }
        }
    }
    free(as);
    free(bs);
    free(ea);
    free(eb);
}

/* ---------------------------------------------------------------------------
 * CRT split form (modular integer emulation; Ozaki, Uchino, Imamura 2025,
 * "Ozaki Scheme II: a GEMM-oriented emulation of floating-point matrix
 * multiplication using an integer modular technique"). No reference
 * counterpart: the split-precision component of the DGEMM engine.
 *
 *   1. Row i of op(A) is scaled to integers a' = rint(a * 2^(pa - ea[i]))
 *      (|a'| <= 2^pa), column j of op(B) likewise with pb, eb[j].
 *   2. For moduli m_0..m_{N-1} (pairwise coprime, <= 256) the symmetric
 *      residues a' mod m_t, b' mod m_t fit in int8; each modular product
 *      C_t = (A'_t B'_t) mod m_t is one integer GEMM.
 *   3. With M = prod m_t > 4 * k * 2^(pa+pb), C' = A'B' is the unique integer
 *      in (-M/2, M/2) with those residues: X = sum_t r_t W_t (W_t = (M/m_t) *
 *      ((M/m_t)^-1 mod m_t)) is formed exactly in four 40-bit limbs,
 *      q = rint(X / M), C' = X - qM limb by limb, limbs balanced and summed
 *      by Horner from the top in double-double, rounded once at the end.
 *   4. C = alpha * C' * 2^(ea+eb-pa-pb) (+ beta*C through fma).
 * ------------------------------------------------------------------------- */
static const int CRT_MODULI[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                   223, 217, 211, 199, 197, 193, 191, 181, 179, 173};
#define CRT_LIMBS 4
#define CRT_LIMB_BITS 40

int oracle_crt_max_moduli(void) { return 20; }
int oracle_crt_modulus(int t) { return (t >= 0 && t < 20) ? CRT_MODULI[t] : 0; }

double oracle_crt_log2_product(int nmod) {
    double s = 0.0;
    for (int t = 0; t < nmod && t < 20; t++) s += log2((double)CRT_MODULI[t]);
    return s;
}

/* Little-endian base-2^32 big integers of 8 words (256 bits). */
static void big_set_small(uint32_t* x, uint32_t v) {
    memset(x, 0, 8 * sizeof(uint32_t));
    x[0] = v;
}
static void big_mul_small(uint32_t* x, uint32_t v) {
    uint64_t carry = 0;
    for (int w = 0; w < 8; w++) {
        uint64_t t = (uint64_t)x[w] * v + carry;
        x[w] = (uint32_t)t;
        carry = t >> 32;
    }
}
static uint32_t big_mod_small(const uint32_t* x, uint32_t m) {
    uint64_t r = 0;
    for (int w = 7; w >= 0; w--) r = ((r << 32) | x[w]) % m;
    return (uint32_t)r;
}
/* Bits [40*j, 40*j+40) of x as a double (exact). */
static double big_limb(const uint32_t* x, int j) {
    double v = 0.0;
    for (int b = CRT_LIMB_BITS - 1; b >= 0; b--) {
        int bit = j * CRT_LIMB_BITS + b;
        v = v * 2.0 + (double)((x[bit >> 5] >> (bit & 31)) & 1u);
    }
    return v;
}
static double big_to_double(const uint32_t* x) {
    double v = 0.0;
    for (int w = 7; w >= 0; w--) v = v * 4294967296.0 + (double)x[w];
    return v;
}

/* Tables of the reconstruction for N moduli: w[t*4 + j] = limb j of W_t,
 * mlimb[j] = limb j of M, finv[j] = 2^(40 j) / M. */
void oracle_crt_tables(int nmod, double* w, double* mlimb, double* finv) {
    uint32_t M[8], Mi[8];
    big_set_small(M, 1);
    for (int t = 0; t < nmod; t++) big_mul_small(M, (uint32_t)CRT_MODULI[t]);
    for (int t = 0; t < nmod; t++) {
        big_set_small(Mi, 1);
        for (int u = 0; u < nmod; u++)
            if (u != t) big_mul_small(Mi, (uint32_t)CRT_MODULI[u]);
        const uint32_t m = (uint32_t)CRT_MODULI[t];
        const uint32_t r = big_mod_small(Mi, m);
        uint32_t y = 0;
        for (uint32_t c = 1; c < m; c++)
            if ((r * c) % m == 1) { y = c; break; }
        big_mul_small(Mi, y);
        for (int j = 0; j < CRT_LIMBS; j++) w[t * CRT_LIMBS + j] = big_limb(Mi, j);
    }
    const double md = big_to_double(M);
    for (int j = 0; j < CRT_LIMBS; j++) {
        mlimb[j] = big_limb(M, j);
        finv[j] = ldexp(1.0, CRT_LIMB_BITS * j) / md;
    }
}

/* Symmetric residue of an integer-valued double |x| < 2^71: in [-128, 127] for
 * m = 256 and [-(m-1)/2, (m-1)/2] for odd m. */
static int crt_residue(double x, int m) {
    __int128 v = (__int128)x;
    int r = (int)(v % m);
    if (r > (m - 1) / 2) r -= m;
    if (r < -(m / 2)) r += m;
    return r;
}

/* Integer image of x in a row with exponent e scaled to p bits. */
static double crt_scaled(double x, int e, int p) {
    double s = ldexp(x, p - e);
    return fabs(s) < 9007199254740992.0 ? rint(s) : s;
}

/* Norm exponent of the CRT form's scaling: with e the row's max exponent,
 * S = sum over entries of ceil(fl(v*v) * 2^40) + 1 (v = |x| 2^-e, 0 entries
 * add 0) is an integer bound of ||x||^2 in units of 2^(2e-40); the result is
 * t = e + ceil((L - 40) / 2) with 2^L >= S, so ||x||_2 <= 2^t. Integer
 * sums: any summation order gives the same t. */
int oracle_crt_norm_exponent(const double* x, int64_t stride, int64_t count, int e) {
    if (e == SPLIT_NONFINITE) return SPLIT_NONFINITE;
    unsigned long long S = 0;
    for (int64_t i = 0; i < count; i++) {
        const double v = ldexp(fabs(x[i * stride]), -e);
        if (v > 0.0) S += (unsigned long long)ceil(v * v * 0x1p40) + 1ull;
    }
    if (S == 0) return 0;
    int L = 0;
    while (L < 64 && (S - 1) >> L) L++; /* bit length of S - 1 */
    const int twice = L - 40;           /* ceil(twice / 2) */
    return e + (twice >= 0 ? (twice + 1) / 2 : -((-twice) / 2));
}

void oracle_crt_residues(const double* x, int64_t stride, int64_t count, int e, int p, int nmod, int8_t* out,
                         int64_t out_stride) {
    for (int64_t q = 0; q < count; q++) {
        const double s = (e == SPLIT_NONFINITE) ? 0.0 : crt_scaled(x[q * stride], e, p);
        for (int t = 0; t < nmod; t++) out[t * out_stride + q] = (int8_t)crt_residue(s, CRT_MODULI[t]);
    }
}

/* Reconstructs the integer with the given residues (|value| < M/4) as a double. */
double oracle_crt_reconstruct(const int* r, int nmod, const double* w, const double* mlimb, const double* finv) {
    double X[CRT_LIMBS];
    for (int j = 0; j < CRT_LIMBS; j++) {
        double s = 0.0;
        for (int t = 0; t < nmod; t++) s += (double)r[t] * w[t * CRT_LIMBS + j];
        X[j] = s;
    }
    double qf = 0.0;
    for (int j = CRT_LIMBS - 1; j >= 0; j--) qf += X[j] * finv[j];
    const double q = rint(qf);
    double D[CRT_LIMBS];
    for (int j = 0; j < CRT_LIMBS; j++) D[j] = fma(-q, mlimb[j], X[j]);
    const double base = ldexp(1.0, CRT_LIMB_BITS);
    for (int j = 0; j < CRT_LIMBS - 1; j++) {
        const double c = rint(ldexp(D[j], -CRT_LIMB_BITS));
        D[j] = D[j] - c * base;
        D[j + 1] = D[j + 1] + c;
    }
    /* Horner from the top limb in double-double (TwoSum per step). */
    double h = D[CRT_LIMBS - 1], l = 0.0;
    for (int j = CRT_LIMBS - 2; j >= 0; j--) {
        const double hb = h * base;
        const double s = hb + D[j];
        const double bb = s - hb;
        const double e = (hb - (s - bb)) + (D[j] - bb);
        l = l * base + e;
        h = s;
    }
    return h + l;
}

void oracle_crt_gemm(int a_layout, int b_layout, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                     int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int nmod, int pa,
                     int pb, int threads) {
    double w[20 * CRT_LIMBS], mlimb[CRT_LIMBS], finv[CRT_LIMBS];
    oracle_crt_tables(nmod, w, mlimb, finv);
    int8_t* ar = (int8_t*)calloc((size_t)(nmod * m * k + 1), 1);
    int8_t* br = (int8_t*)calloc((size_t)(nmod * n * k + 1), 1);
    int* ea = (int*)malloc(sizeof(int) * (size_t)(m + 1));
    int* eb = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    for (int64_t i = 0; i < m; i++) {
        const double* row = a_layout == ORACLE_ROW ? a + i * lda : a + i;
        int64_t st = a_layout == ORACLE_ROW ? 1 : lda;
        ea[i] = oracle_crt_norm_exponent(row, st, k, oracle_split_exponent(row, st, k));
        oracle_crt_residues(row, st, k, ea[i], pa, nmod, ar + i * k, m * k);
    }
    for (int64_t j = 0; j < n; j++) {
        const double* col = b_layout == ORACLE_ROW ? b + j : b + j * ldb;
        int64_t st = b_layout == ORACLE_ROW ? ldb : 1;
        eb[j] = oracle_crt_norm_exponent(col, st, k, oracle_split_exponent(col, st, k));
        oracle_crt_residues(col, st, k, eb[j], pb, nmod, br + j * k, n * k);
    }
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t i = 0; i < m; i++) {
        for (int64_t j = 0; j < n; j++) {
            int r[20];
            for (int t = 0; t < nmod; t++) {
                const int8_t* x = ar + (int64_t)t * m * k + i * k;
                const int8_t* y = br + (int64_t)t * n * k + j * k;
                int64_t s = 0;
                for (int64_t q = 0; q < k; q++) s += (int64_t)x[q] * y[q];
                const int mt = CRT_MODULI[t];
                int rr = (int)(s % mt);
                if (rr > (mt - 1) / 2) rr -= mt;
                if (rr < -(mt / 2)) rr += mt;
                r[t] = rr;
            }
            const double cp = oracle_crt_reconstruct(r, nmod, w, mlimb, finv);
            const double v = (ea[i] == SPLIT_NONFINITE || eb[j] == SPLIT_NONFINITE)
                                 ? NAN
                                 : ldexp(cp, ea[i] + eb[j] - pa - pb);
            double* p = &c[i * ldc + j];
            *p = (beta == 0.0) ? alpha * v : fma(alpha, v, beta * *p);
        }
    }
    free(ar);
    free(br);
    free(ea);
    free(eb);
}
