/*
 * CPU oracle for zml's dense factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, in plain C, the
 * algorithms of zml's BLAS/LAPACK files so that the GPU library can be
 * checked against them. Nothing in the product (the HIP library, its C ABI
 * or the Python host layer) links, loads or calls it; only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark do.
 *
 * Every routine cites the reference file and lines it follows. All
 * arithmetic is IEEE binary64 with separate multiply and add: the library
 * must be compiled with -ffp-contract=off (the Makefile does), because zml's
 * f64 ops never contract (src/float.zig:114-118).
 *
 * Enumerations use the CBLAS/LAPACKE numeric values that zml maps to
 * (include/cblas.h:11-37, src/types.zig:59-170, src/linalg.zig:79-162).
 */
#ifndef ZML_ORACLE_H
#define ZML_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ZR_ROW = 101, ZR_COL = 102 };
enum { ZR_NOTRANS = 111, ZR_TRANS = 112, ZR_CONJTRANS = 113, ZR_CONJNOTRANS = 114 };
enum { ZR_UPPER = 121, ZR_LOWER = 122 };
enum { ZR_NONUNIT = 131, ZR_UNIT = 132 };
enum { ZR_LEFT = 141, ZR_RIGHT = 142 };

/* Error code returned where zml raises InvalidArgument. */
#define ZR_EINVAL (-1)

/* ---- machine constants and block sizes ---------------------------------- */
double zr_lamch(char cmach);            /* lapack/lamch.zig:15-78: zml's values,
                                           N = 52, M = -1022, L = 1023 for f64 */
int zr_ilaenv_nb(const char* name);     /* lapack/ilaenv.zig:143-148,227-233 */

/* ---- BLAS (src/linalg/blas/) --------------------------------------- */
int zr_iamax(int n, const double* x, int incx);               /* iamax.zig:11-68 */
int zr_scal(int n, double alpha, double* x, int incx);         /* scal.zig:9-41   */
int zr_swap(int n, double* x, int incx, double* y, int incy);  /* swap.zig        */
int zr_gemm(int order, int transa, int transb, int m, int n, int k, double alpha,
            const double* a, int lda, const double* b, int ldb, double beta,
            double* c, int ldc);                               /* gemm.zig:14-1033 */
int zr_trsm(int order, int side, int uplo, int transa, int diag, int m, int n,
            double alpha, const double* a, int lda, double* b, int ldb); /* trsm.zig */
int zr_syrk(int order, int uplo, int trans, int n, int k, double alpha,
            const double* a, int lda, double beta, double* c, int ldc); /* syrk.zig */
int zr_gemv(int order, int trans, int m, int n, double alpha, const double* a, int lda,
            const double* x, int incx, double beta, double* y, int incy); /* gemv.zig */

int zr_trsv(int order, int uplo, int trans, int diag, int n, const double* a, int lda,
            double* x, int incx);                               /* trsv.zig */

/* trsm.zig:137-142,221-229 divide b[k + j*lda] (not ldb) by the diagonal in
 * the left/no-transpose/non-unit branches. 1 (default) restates that read
 * faithfully; 0 reads b[k + j*ldb], the evidently intended element. */
void zr_set_trsm_faithful_lda(int on);

/* ---- LAPACK (src/linalg/lapack/) ----------------------------------- */
int zr_laswp(int order, int n, double* a, int lda, int k1, int k2, const int32_t* ipiv, int incx);
int zr_getrf2(int order, int m, int n, double* a, int lda, int32_t* ipiv);
int zr_getrf(int order, int m, int n, double* a, int lda, int32_t* ipiv);
int zr_getrs(int order, int trans, int n, int nrhs, const double* a, int lda,
             const int32_t* ipiv, double* b, int ldb);
int zr_gesv(int order, int n, int nrhs, double* a, int lda, int32_t* ipiv, double* b, int ldb);
int zr_potrf2(int order, int uplo, int n, double* a, int lda);
int zr_potrf(int order, int uplo, int n, double* a, int lda);
int zr_potrs(int order, int uplo, int n, int nrhs, const double* a, int lda, double* b, int ldb);
int zr_posv(int order, int uplo, int n, int nrhs, double* a, int lda, double* b, int ldb);

/* ---- f32: the same routines at zml's T = f32 (zr_s*, one generic source) -- */
float zr_slamch(char cmach);
int zr_siamax(int n, const float* x, int incx);
int zr_sscal(int n, float alpha, float* x, int incx);
int zr_sswap(int n, float* x, int incx, float* y, int incy);
int zr_sgemm(int order, int transa, int transb, int m, int n, int k, float alpha,
             const float* a, int lda, const float* b, int ldb, float beta, float* c, int ldc);
int zr_strsm(int order, int side, int uplo, int transa, int diag, int m, int n,
             float alpha, const float* a, int lda, float* b, int ldb);
int zr_ssyrk(int order, int uplo, int trans, int n, int k, float alpha,
             const float* a, int lda, float beta, float* c, int ldc);
int zr_sgemv(int order, int trans, int m, int n, float alpha, const float* a, int lda,
             const float* x, int incx, float beta, float* y, int incy);
int zr_strsv(int order, int uplo, int trans, int diag, int n, const float* a, int lda,
             float* x, int incx);
int zr_slaswp(int order, int n, float* a, int lda, int k1, int k2, const int32_t* ipiv, int incx);
int zr_sgetrf2(int order, int m, int n, float* a, int lda, int32_t* ipiv);
int zr_sgetrf(int order, int m, int n, float* a, int lda, int32_t* ipiv);
int zr_sgetrs(int order, int trans, int n, int nrhs, const float* a, int lda,
              const int32_t* ipiv, float* b, int ldb);
int zr_sgesv(int order, int n, int nrhs, float* a, int lda, int32_t* ipiv, float* b, int ldb);
int zr_spotrf2(int order, int uplo, int n, float* a, int lda);
int zr_spotrf(int order, int uplo, int n, float* a, int lda);
int zr_spotrs(int order, int uplo, int n, int nrhs, const float* a, int lda, float* b, int ldb);
int zr_sposv(int order, int uplo, int n, int nrhs, float* a, int lda, float* b, int ldb);
int zr_sgetrf_unblocked(int order, int m, int n, float* a, int lda, int32_t* ipiv);
int zr_spotrf_lower_unblocked(int n, float* a, int lda);

/* ---- closed forms (derived, not zml files) ------------------------------ */
/* Unblocked right-looking LU with zml's pivot, scaling and update rules:
 * bitwise equal to zr_getrf on any input (up to the sign of exact zeros). */
int zr_getrf_unblocked(int order, int m, int n, double* a, int lda, int32_t* ipiv);
/* Same per-element operation sequence, blocked for cache and threaded with
 * OpenMP over rows; bitwise equal to zr_getrf_unblocked. Column-major only. */
int zr_getrf_fast(int m, int n, double* a, int lda, int32_t* ipiv, int nthreads);
/* Unblocked right-looking lower Cholesky (column-major): bitwise equal to
 * zr_potrf(ZR_COL, ZR_LOWER, ...). */
int zr_potrf_lower_unblocked(int n, double* a, int lda);

/* ---- synthetic inputs and error metrics --------------------------------- */
uint64_t zr_rng_u64(uint64_t seed, uint64_t ctr);
double zr_rng_uniform(uint64_t seed, uint64_t ctr);      /* [0, 1), 53-bit */
/* Fill an m x n matrix (logical element (i, j) draws counter i*n + j) with
 * lo + (hi - lo) * u, stored in the given layout with leading dimension ld. */
void zr_fill_uniform(int order, int m, int n, double* a, int ld, uint64_t seed, double lo, double hi);
/* ||A||_inf (max row sum of |a_ij|) of an m x n matrix. */
double zr_lange(int order, char norm, int m, int n, const double* a, int ld); /* no zml counterpart */
double zr_norm_inf(int order, int m, int n, const double* a, int ld);
/* r = b - A x in FP64 (row dots, ascending j), returns ||r||_inf. */
double zr_residual_inf(int order, int n, const double* a, int lda, const double* x, const double* b, double* r);
/* ||b - A x||_inf / (||A||_inf ||x||_inf n eps), eps = 2^-53. */
double zr_backward_error(int order, int n, const double* a, int lda, const double* x, const double* b);
/* Minimum over pivot steps k of (|p_k| - |second largest|)/|p_k| recorded
 * during zr_getrf_unblocked on a copy of A (diagnostic for pivot identity). */
double zr_min_pivot_gap(int order, int n, const double* a, int lda);

#ifdef __cplusplus
}
#endif
#endif
