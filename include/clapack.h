/*
 * LAPACKE entry points exported by libzmlhip.so. zml consumes these through
 * `-Dlink_lapacke=<lib>` (reference build.zig:21-24, src/c.zig:1-4); the
 * reference's own include/clapack.h is empty, so the signatures below are the
 * ones zml's call sites use (LP64: lapack_int is int, ipiv is int32, 1-based).
 *
 * Host pointers, synchronous. Return value: 0 on success, > 0 the LAPACK info
 * (first zero pivot for getrf, first non-positive pivot for potrf), -i when
 * argument i is invalid (arguments counted from 1 = matrix_layout, as LAPACKE
 * does), -1000 and below for a device error (text in zmlhip_last_error()).
 * matrix_layout is 101 (row-major) or 102 (column-major).
 */
#ifndef ZMLHIP_CLAPACK_H
#define ZMLHIP_CLAPACK_H

#ifdef __cplusplus
extern "C" {
#endif

/* replaces the call at src/linalg/lapack.zig:92 (lamch.zig:15-78): zml's own
 * values for f64, i.e. 'N' = 52 (stored mantissa bits), 'M' = -1022 and
 * 'L' = 1023 (std.math floatExponentMin/Max), where netlib's dlamch gives
 * 53, -1021 and 1024 */
double LAPACKE_dlamch(char cmach);
/* replaces src/linalg/lapack.zig:375 (laswp.zig:12-95) */
int LAPACKE_dlaswp(int matrix_layout, int n, double *a, int lda, int k1, int k2, const int *ipiv,
                   int incx);
/* replaces src/linalg/lapack.zig:758 (getrf2.zig:14-210) */
int LAPACKE_dgetrf2(int matrix_layout, int m, int n, double *a, int lda, int *ipiv);
/* replaces src/linalg/lapack.zig:1169 (getrf.zig:14-152) */
int LAPACKE_dgetrf(int matrix_layout, int m, int n, double *a, int lda, int *ipiv);
/* replaces src/linalg/lapack.zig:1575 (getrs.zig:13-142); trans 'N', 'T' or 'C' */
int LAPACKE_dgetrs(int matrix_layout, char trans, int n, int nrhs, const double *a, int lda,
                   const int *ipiv, double *b, int ldb);
/* replaces src/linalg/lapack.zig:2103 (gesv.zig:13-70) */
int LAPACKE_dgesv(int matrix_layout, int n, int nrhs, double *a, int lda, int *ipiv, double *b,
                  int ldb);
/* replaces src/linalg/lapack.zig:2541 (potrf2.zig:17-217); uplo 'U' or 'L' */
int LAPACKE_dpotrf2(int matrix_layout, char uplo, int n, double *a, int lda);
/* replaces src/linalg/lapack.zig:3012 (potrf.zig:17-240) */
int LAPACKE_dpotrf(int matrix_layout, char uplo, int n, double *a, int lda);
/* replaces src/linalg/lapack.zig:1160 (getrf.zig:14-152 at T = f32) */
int LAPACKE_sgetrf(int matrix_layout, int m, int n, float *a, int lda, int *ipiv);
/* replaces src/linalg/lapack.zig:3004 (potrf.zig:17-240 at T = f32) */
int LAPACKE_spotrf(int matrix_layout, char uplo, int n, float *a, int lda);
/* LAPACK dlange: 'M' max |a_ij|, 'O'/'1' largest column sum of |a_ij|, 'I'
 * largest row sum, 'F'/'E' Frobenius norm; a NaN entry gives NaN. No zml
 * counterpart (zml has no lange; the residual and backward-error step of a
 * solve needs ||A||). Returns the norm, or -i (as a double) for an invalid
 * argument i. */
double LAPACKE_dlange(int matrix_layout, char norm, int m, int n, const double *a, int lda);
/* replaces src/linalg/lapack.zig:3466 (potrs.zig:15-117) */
int LAPACKE_dpotrs(int matrix_layout, char uplo, int n, int nrhs, const double *a, int lda,
                   double *b, int ldb);
/* replaces src/linalg/lapack.zig:3948 (posv.zig:15-70) */
int LAPACKE_dposv(int matrix_layout, char uplo, int n, int nrhs, double *a, int lda, double *b,
                  int ldb);

#ifdef __cplusplus
}
#endif
#endif
