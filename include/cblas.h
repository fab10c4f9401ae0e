/*
 * CBLAS entry points exported by libzmlhip.so, the MI355X (gfx950) build of
 * zml's dense factor-and-solve path. Each one is a drop-in for the symbol of
 * the same name that zml's own `zig build cblas` library exports, with the
 * same signature and the same enumeration values (reference
 * include/cblas.h:11-37). Host pointers, synchronous: results are in host
 * memory when the call returns. Invalid arguments are ignored silently, as
 * zml's exports do (`catch {}` in src/linalg/blas.zig:14156).
 *
 * Device-resident, stream-ordered versions are in zmlhip.h.
 */
#ifndef ZMLHIP_CBLAS_H
#define ZMLHIP_CBLAS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum CBLAS_ORDER { CblasRowMajor = 101, CblasColMajor = 102 } CBLAS_ORDER;
typedef enum CBLAS_TRANSPOSE {
    CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113, CblasConjNoTrans = 114
} CBLAS_TRANSPOSE;
typedef enum CBLAS_UPLO { CblasUpper = 121, CblasLower = 122 } CBLAS_UPLO;
typedef enum CBLAS_DIAG { CblasNonUnit = 131, CblasUnit = 132 } CBLAS_DIAG;
typedef enum CBLAS_SIDE { CblasLeft = 141, CblasRight = 142 } CBLAS_SIDE;
typedef CBLAS_ORDER CBLAS_LAYOUT;

/* replaces src/cblas.zig:213 (include/cblas.h:96) */
void cblas_dscal(const int n, const double alpha, double *x, const int incx);
/* replaces src/cblas.zig:234 (include/cblas.h:103) */
void cblas_dswap(const int n, double *x, const int incx, double *y, const int incy);
/* replaces src/cblas.zig:247 (include/cblas.h:108); 0-based, first index on ties */
unsigned int cblas_idamax(const int n, const double *x, const int incx);
/* replaces src/cblas.zig:291 (include/cblas.h:127) */
void cblas_dgemv(const CBLAS_ORDER order, const CBLAS_TRANSPOSE transa, const int m, const int n,
                 const double alpha, const double *a, const int lda, const double *x, const int incx,
                 const double beta, double *y, const int incy);
/* replaces src/cblas.zig:512 (include/cblas.h:208); divides by the diagonal
 * as trsv.zig does; incx may be negative */
void cblas_dtrsv(const CBLAS_ORDER order, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                 const CBLAS_DIAG diag, const int n, const double *a, const int lda, double *x,
                 const int incx);
/* replaces src/cblas.zig:523 (include/cblas.h:216): FP32, FP32 MFMA */
void cblas_sgemm(const CBLAS_ORDER order, const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                 const int m, const int n, const int k, const float alpha, const float *a,
                 const int lda, const float *b, const int ldb, const float beta, float *c,
                 const int ldc);
/* replaces src/cblas.zig:584 (include/cblas.h:235) */
void cblas_ssyrk(const CBLAS_ORDER order, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE trans,
                 const int n, const int k, const float alpha, const float *a, const int lda,
                 const float beta, float *c, const int ldc);
/* replaces src/cblas.zig:633 (include/cblas.h:250) */
void cblas_strsm(const CBLAS_ORDER order, const CBLAS_SIDE side, const CBLAS_UPLO uplo,
                 const CBLAS_TRANSPOSE transa, const CBLAS_DIAG diag, const int m, const int n,
                 const float alpha, const float *a, const int lda, float *b, const int ldb);
/* replaces src/cblas.zig:526 (include/cblas.h:217) */
void cblas_dgemm(const CBLAS_ORDER order, const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                 const int m, const int n, const int k, const double alpha, const double *a,
                 const int lda, const double *b, const int ldb, const double beta, double *c,
                 const int ldc);
/* replaces src/cblas.zig:587 (include/cblas.h:236) */
void cblas_dsyrk(const CBLAS_ORDER order, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE trans,
                 const int n, const int k, const double alpha, const double *a, const int lda,
                 const double beta, double *c, const int ldc);
/* replaces src/cblas.zig:636 (include/cblas.h:251) */
void cblas_dtrsm(const CBLAS_ORDER order, const CBLAS_SIDE side, const CBLAS_UPLO uplo,
                 const CBLAS_TRANSPOSE transa, const CBLAS_DIAG diag, const int m, const int n,
                 const double alpha, const double *a, const int lda, double *b, const int ldb);

#ifdef __cplusplus
}
#endif
#endif
