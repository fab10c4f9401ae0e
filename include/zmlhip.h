/*
 * libzmlhip device-resident API (MI355X / gfx950).
 *
 * Same operations as the CBLAS/LAPACKE exports (cblas.h, clapack.h) on
 * device pointers, ordered on a caller-supplied HIP stream (`stream` is a
 * hipStream_t passed as void*; NULL is the default stream). They return as
 * soon as the work is queued; LAPACK info values are written to device
 * memory. Return value: 0 when the work was queued, -i for an invalid
 * argument i (matrix_layout counted as 1, as in clapack.h), or one of the
 * ZMLHIP_E* codes below (message in zmlhip_last_error()).
 *
 * Enumerations: layout 101/102, trans 111/112/113, uplo 121/122,
 * diag 131/132, side 141/142 (reference include/cblas.h:11-37); the LAPACK
 * style routines take the LAPACKE characters ('N','T','C' / 'U','L').
 */
#ifndef ZMLHIP_H
#define ZMLHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ZMLHIP_EHIP (-1000)      /* a HIP runtime call failed */
#define ZMLHIP_EALLOC (-1001)    /* device allocation failed */
#define ZMLHIP_ESYNC (-1002)     /* a bounded device-side wait timed out */
#define ZMLHIP_ENOTSUP (-1003)   /* shape or option not supported */

/* Calls of one host thread share that thread's device workspaces and internal
 * streams. Each call therefore starts only after the previous call of the
 * same thread has finished on the device, whatever stream either was given:
 * calls queued on two streams by one thread run one after the other. Calls
 * from different host threads use different workspaces and may overlap. */

/* Library version (major*10000 + minor*100 + patch). */
int zmlhip_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char *zmlhip_last_error(void);
/* Waits for all work queued on `stream`; returns 0 or ZMLHIP_EHIP. */
int zmlhip_stream_synchronize(void *stream);

/* Per-launch timing of the dominant kernels, for measurement tools. When
 * enabled, HIP events are recorded around every launch of the kinds below on
 * the stream it runs on; collect sums (and forgets) the records of one kind:
 * total milliseconds, launches, algorithmic flops and bytes.
 * No counterpart in the reference (zml has no GPU kernels to time). */
#define ZMLHIP_PROF_GEMM_F64 1     /* FP64 trailing update of getrf */
#define ZMLHIP_PROF_GEMM_LOWP 2    /* low-precision trailing update of the mixed LU */
#define ZMLHIP_PROF_PANEL 3        /* panel factorization */
int zmlhip_profile_enable(int on);
int zmlhip_profile_collect(int kind, double *ms, long long *launches, double *flops,
                           double *bytes);

/* ---- BLAS (device pointers) ------------------------------------------- */
/* cblas_dgemm (src/cblas.zig:526; blas/gemm.zig:14-1033) */
int zmlhip_dgemm(int layout, int transa, int transb, int m, int n, int k, double alpha,
                 const double *a, int lda, const double *b, int ldb, double beta, double *c,
                 int ldc, void *stream);
/* cblas_dtrsm (src/cblas.zig:636; blas/trsm.zig:18-735) */
int zmlhip_dtrsm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha,
                 const double *a, int lda, double *b, int ldb, void *stream);
/* cblas_dsyrk (src/cblas.zig:587; blas/syrk.zig:15-366) */
int zmlhip_dsyrk(int layout, int uplo, int trans, int n, int k, double alpha, const double *a,
                 int lda, double beta, double *c, int ldc, void *stream);
/* cblas_idamax (src/cblas.zig:247; blas/iamax.zig:11-68): writes the 0-based
 * index to *result (device int). */
int zmlhip_idamax(int n, const double *x, int incx, int *result, void *stream);
/* cblas_dscal (src/cblas.zig:213; blas/scal.zig:9-41) */
int zmlhip_dscal(int n, double alpha, double *x, int incx, void *stream);
/* cblas_dswap (src/cblas.zig:234; blas/swap.zig) */
int zmlhip_dswap(int n, double *x, int incx, double *y, int incy, void *stream);
/* cblas_dgemv (src/cblas.zig:291; blas/gemv.zig:14-360) */
int zmlhip_dgemv(int layout, int transa, int m, int n, double alpha, const double *a, int lda,
                 const double *x, int incx, double beta, double *y, int incy, void *stream);

/* zmlhip_dgemv without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that refine
 * solutions, harnesses that forbid hipMalloc). No counterpart in the
 * reference. Same arguments and results as zmlhip_dgemv (the same kernels;
 * bit-identical results while m * splits fits), but the partial sums of the
 * no-transpose form live in a static per-device array of 2^20 doubles
 * (8 MiB, held by each device the library runs on until the process ends)
 * instead of a library workspace. Returns ZMLHIP_ENOTSUP, with nothing
 * queued, when the no-transpose view (column-major, or the transposed view of
 * a row-major call) has more than 2^20 rows. Calls that overlap on one
 * device share that array and race; order them on one stream. */
int zmlhip_dgemv_noalloc(int layout, int transa, int m, int n, double alpha, const double *a, int lda,
                         const double *x, int incx, double beta, double *y, int incy, void *stream);
/* cblas_dtrsv (src/cblas.zig:512; blas/trsv.zig:16-587) */
int zmlhip_dtrsv(int layout, int uplo, int transa, int diag, int n, const double *a, int lda, double *x,
                 int incx, void *stream);
/* FP32 forms: cblas_sgemm (src/cblas.zig:523), cblas_strsm (:633),
 * cblas_ssyrk (:584); zml's generic gemm/trsm/syrk at T = f32 */
int zmlhip_sgemm(int layout, int transa, int transb, int m, int n, int k, float alpha, const float *a,
                 int lda, const float *b, int ldb, float beta, float *c, int ldc, void *stream);
int zmlhip_strsm(int layout, int side, int uplo, int transa, int diag, int m, int n, float alpha,
                 const float *a, int lda, float *b, int ldb, void *stream);
int zmlhip_ssyrk(int layout, int uplo, int trans, int n, int k, float alpha, const float *a, int lda,
                 float beta, float *c, int ldc, void *stream);
/* LAPACK dlange norms ('M', 'O'/'1', 'I', 'F'/'E') of an m x n matrix into
 * *result (device double). No counterpart in zml (it has no lange). */
int zmlhip_dlange(int layout, char norm, int m, int n, const double *a, int lda, double *result,
                  void *stream);

/* ---- LAPACK (device pointers; ipiv and info are device memory) -------- */
/* LAPACKE_dlaswp (src/linalg/lapack.zig:375; lapack/laswp.zig:12-95) */
int zmlhip_dlaswp(int layout, int n, double *a, int lda, int k1, int k2, const int *ipiv,
                  int incx, void *stream);
/* LAPACKE_dgetrf (src/linalg/lapack.zig:1169; lapack/getrf.zig:14-152).
 * *info: 0, the first zero pivot (1-based), or ZMLHIP_ESYNC when a panel's
 * bounded cross-workgroup wait gave up (the factors are then invalid; the
 * condition is cleared at the start of the next factorization). */
int zmlhip_dgetrf(int layout, int m, int n, double *a, int lda, int *ipiv, int *info,
                  void *stream);
/* LAPACKE_sgetrf (src/linalg/lapack.zig:1160): the same at T = f32 */
int zmlhip_sgetrf(int layout, int m, int n, float *a, int lda, int *ipiv, int *info, void *stream);
/* LAPACKE_spotrf (src/linalg/lapack.zig:3004): potrf at T = f32 */
int zmlhip_spotrf(int layout, char uplo, int n, float *a, int lda, int *info, void *stream);
/* Polls a panel's cross-workgroup wait makes before it gives up (default
 * 2000000; 0 restores the default). A negative value makes every
 * multi-workgroup panel give up at its first wait, with every kernel still
 * ending normally: a test hook for the ZMLHIP_ESYNC path. Process-wide.
 * No counterpart in the reference. */
int zmlhip_set_panel_spin_limit(long polls);
/* FP64 LU block columns with more than `rows` rows (m - j0) are factored in
 * one launch, the others leaf by leaf with GEMM updates in between; both give
 * the same pivots. Default 16384 (the faster choice per height measured on
 * MI355X); a negative value restores it, 0 makes every block column take the
 * one-launch panel. Process-wide. No counterpart in the reference. */
int zmlhip_set_panel_one_launch_rows(long rows);
/* FP64 GEMMs (zmlhip_dgemm and the GEMMs inside the FP64 factorizations)
 * whose shapes allow it move their operand slabs from global memory to LDS
 * with the load unit (on = 1, the default); on = 0 makes every FP64 GEMM stage
 * them through registers. The two kernels give the same results bit for bit.
 * Returns the previous setting. Process-wide. No counterpart in the reference. */
int zmlhip_set_gemm_lds_dma(int on);
/* LU panels of more than one workgroup publish column k + 1's pivot
 * candidates before they finish column k's update (on = 1, the default), or
 * finish every column first (on = 0). Both give the same factors and pivots
 * bit for bit. Returns the previous setting. Process-wide. No counterpart in
 * the reference. */
int zmlhip_set_leaf_lookahead(int on);
/* LAPACKE_dgetrs (src/linalg/lapack.zig:1575; lapack/getrs.zig:13-142) */
int zmlhip_dgetrs(int layout, char trans, int n, int nrhs, const double *a, int lda,
                  const int *ipiv, double *b, int ldb, void *stream);
/* LAPACKE_dpotrf (src/linalg/lapack.zig:3012; lapack/potrf.zig:17-240) */
int zmlhip_dpotrf(int layout, char uplo, int n, double *a, int lda, int *info, void *stream);
/* LAPACKE_dpotrs (src/linalg/lapack.zig:3466; lapack/potrs.zig:15-117) */
int zmlhip_dpotrs(int layout, char uplo, int n, int nrhs, const double *a, int lda, double *b,
                  int ldb, void *stream);

/* ---- mixed precision + iterative refinement ------------------------------ */
/* No counterpart in the reference: zml has no mixed-precision factorization,
 * refinement or GMRES. The result is the FP64 solution of the dgesv/dposv
 * problem (src/linalg/lapack.zig:2051, :3896); A is not modified. */
typedef struct zmlhip_ir_opts {
  int precision;       /* trailing update: 0 = FP32, 1 = FP16 (default), 2 = BF16 */
  int max_iter;        /* classic refinement steps before GMRES-IR (default 30) */
  int gmres_restart;   /* FGMRES restart length (default 30) */
  int gmres_max_iter;  /* total FGMRES iterations (default 300) */
  int no_gmres;        /* 1: on stagnation go straight to the FP64 factorization */
  double tol;          /* stop at ||b-Ax||inf <= tol ||A||inf ||x||inf n u, u = 2^-53 (default 1) */
  int no_optimistic;   /* 1: every block column takes the full pivot search (no diagonal-block try) */
  double pivot_threshold; /* diagonal-block try accepted when every |l_ij| below the block is <= this
                             (default 16; 1 keeps exactly the partial-pivoting choice) */
} zmlhip_ir_opts;

typedef struct zmlhip_ir_result {
  int info;              /* first zero pivot of the low-precision factorization, 0 if none */
  int iterations;        /* classic refinement steps (largest over the right-hand sides) */
  int gmres_iterations;  /* FGMRES iterations (largest over the right-hand sides) */
  int fallback;          /* 0 none, 1 GMRES-IR was needed, 2 the FP64 factorization was used */
  double backward_error; /* ||b-Ax||inf / (||A||inf ||x||inf n u), worst right-hand side */
  int panels;            /* block columns tried with pivots inside the diagonal block */
  int panels_optimistic; /* ... of which needed no pivot from below the block */
} zmlhip_ir_result;

/* A x = b for general A (n x n, device, layout 101/102), nrhs right-hand sides
 * b (ldb), solution into x (ldx, same layout). Synchronous with the host
 * (returns when x is final); work is ordered on `stream`. opts may be NULL.
 * Returns 0; k > 0 when the FP64 factorization that a right-hand side fell
 * back to has an exactly zero pivot at step k (A singular: x is then not a
 * solution; a column-major A is factored as A^T, so k is that factorization's
 * step); -i for an invalid argument; or a ZMLHIP_E* code. result->info
 * reports the low-precision factorization's first zero pivot separately. */
int zmlhip_dgesv_mixed(int layout, int n, int nrhs, const double *a, int lda, const double *b,
                       int ldb, double *x, int ldx, const zmlhip_ir_opts *opts,
                       zmlhip_ir_result *result, void *stream);

/* A x = b for symmetric positive definite A (n x n, device, layout 101/102,
 * only the `uplo` ('L' or 'U') triangle is read): FP32 Cholesky whose
 * diagonal blocks are factored in FP64 and whose trailing updates run in
 * opts->precision, then FP64 refinement of the right-hand sides in groups of
 * 16 (classic, GMRES-IR, FP64 Cholesky as in zmlhip_dgesv_mixed). result->info
 * is the first non-positive pivot of the unshifted low-precision
 * factorization (0 if none); a failure there is answered by factoring again
 * with a small diagonal shift. Synchronous with the host. Returns 0; k > 0
 * when A is not positive definite (LAPACK's dposv info: the unshifted
 * low-precision factorization's first non-positive pivot k when four
 * shifted attempts fail too, x untouched; or that of the FP64 Cholesky a
 * right-hand side fell back to); -i for an invalid argument; or a ZMLHIP_E*
 * code. */
int zmlhip_dposv_mixed(int layout, char uplo, int n, int nrhs, const double *a, int lda,
                       const double *b, int ldb, double *x, int ldx, const zmlhip_ir_opts *opts,
                       zmlhip_ir_result *result, void *stream);

/* Batched solves of many small independent systems (n <= 256 each), FP32
 * factorization + FP64 refinement per system, all in one hipGraph launch.
 * A plan records the systems (host arrays of sizes and of device pointers;
 * every matrix column-major with ld = n, b and x column-major n x nrhs with
 * ld = n; spd[i] = 1 selects elimination without interchanges for a
 * symmetric positive definite A, whose full matrix must be stored), picks the
 * thread-grid shape of each size bucket (by timing the candidates when
 * `autotune` is set) and captures the graph; execute replays it on `stream`.
 * opts: max_iter (default 10) and tol (default 1) are used. status (host,
 * count ints) receives per system the refinement steps of its worst
 * right-hand side, or -1 when the refinement did not reach the target. */
typedef struct zmlhip_batch_plan zmlhip_batch_plan;
int zmlhip_batch_plan_create(zmlhip_batch_plan **plan, int count, const int *n, const int *nrhs,
                             const int *spd, const double *const *a, const double *const *b,
                             double *const *x, const zmlhip_ir_opts *opts, int autotune, void *stream);
int zmlhip_batch_plan_execute(zmlhip_batch_plan *plan, void *stream);
int zmlhip_batch_plan_status(zmlhip_batch_plan *plan, int *status, void *stream);
/* thread rows of the grid, matrix rows per thread (their product bounds the
 * bucket's n) and system count of the 4 size buckets (n <= 32, 64, 128, 256) */
int zmlhip_batch_plan_shapes(const zmlhip_batch_plan *plan, int *tg, int *r, int *count);
int zmlhip_batch_plan_destroy(zmlhip_batch_plan *plan);

#ifdef __cplusplus
}
#endif
#endif
