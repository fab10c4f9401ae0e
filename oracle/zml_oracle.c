/*
 * CPU oracle for zml's dense factor-and-solve path (TEST INFRASTRUCTURE ONLY;
 * see zml_oracle.h). Each function restates the cited zml file in C with the
 * same loop order, blocking, recursion split and reciprocal-versus-division
 * choices. Indices are 64-bit so that n = 32768 stays in range.
 *
 * Real types only (f64 as zr_*, f32 as zr_s*, both from zml_oracle_real.inc):
 * conj_no_trans behaves as no_trans and conj_trans as trans, as zml's noconj
 * branches do for real types.
 *
 * Pins (tests/test_oracle.py): zml's own test vectors for gemm, trsm, syrk,
 * gemv, trsv, iamax, scal and swap; the unblocked closed forms of getrf and of
 * column-major lower potrf (bitwise, f64 and f32); hand-solved systems, IEEE
 * constants for lamch, numpy for lange, a planted mistake that must fail.
 * Parity unpinned: the bits of row-major and upper potrf / potrf2, which
 * follow zml's blocked dot-form order with no reference-held vector or closed
 * form to check them against (they are checked to 1e-12 only).
 */
#include "zml_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long long i64;

/* utils.zig:4-19 */
static inline i64 idx(int order, i64 i, i64 j, i64 ld) { return order == ZR_COL ? i + j * ld : i * ld + j; }
static inline int col_ld(int order, int ld) { return order == ZR_COL ? 1 : ld; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

static inline int is_notrans(int t) { return t == ZR_NOTRANS || t == ZR_CONJNOTRANS; }
/* linalg.zig:112-119 Transpose.invert; types.zig:131-136 Uplo.invert */
static inline int trans_invert(int t) {
    switch (t) {
    case ZR_NOTRANS: return ZR_TRANS;
    case ZR_TRANS: return ZR_NOTRANS;
    case ZR_CONJTRANS: return ZR_CONJNOTRANS;
    default: return ZR_CONJTRANS;
    }
}
static inline int uplo_invert(int u) { return u == ZR_UPPER ? ZR_LOWER : ZR_UPPER; }
static inline int side_invert(int s) { return s == ZR_LEFT ? ZR_RIGHT : ZR_LEFT; }

/* ilaenv.zig:143-148 (xGETRF) and :227-233 (xPOTRF): NB = 64. */
int zr_ilaenv_nb(const char* name) {
    (void)name;
    return 64;
}

/* trsm.zig:137-142,221-229 divide b[k + j*lda] in the left no-transpose
 * branches; zr_set_trsm_faithful_lda(0) reads b[k + j*ldb] instead. */
static int g_trsm_faithful_lda = 1;
void zr_set_trsm_faithful_lda(int on) { g_trsm_faithful_lda = on; }

/* ---- f64 instantiation: zr_* ------------------------------------------- */
#define R double
#define F(name) zr_##name
#define RABS fabs
#define RSQRT sqrt
#define RIS_NAN isnan
#define R_EPS_MACH DBL_EPSILON
#define R_MIN_NORM DBL_MIN
#define R_MAX_FIN DBL_MAX
#define R_MANT_BITS (DBL_MANT_DIG - 1)
#define R_EXP_MIN (DBL_MIN_EXP - 1)
#define R_EXP_MAX (DBL_MAX_EXP - 1)
#include "zml_oracle_real.inc"
#undef R
#undef F
#undef RABS
#undef RSQRT
#undef RIS_NAN
#undef R_EPS_MACH
#undef R_MIN_NORM
#undef R_MAX_FIN
#undef R_MANT_BITS
#undef R_EXP_MIN
#undef R_EXP_MAX

/* ---- f32 instantiation: zr_s* (zml's generic routines at T = f32) ------- */
#define R float
#define F(name) zr_s##name
#define RABS fabsf
#define RSQRT sqrtf
#define RIS_NAN isnan
#define R_EPS_MACH FLT_EPSILON
#define R_MIN_NORM FLT_MIN
#define R_MAX_FIN FLT_MAX
#define R_MANT_BITS (FLT_MANT_DIG - 1)
#define R_EXP_MIN (FLT_MIN_EXP - 1)
#define R_EXP_MAX (FLT_MAX_EXP - 1)
#include "zml_oracle_real.inc"
#undef R
#undef F
#undef RABS
#undef RSQRT
#undef RIS_NAN
#undef R_EPS_MACH
#undef R_MIN_NORM
#undef R_MAX_FIN
#undef R_MANT_BITS
#undef R_EXP_MIN
#undef R_EXP_MAX

/* Blocked, threaded form of zr_getrf_unblocked (column-major). Right-looking
 * with panel width NBF: the panel is factored column by column exactly as
 * above; the row swaps are applied to the rest of the matrix; U12 is formed
 * by forward substitution with k ascending; the trailing matrix receives
 * c <- fl(c - fl(l_ik u_kj)) for k ascending inside the block. The per-element
 * operation sequence is therefore unchanged and so are the bits. */
#define NBF 64
int zr_getrf_fast(int m, int n, double* a, int lda, int32_t* ipiv, int nthreads) {
    if (m < 0 || n < 0 || lda < imax(1, m)) return ZR_EINVAL;
    if (m == 0 || n == 0) return 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const double sfmin = zr_lamch('S');
    const i64 ld = lda;
    int info = 0;
    const int mn = imin(m, n);
    for (int j0 = 0; j0 < mn; j0 += NBF) {
        const int jb = imin(NBF, mn - j0);
        /* panel: columns j0..j0+jb-1, rows j0..m-1 */
        for (int k = j0; k < j0 + jb; ++k) {
            int p = k;
            double mx = fabs(a[k + k * ld]);
            for (int i = k + 1; i < m; ++i) {
                const double v = fabs(a[i + k * ld]);
                if (v > mx) { mx = v; p = i; }
            }
            ipiv[k] = p + 1;
            const double piv = a[p + k * ld];
            if (piv != 0.0) {
                if (p != k)
                    for (int j = j0; j < j0 + jb; ++j) { double t = a[k + j * ld]; a[k + j * ld] = a[p + j * ld]; a[p + j * ld] = t; }
                const double d = a[k + k * ld];
                if (fabs(d) >= sfmin) {
                    const double r = 1.0 / d;
                    if (r != 1.0) for (int i = k + 1; i < m; ++i) a[i + k * ld] = a[i + k * ld] * r;
                } else {
                    for (int i = k + 1; i < m; ++i) a[i + k * ld] = a[i + k * ld] / d;
                }
            } else if (info == 0) {
                info = k + 1;
            }
            for (int j = k + 1; j < j0 + jb; ++j) {
                const double u = a[k + j * ld];
                double* cj = a + j * ld;
                const double* lk = a + k * ld;
                for (int i = k + 1; i < m; ++i) cj[i] = cj[i] - lk[i] * u;
            }
        }
        /* swaps on columns outside the panel, U12, trailing update (per column, parallel) */
        const int jr = j0 + jb;
#pragma omp parallel for schedule(dynamic, 8)
        for (int j = 0; j < n; ++j) {
            if (j >= j0 && j < jr) continue;
            double* cj = a + (i64)j * ld;
            for (int k = j0; k < jr; ++k) {
                const int p = ipiv[k] - 1;
                if (p != k) { double t = cj[k]; cj[k] = cj[p]; cj[p] = t; }
            }
            if (j < jr) continue;
            /* forward substitution with the unit lower panel block */
            for (int k = j0; k < jr; ++k) {
                const double u = cj[k];
                const double* lk = a + (i64)k * ld;
                for (int i = k + 1; i < jr; ++i) cj[i] = cj[i] - lk[i] * u;
            }
            /* trailing rows */
            for (int k = j0; k < jr; ++k) {
                const double u = cj[k];
                const double* lk = a + (i64)k * ld;
                for (int i = jr; i < m; ++i) cj[i] = cj[i] - lk[i] * u;
            }
        }
    }
    return info;
}

/* ------------------------------------------------------------------------ */
/* Counter-based generator: a splitmix64 finalizer over (seed, counter).    */
uint64_t zr_rng_u64(uint64_t seed, uint64_t ctr) {
    uint64_t z = seed * 0xD1B54A32D192ED03ull + (ctr + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double zr_rng_uniform(uint64_t seed, uint64_t ctr) {
    return (double)(zr_rng_u64(seed, ctr) >> 11) * 0x1p-53;
}

void zr_fill_uniform(int order, int m, int n, double* a, int ld, uint64_t seed, double lo, double hi) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j)
            a[idx(order, i, j, ld)] = lo + (hi - lo) * zr_rng_uniform(seed, (uint64_t)i * (uint64_t)n + (uint64_t)j);
}

double zr_norm_inf(int order, int m, int n, const double* a, int ld) {
    double mx = 0.0;
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += fabs(a[idx(order, i, j, ld)]);
        if (s > mx || isnan(s)) mx = s;
    }
    return mx;
}

double zr_residual_inf(int order, int n, const double* a, int lda, const double* x, const double* b, double* r) {
    double mx = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += a[idx(order, i, j, lda)] * x[j];
        const double ri = b[i] - s;
        if (r) r[i] = ri;
        if (fabs(ri) > mx || isnan(ri)) mx = fabs(ri);
    }
    return mx;
}

/* Matrix norm of an m x n matrix, LAPACK dlange's definitions (zml has no
 * lange): 'M' max |a_ij|, 'O'/'1' max column sum of |a_ij|, 'I' max row sum,
 * 'F'/'E' Frobenius (sqrt of the plain sum of squares). A NaN entry makes
 * the result NaN; m or n of 0 gives 0; another norm character gives -1. */
double zr_lange(int order, char norm, int m, int n, const double* a, int ld) {
    if (m <= 0 || n <= 0) return 0.0;
    double r = 0.0;
    switch (norm) {
    case 'M': case 'm':
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j) {
                const double v = fabs(a[idx(order, i, j, ld)]);
                if (v > r || isnan(v)) r = v;
                if (isnan(r)) return r;
            }
        return r;
    case 'O': case 'o': case '1':
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += fabs(a[idx(order, i, j, ld)]);
            if (s > r || isnan(s)) r = s;
            if (isnan(r)) return r;
        }
        return r;
    case 'I': case 'i':
        return zr_norm_inf(order, m, n, a, ld);
    case 'F': case 'f': case 'E': case 'e': {
        double s = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j) { const double v = a[idx(order, i, j, ld)]; s += v * v; }
        return sqrt(s);
    }
    default:
        return -1.0;
    }
}

double zr_backward_error(int order, int n, const double* a, int lda, const double* x, const double* b) {
    const double rn = zr_residual_inf(order, n, a, lda, x, b, 0);
    const double an = zr_norm_inf(order, n, n, a, lda);
    double xn = 0.0;
    for (int i = 0; i < n; ++i) if (fabs(x[i]) > xn) xn = fabs(x[i]);
    return rn / (an * xn * (double)n * 0x1p-53);
}

double zr_min_pivot_gap(int order, int n, const double* a0, int lda) {
    double* a = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
    if (!a) return -1.0;
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) a[i + (i64)j * n] = a0[idx(order, i, j, lda)];
    const double sfmin = zr_lamch('S');
    double gap = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double m1 = fabs(a[k + (i64)k * n]), m2 = 0.0;
        for (int i = k + 1; i < n; ++i) {
            const double v = fabs(a[i + (i64)k * n]);
            if (v > m1) { m2 = m1; m1 = v; p = i; } else if (v > m2) { m2 = v; }
        }
        if (m1 > 0.0 && k + 1 < n) { const double g = (m1 - m2) / m1; if (g < gap) gap = g; }
        if (m1 == 0.0) continue;
        if (p != k) for (int j = 0; j < n; ++j) { double t = a[k + (i64)j * n]; a[k + (i64)j * n] = a[p + (i64)j * n]; a[p + (i64)j * n] = t; }
        const double d = a[k + (i64)k * n];
        if (fabs(d) >= sfmin) { const double r = 1.0 / d; for (int i = k + 1; i < n; ++i) a[i + (i64)k * n] *= r; }
        else for (int i = k + 1; i < n; ++i) a[i + (i64)k * n] /= d;
        for (int j = k + 1; j < n; ++j) {
            const double u = a[k + (i64)j * n];
            for (int i = k + 1; i < n; ++i) a[i + (i64)j * n] = a[i + (i64)j * n] - a[i + (i64)k * n] * u;
        }
    }
    free(a);
    return gap;
}
