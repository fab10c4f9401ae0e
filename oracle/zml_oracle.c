/*
 * CPU oracle for zml's dense factor-and-solve path (TEST INFRASTRUCTURE ONLY;
 * see zml_oracle.h). Each function restates the cited zml file in C with the
 * same loop order, blocking, recursion split and reciprocal-versus-division
 * choices. Indices are 64-bit so that n = 32768 stays in range.
 *
 * Real f64 only: conj_no_trans behaves as no_trans and conj_trans as trans,
 * as zml's noconj branches do for real types.
 */
#include "zml_oracle.h"

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

/* ------------------------------------------------------------------------ */
/* lamch.zig:15-78 with rnd = 1: eps = 2^-53, sfmin = tiny (1/huge < tiny).  */
double zr_lamch(char c) {
    const double eps = 0x1p-53;
    switch (c) {
    case 'E': case 'e': return eps;
    case 'S': case 's': {
        double sfmin = 0x1p-1022;
        double small = 1.0 / 0x1.fffffffffffffp+1023;
        if (small >= sfmin) sfmin = small * (1.0 + eps);
        return sfmin;
    }
    case 'B': case 'b': return 2.0;
    case 'P': case 'p': return eps * 2.0;
    case 'N': case 'n': return 53.0;
    case 'R': case 'r': return 1.0;
    case 'M': case 'm': return -1021.0;
    case 'U': case 'u': return 0x1p-1022;
    case 'L': case 'l': return 1024.0;
    case 'O': case 'o': return 0x1.fffffffffffffp+1023;
    default: return 0.0;
    }
}

/* ilaenv.zig:143-148 (xGETRF) and :227-233 (xPOTRF): NB = 64. */
int zr_ilaenv_nb(const char* name) {
    (void)name;
    return 64;
}

/* ------------------------------------------------------------------------ */
/* iamax.zig:11-68: first index of max |x| (strict >), 0-based.             */
int zr_iamax(int n, const double* x, int incx) {
    if (n <= 0 || incx <= 0) return ZR_EINVAL;
    if (n == 1) return 0;
    int imaxi = 0;
    double mx = fabs(x[0]);
    i64 ix = incx;
    for (int i = 1; i < n; ++i) {
        double ax = fabs(x[ix]);
        if (ax > mx) { mx = ax; imaxi = i; }
        ix += incx;
    }
    return imaxi;
}

/* scal.zig:9-41: skips alpha == 1. */
int zr_scal(int n, double alpha, double* x, int incx) {
    if (n < 0 || incx <= 0) return ZR_EINVAL;
    if (n == 0 || alpha == 1.0) return 0;
    i64 ix = 0;
    for (int i = 0; i < n; ++i) { x[ix] = x[ix] * alpha; ix += incx; }
    return 0;
}

/* swap.zig */
int zr_swap(int n, double* x, int incx, double* y, int incy) {
    if (n < 0) return ZR_EINVAL;
    i64 ix = incx < 0 ? (i64)(-n + 1) * incx : 0;
    i64 iy = incy < 0 ? (i64)(-n + 1) * incy : 0;
    for (int i = 0; i < n; ++i) {
        double t = x[ix]; x[ix] = y[iy]; y[iy] = t;
        ix += incx; iy += incy;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* gemm.zig:68-1033 (k_gemm, column-major view).                            */
static int k_gemm(int transa, int transb, int m, int n, int k, double alpha,
                  const double* a, i64 lda, const double* b, i64 ldb, double beta,
                  double* c, i64 ldc) {
    const int nota = is_notrans(transa), notb = is_notrans(transb);
    const int nrowa = nota ? m : k, nrowb = notb ? k : n;
    if (m < 0 || n < 0 || k < 0 || lda < imax(1, nrowa) || ldb < imax(1, nrowb) || ldc < imax(1, m))
        return ZR_EINVAL;
    if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return 0;
    if (alpha == 0.0) { /* gemm.zig:110-140 */
        for (i64 j = 0; j < n; ++j)
            for (i64 i = 0; i < m; ++i)
                c[i + j * ldc] = beta == 0.0 ? 0.0 : c[i + j * ldc] * beta;
        return 0;
    }
    if (notb && nota) { /* gemm.zig:142-191: axpy form */
        for (i64 j = 0; j < n; ++j) {
            if (beta == 0.0) { for (i64 i = 0; i < m; ++i) c[i + j * ldc] = 0.0; }
            else if (beta != 1.0) { for (i64 i = 0; i < m; ++i) c[i + j * ldc] = c[i + j * ldc] * beta; }
            for (i64 l = 0; l < k; ++l) {
                const double temp = alpha * b[l + j * ldb];
                for (i64 i = 0; i < m; ++i) c[i + j * ldc] = c[i + j * ldc] + temp * a[i + l * lda];
            }
        }
    } else if (notb) { /* gemm.zig:337-398: dot form, A transposed */
        for (i64 j = 0; j < n; ++j)
            for (i64 i = 0; i < m; ++i) {
                double temp = 0.0;
                for (i64 l = 0; l < k; ++l) temp = temp + a[l + i * lda] * b[l + j * ldb];
                if (beta == 0.0) c[i + j * ldc] = alpha * temp;
                else if (beta == 1.0) c[i + j * ldc] = c[i + j * ldc] + alpha * temp;
                else { c[i + j * ldc] = c[i + j * ldc] * beta; c[i + j * ldc] = c[i + j * ldc] + alpha * temp; }
            }
    } else if (nota) { /* gemm.zig:584-633: axpy form, B transposed */
        for (i64 j = 0; j < n; ++j) {
            if (beta == 0.0) { for (i64 i = 0; i < m; ++i) c[i + j * ldc] = 0.0; }
            else if (beta != 1.0) { for (i64 i = 0; i < m; ++i) c[i + j * ldc] = c[i + j * ldc] * beta; }
            for (i64 l = 0; l < k; ++l) {
                const double temp = b[j + l * ldb] * alpha;
                for (i64 i = 0; i < m; ++i) c[i + j * ldc] = c[i + j * ldc] + a[i + l * lda] * temp;
            }
        }
    } else { /* gemm.zig:780-900: dot form, both transposed */
        for (i64 j = 0; j < n; ++j)
            for (i64 i = 0; i < m; ++i) {
                double temp = 0.0;
                for (i64 l = 0; l < k; ++l) temp = temp + a[l + i * lda] * b[j + l * ldb];
                if (beta == 0.0) c[i + j * ldc] = alpha * temp;
                else if (beta == 1.0) c[i + j * ldc] = c[i + j * ldc] + alpha * temp;
                else { c[i + j * ldc] = c[i + j * ldc] * beta; c[i + j * ldc] = c[i + j * ldc] + alpha * temp; }
            }
    }
    return 0;
}

/* gemm.zig:14-65: row-major swaps (A, B) and (m, n). */
int zr_gemm(int order, int transa, int transb, int m, int n, int k, double alpha,
            const double* a, int lda, const double* b, int ldb, double beta, double* c, int ldc) {
    if (order == ZR_COL) return k_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    return k_gemm(transb, transa, n, m, k, alpha, b, ldb, a, lda, beta, c, ldc);
}

/* ------------------------------------------------------------------------ */
/* trsm.zig:66-735 (k_trsm, column-major view).                             */
static int g_trsm_faithful_lda = 1;
void zr_set_trsm_faithful_lda(int on) { g_trsm_faithful_lda = on; }

static int k_trsm(int side, int uplo, int transa, int diag, int m, int n, double alpha,
                  const double* a, i64 lda, double* b, i64 ldb) {
    const int nt = is_notrans(transa), nonunit = diag == ZR_NONUNIT;
    const int nrowa = side == ZR_LEFT ? m : n;
    /* the diagonal divide in the left/no-trans branches reads b[k + j*lda] */
    const i64 ldd = g_trsm_faithful_lda ? lda : ldb;
    if (m < 0 || n < 0 || lda < imax(1, nrowa) || ldb < imax(1, m)) return ZR_EINVAL;
    if (m == 0 || n == 0) return 0;
    if (alpha == 0.0) { /* trsm.zig:99-113 */
        for (i64 j = 0; j < n; ++j) for (i64 i = 0; i < m; ++i) b[i + j * ldb] = 0.0;
        return 0;
    }
    if (side == ZR_LEFT) {
        if (nt) {
            if (uplo == ZR_UPPER) { /* trsm.zig:117-159 */
                for (i64 j = 0; j < n; ++j) {
                    if (alpha != 1.0) for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] * alpha;
                    for (i64 k = (i64)m - 1; k >= 0; --k) {
                        if (b[k + j * ldb] != 0.0) {
                            if (nonunit) b[k + j * ldb] = b[k + j * ldd] / a[k + k * lda];
                            for (i64 i = 0; i < k; ++i) b[i + j * ldb] = b[i + j * ldb] - b[k + j * ldb] * a[i + k * lda];
                        }
                    }
                }
            } else { /* trsm.zig:202-270 */
                for (i64 j = 0; j < n; ++j) {
                    if (alpha != 1.0) for (i64 i = 0; i < m; ++i) b[i + j * ldb] = alpha * b[i + j * ldb];
                    for (i64 k = 0; k < m; ++k) {
                        if (b[k + j * ldb] != 0.0) {
                            if (nonunit) b[k + j * ldb] = b[k + j * ldd] / a[k + k * lda];
                            for (i64 i = k + 1; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] - b[k + j * ldb] * a[i + k * lda];
                        }
                    }
                }
            }
        } else {
            if (uplo == ZR_UPPER) { /* trsm.zig:272-337 */
                for (i64 j = 0; j < n; ++j)
                    for (i64 i = 0; i < m; ++i) {
                        double temp = b[i + j * ldb] * alpha;
                        for (i64 k = 0; k < i; ++k) temp = temp - a[k + i * lda] * b[k + j * ldb];
                        if (nonunit) temp = temp / a[i + i * lda];
                        b[i + j * ldb] = temp;
                    }
            } else { /* trsm.zig:338-404 */
                for (i64 j = 0; j < n; ++j)
                    for (i64 i = (i64)m - 1; i >= 0; --i) {
                        double temp = b[i + j * ldb] * alpha;
                        for (i64 k = i + 1; k < m; ++k) temp = temp - a[k + i * lda] * b[k + j * ldb];
                        if (nonunit) temp = temp / a[i + i * lda];
                        b[i + j * ldb] = temp;
                    }
            }
        }
    } else {
        if (nt) {
            if (uplo == ZR_UPPER) { /* trsm.zig:408-487 */
                for (i64 j = 0; j < n; ++j) {
                    if (alpha != 1.0) for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] * alpha;
                    for (i64 k = 0; k < j; ++k)
                        if (a[k + j * lda] != 0.0)
                            for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] - a[k + j * lda] * b[i + k * ldb];
                    if (nonunit) {
                        const double temp = 1.0 / a[j + j * lda];
                        for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] * temp;
                    }
                }
            } else { /* trsm.zig:488-568 */
                for (i64 j = (i64)n - 1; j >= 0; --j) {
                    if (alpha != 1.0) for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] * alpha;
                    for (i64 k = j + 1; k < n; ++k)
                        if (a[k + j * lda] != 0.0)
                            for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] - a[k + j * lda] * b[i + k * ldb];
                    if (nonunit) {
                        const double temp = 1.0 / a[j + j * lda];
                        for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] * temp;
                    }
                }
            }
        } else {
            if (uplo == ZR_UPPER) { /* trsm.zig:570-647 */
                for (i64 k = (i64)n - 1; k >= 0; --k) {
                    if (nonunit) {
                        const double temp = 1.0 / a[k + k * lda];
                        for (i64 i = 0; i < m; ++i) b[i + k * ldb] = b[i + k * ldb] * temp;
                    }
                    for (i64 j = 0; j < k; ++j)
                        if (a[j + k * lda] != 0.0) {
                            const double temp = a[j + k * lda];
                            for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] - temp * b[i + k * ldb];
                        }
                    if (alpha != 1.0) for (i64 i = 0; i < m; ++i) b[i + k * ldb] = b[i + k * ldb] * alpha;
                }
            } else { /* trsm.zig:648-727 */
                for (i64 k = 0; k < n; ++k) {
                    if (nonunit) {
                        const double temp = 1.0 / a[k + k * lda];
                        for (i64 i = 0; i < m; ++i) b[i + k * ldb] = b[i + k * ldb] * temp;
                    }
                    for (i64 j = k + 1; j < n; ++j)
                        if (a[j + k * lda] != 0.0) {
                            const double temp = a[j + k * lda];
                            for (i64 i = 0; i < m; ++i) b[i + j * ldb] = b[i + j * ldb] - temp * b[i + k * ldb];
                        }
                    if (alpha != 1.0) for (i64 i = 0; i < m; ++i) b[i + k * ldb] = b[i + k * ldb] * alpha;
                }
            }
        }
    }
    return 0;
}

/* trsm.zig:18-63: row-major inverts side and uplo and swaps m, n. */
int zr_trsm(int order, int side, int uplo, int transa, int diag, int m, int n,
            double alpha, const double* a, int lda, double* b, int ldb) {
    if (order == ZR_COL) return k_trsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
    return k_trsm(side_invert(side), uplo_invert(uplo), transa, diag, n, m, alpha, a, lda, b, ldb);
}

/* ------------------------------------------------------------------------ */
/* syrk.zig:60-366 (k_syrk, column-major view).                             */
static int k_syrk(int uplo, int trans, int n, int k, double alpha, const double* a, i64 lda,
                  double beta, double* c, i64 ldc) {
    const int nrowa = trans == ZR_NOTRANS ? n : k;
    if (trans == ZR_CONJNOTRANS || trans == ZR_CONJTRANS || n < 0 || k < 0 ||
        lda < imax(1, nrowa) || ldc < imax(1, n))
        return ZR_EINVAL;
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return 0;
    const int up = uplo == ZR_UPPER;
    if (alpha == 0.0) { /* syrk.zig:93-150 */
        for (i64 j = 0; j < n; ++j) {
            const i64 i0 = up ? 0 : j, i1 = up ? j + 1 : n;
            for (i64 i = i0; i < i1; ++i) c[i + j * ldc] = beta == 0.0 ? 0.0 : c[i + j * ldc] * beta;
        }
        return 0;
    }
    if (trans == ZR_NOTRANS) { /* syrk.zig:153-252: axpy form */
        for (i64 j = 0; j < n; ++j) {
            const i64 i0 = up ? 0 : j, i1 = up ? j + 1 : n;
            if (beta == 0.0) { for (i64 i = i0; i < i1; ++i) c[i + j * ldc] = 0.0; }
            else if (beta != 1.0) { for (i64 i = i0; i < i1; ++i) c[i + j * ldc] = c[i + j * ldc] * beta; }
            for (i64 l = 0; l < k; ++l)
                if (a[j + l * lda] != 0.0) {
                    const double temp = alpha * a[j + l * lda];
                    for (i64 i = i0; i < i1; ++i) c[i + j * ldc] = c[i + j * ldc] + temp * a[i + l * lda];
                }
        }
    } else { /* syrk.zig:253-359: dot form */
        for (i64 j = 0; j < n; ++j) {
            const i64 i0 = up ? 0 : j, i1 = up ? j + 1 : n;
            for (i64 i = i0; i < i1; ++i) {
                double temp = 0.0;
                for (i64 l = 0; l < k; ++l) temp = temp + a[l + i * lda] * a[l + j * lda];
                if (beta == 0.0) c[i + j * ldc] = alpha * temp;
                else if (beta == 1.0) c[i + j * ldc] = c[i + j * ldc] + alpha * temp;
                else { c[i + j * ldc] = c[i + j * ldc] * beta; c[i + j * ldc] = c[i + j * ldc] + alpha * temp; }
            }
        }
    }
    return 0;
}

/* syrk.zig:15-56: row-major inverts uplo and trans. */
int zr_syrk(int order, int uplo, int trans, int n, int k, double alpha, const double* a, int lda,
            double beta, double* c, int ldc) {
    if (order == ZR_COL) return k_syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
    return k_syrk(uplo_invert(uplo), trans_invert(trans), n, k, alpha, a, lda, beta, c, ldc);
}

/* ------------------------------------------------------------------------ */
/* gemv.zig:62-360 (k_gemv, column-major view); row-major inverts trans.    */
static int k_gemv(int trans, int m, int n, double alpha, const double* a, i64 lda,
                  const double* x, int incx, double beta, double* y, int incy) {
    if (m < 0 || n < 0 || lda < imax(1, m) || incx == 0 || incy == 0) return ZR_EINVAL;
    if (m == 0 || n == 0 || (alpha == 0.0 && beta == 1.0)) return 0;
    const int nt = is_notrans(trans);
    const int lenx = nt ? n : m, leny = nt ? m : n;
    const i64 kx = incx < 0 ? (i64)(-lenx + 1) * incx : 0;
    const i64 ky = incy < 0 ? (i64)(-leny + 1) * incy : 0;
    if (beta != 1.0) {
        i64 iy = ky;
        for (int i = 0; i < leny; ++i) { y[iy] = beta == 0.0 ? 0.0 : y[iy] * beta; iy += incy; }
    }
    if (alpha == 0.0) return 0;
    if (nt) {
        i64 jx = kx;
        for (i64 j = 0; j < n; ++j) {
            const double temp = alpha * x[jx];
            i64 iy = ky;
            for (i64 i = 0; i < m; ++i) { y[iy] = y[iy] + temp * a[i + j * lda]; iy += incy; }
            jx += incx;
        }
    } else {
        i64 jy = ky;
        for (i64 j = 0; j < n; ++j) {
            double temp = 0.0;
            i64 ix = kx;
            for (i64 i = 0; i < m; ++i) { temp = temp + a[i + j * lda] * x[ix]; ix += incx; }
            y[jy] = y[jy] + alpha * temp;
            jy += incy;
        }
    }
    return 0;
}

int zr_gemv(int order, int trans, int m, int n, double alpha, const double* a, int lda,
            const double* x, int incx, double beta, double* y, int incy) {
    if (order == ZR_COL) return k_gemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
    return k_gemv(trans_invert(trans), n, m, alpha, a, lda, x, incx, beta, y, incy);
}

/* ------------------------------------------------------------------------ */
/* laswp.zig:12-95: interchanges in 32-column chunks, ipiv 1-based.         */
int zr_laswp(int order, int n, double* a, int lda, int k1, int k2, const int32_t* ipiv, int incx) {
    if (n < 0 || (order == ZR_ROW && lda < imax(1, n)) || incx == 0) return ZR_EINVAL;
    if (n == 0) return 0;
    int ix0, i1, i2, inc;
    if (incx > 0) { ix0 = k1 - 1; i1 = k1 - 1; i2 = k2 - 1; inc = 1; }
    else { ix0 = k1 - 1 + (k1 - k2) * incx; i1 = k2 - 1; i2 = k1 - 1; inc = -1; }
    const int n32 = (n >> 5) << 5;
    for (int j = 0; j < n32; j += 32) {
        int ix = ix0;
        for (int i = i1; i != i2 + inc; i += inc) {
            const int ip = ipiv[ix] - 1;
            if (ip != i)
                for (int k = j; k < j + 32; ++k) {
                    double t = a[idx(order, i, k, lda)];
                    a[idx(order, i, k, lda)] = a[idx(order, ip, k, lda)];
                    a[idx(order, ip, k, lda)] = t;
                }
            ix += incx;
        }
    }
    if (n32 != n) {
        int ix = ix0;
        for (int i = i1; i != i2 + inc; i += inc) {
            const int ip = ipiv[ix] - 1;
            if (ip != i)
                for (int k = n32; k < n; ++k) {
                    double t = a[idx(order, i, k, lda)];
                    a[idx(order, i, k, lda)] = a[idx(order, ip, k, lda)];
                    a[idx(order, ip, k, lda)] = t;
                }
            ix += incx;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* getrf2.zig:14-210: recursive panel LU.                                   */
int zr_getrf2(int order, int m, int n, double* a, int lda, int32_t* ipiv) {
    if (m < 0 || n < 0 || lda < imax(1, order == ZR_COL ? m : n)) return ZR_EINVAL;
    int info = 0;
    if (m == 0 || n == 0) return 0;
    if (m == 1) { /* getrf2.zig:35-41 */
        ipiv[0] = 1;
        if (a[0] == 0.0) info = 1;
    } else if (n == 1) { /* getrf2.zig:41-90 */
        const double sfmin = zr_lamch('S');
        const int ii = zr_iamax(m, a, col_ld(order, lda));
        ipiv[0] = ii + 1;
        if (a[idx(order, ii, 0, lda)] != 0.0) {
            if (ii != 0) {
                double t = a[0];
                a[0] = a[idx(order, ii, 0, lda)];
                a[idx(order, ii, 0, lda)] = t;
            }
            if (fabs(a[0]) >= sfmin) {
                zr_scal(m - 1, 1.0 / a[0], a + idx(order, 1, 0, lda), col_ld(order, lda));
            } else {
                for (int i = 0; i < m - 1; ++i) a[idx(order, i + 1, 0, lda)] = a[idx(order, i + 1, 0, lda)] / a[0];
            }
        } else {
            info = 1;
        }
    } else { /* getrf2.zig:93-202 */
        const int n1 = imin(m, n) / 2, n2 = n - n1;
        int iinfo = zr_getrf2(order, m, n1, a, lda, ipiv);
        if (info == 0 && iinfo > 0) info = iinfo;
        zr_laswp(order, n2, a + idx(order, 0, n1, lda), lda, 1, n1, ipiv, 1);
        zr_trsm(order, ZR_LEFT, ZR_LOWER, ZR_NOTRANS, ZR_UNIT, n1, n2, 1.0, a, lda,
                a + idx(order, 0, n1, lda), lda);
        zr_gemm(order, ZR_NOTRANS, ZR_NOTRANS, m - n1, n2, n1, -1.0, a + idx(order, n1, 0, lda), lda,
                a + idx(order, 0, n1, lda), lda, 1.0, a + idx(order, n1, n1, lda), lda);
        iinfo = zr_getrf2(order, m - n1, n2, a + idx(order, n1, n1, lda), lda, ipiv + n1);
        if (info == 0 && iinfo > 0) info = iinfo + n1;
        for (int i = n1; i < imin(m, n); ++i) ipiv[i] += n1;
        zr_laswp(order, n1, a, lda, n1 + 1, imin(m, n), ipiv, 1);
    }
    return info;
}

/* getrf.zig:14-152: blocked right-looking LU, NB = 64. */
int zr_getrf(int order, int m, int n, double* a, int lda, int32_t* ipiv) {
    if (m < 0 || n < 0 || lda < imax(1, order == ZR_COL ? m : n)) return ZR_EINVAL;
    int info = 0;
    if (m == 0 || n == 0) return 0;
    const int nb = zr_ilaenv_nb("DGETRF");
    const int mn = imin(m, n);
    if (nb <= 1 || nb >= mn) return zr_getrf2(order, m, n, a, lda, ipiv);
    for (int j = 0; j < mn; j += nb) {
        const int jb = imin(mn - j, nb);
        const int iinfo = zr_getrf2(order, m - j, jb, a + idx(order, j, j, lda), lda, ipiv + j);
        if (info == 0 && iinfo > 0) info = iinfo + j;
        for (int i = j; i < imin(m, j + jb); ++i) ipiv[i] += j;
        zr_laswp(order, j, a, lda, j + 1, j + jb, ipiv, 1);
        if (j + jb < n) {
            zr_laswp(order, n - j - jb, a + idx(order, 0, j + jb, lda), lda, j + 1, j + jb, ipiv, 1);
            zr_trsm(order, ZR_LEFT, ZR_LOWER, ZR_NOTRANS, ZR_UNIT, jb, n - j - jb, 1.0,
                    a + idx(order, j, j, lda), lda, a + idx(order, j, j + jb, lda), lda);
            if (j + jb < m)
                zr_gemm(order, ZR_NOTRANS, ZR_NOTRANS, m - j - jb, n - j - jb, jb, -1.0,
                        a + idx(order, j + jb, j, lda), lda, a + idx(order, j, j + jb, lda), lda, 1.0,
                        a + idx(order, j + jb, j + jb, lda), lda);
        }
    }
    return info;
}

/* getrs.zig:13-142 */
int zr_getrs(int order, int trans, int n, int nrhs, const double* a, int lda, const int32_t* ipiv,
             double* b, int ldb) {
    if (n < 0 || nrhs < 0 || lda < imax(1, n) || ldb < imax(1, order == ZR_COL ? n : nrhs)) return ZR_EINVAL;
    if (n == 0 || nrhs == 0) return 0;
    if (is_notrans(trans)) {
        zr_laswp(order, nrhs, b, ldb, 1, n, ipiv, 1);
        zr_trsm(order, ZR_LEFT, ZR_LOWER, trans, ZR_UNIT, n, nrhs, 1.0, a, lda, b, ldb);
        zr_trsm(order, ZR_LEFT, ZR_UPPER, trans, ZR_NONUNIT, n, nrhs, 1.0, a, lda, b, ldb);
    } else {
        zr_trsm(order, ZR_LEFT, ZR_UPPER, trans, ZR_NONUNIT, n, nrhs, 1.0, a, lda, b, ldb);
        zr_trsm(order, ZR_LEFT, ZR_LOWER, trans, ZR_UNIT, n, nrhs, 1.0, a, lda, b, ldb);
        zr_laswp(order, nrhs, b, ldb, 1, n, ipiv, -1);
    }
    return 0;
}

/* gesv.zig:13-70 */
int zr_gesv(int order, int n, int nrhs, double* a, int lda, int32_t* ipiv, double* b, int ldb) {
    if (n < 0 || nrhs < 0 || lda < imax(1, n) || ldb < imax(1, order == ZR_COL ? n : nrhs)) return ZR_EINVAL;
    const int info = zr_getrf(order, n, n, a, lda, ipiv);
    if (info == 0) zr_getrs(order, ZR_NOTRANS, n, nrhs, a, lda, ipiv, b, ldb);
    return info;
}

/* ------------------------------------------------------------------------ */
/* potrf2.zig:17-217: recursive Cholesky, n1 = n/2.                         */
int zr_potrf2(int order, int uplo, int n, double* a, int lda) {
    if (n < 0 || lda < imax(1, n)) return ZR_EINVAL;
    if (n == 0) return 0;
    if (n == 1) { /* potrf2.zig:37-53; sqrt is correctly rounded (float/sqrt.zig:31-41) */
        const double ajj = a[0];
        if (ajj <= 0.0 || isnan(ajj)) return 1;
        a[0] = sqrt(ajj);
        return 0;
    }
    const int n1 = n / 2, n2 = n - n1;
    int iinfo = zr_potrf2(order, uplo, n1, a, lda);
    if (iinfo != 0) return iinfo;
    if (uplo == ZR_UPPER) { /* potrf2.zig:75-141 */
        zr_trsm(order, ZR_LEFT, ZR_UPPER, ZR_CONJTRANS, ZR_NONUNIT, n1, n2, 1.0, a, lda,
                a + idx(order, 0, n1, lda), lda);
        zr_syrk(order, uplo, ZR_TRANS, n2, n1, -1.0, a + idx(order, 0, n1, lda), lda, 1.0,
                a + idx(order, n1, n1, lda), lda);
    } else { /* potrf2.zig:142-209 */
        zr_trsm(order, ZR_RIGHT, ZR_LOWER, ZR_CONJTRANS, ZR_NONUNIT, n2, n1, 1.0, a, lda,
                a + idx(order, n1, 0, lda), lda);
        zr_syrk(order, uplo, ZR_NOTRANS, n2, n1, -1.0, a + idx(order, n1, 0, lda), lda, 1.0,
                a + idx(order, n1, n1, lda), lda);
    }
    iinfo = zr_potrf2(order, uplo, n2, a + idx(order, n1, n1, lda), lda);
    if (iinfo != 0) return iinfo + n1;
    return 0;
}

/* potrf.zig:17-240: blocked left-looking Cholesky, NB = 64. */
int zr_potrf(int order, int uplo, int n, double* a, int lda) {
    if (n < 0 || lda < imax(1, n)) return ZR_EINVAL;
    if (n == 0) return 0;
    const int nb = zr_ilaenv_nb("DPOTRF");
    if (nb <= 1 || nb >= n) return zr_potrf2(order, uplo, n, a, lda);
    for (int j = 0; j < n; j += nb) {
        const int jb = imin(nb, n - j);
        int info;
        if (uplo == ZR_UPPER) { /* potrf.zig:52-141 */
            zr_syrk(order, ZR_UPPER, ZR_TRANS, jb, j, -1.0, a + idx(order, 0, j, lda), lda, 1.0,
                    a + idx(order, j, j, lda), lda);
            info = zr_potrf2(order, ZR_UPPER, jb, a + idx(order, j, j, lda), lda);
            if (info != 0) return info + j;
            if (j + jb < n) {
                zr_gemm(order, ZR_CONJTRANS, ZR_NOTRANS, jb, n - j - jb, j, -1.0, a + idx(order, 0, j, lda), lda,
                        a + idx(order, 0, j + jb, lda), lda, 1.0, a + idx(order, j, j + jb, lda), lda);
                zr_trsm(order, ZR_LEFT, ZR_UPPER, ZR_CONJTRANS, ZR_NONUNIT, jb, n - j - jb, 1.0,
                        a + idx(order, j, j, lda), lda, a + idx(order, j, j + jb, lda), lda);
            }
        } else { /* potrf.zig:144-231 */
            zr_syrk(order, ZR_LOWER, ZR_NOTRANS, jb, j, -1.0, a + idx(order, j, 0, lda), lda, 1.0,
                    a + idx(order, j, j, lda), lda);
            info = zr_potrf2(order, ZR_LOWER, jb, a + idx(order, j, j, lda), lda);
            if (info != 0) return info + j;
            if (j + jb < n) {
                zr_gemm(order, ZR_NOTRANS, ZR_CONJTRANS, n - j - jb, jb, j, -1.0, a + idx(order, j + jb, 0, lda), lda,
                        a + idx(order, j, 0, lda), lda, 1.0, a + idx(order, j + jb, j, lda), lda);
                zr_trsm(order, ZR_RIGHT, ZR_LOWER, ZR_CONJTRANS, ZR_NONUNIT, n - j - jb, jb, 1.0,
                        a + idx(order, j, j, lda), lda, a + idx(order, j + jb, j, lda), lda);
            }
        }
    }
    return 0;
}

/* potrs.zig:15-117 (ldb checked against n in both layouts, as zml does). */
int zr_potrs(int order, int uplo, int n, int nrhs, const double* a, int lda, double* b, int ldb) {
    if (n < 0 || nrhs < 0 || lda < imax(1, n) || ldb < imax(1, n)) return ZR_EINVAL;
    if (n == 0 || nrhs == 0) return 0;
    if (uplo == ZR_UPPER) {
        zr_trsm(order, ZR_LEFT, ZR_UPPER, ZR_CONJTRANS, ZR_NONUNIT, n, nrhs, 1.0, a, lda, b, ldb);
        zr_trsm(order, ZR_LEFT, ZR_UPPER, ZR_NOTRANS, ZR_NONUNIT, n, nrhs, 1.0, a, lda, b, ldb);
    } else {
        zr_trsm(order, ZR_LEFT, ZR_LOWER, ZR_NOTRANS, ZR_NONUNIT, n, nrhs, 1.0, a, lda, b, ldb);
        zr_trsm(order, ZR_LEFT, ZR_LOWER, ZR_CONJTRANS, ZR_NONUNIT, n, nrhs, 1.0, a, lda, b, ldb);
    }
    return 0;
}

/* posv.zig:15-70 */
int zr_posv(int order, int uplo, int n, int nrhs, double* a, int lda, double* b, int ldb) {
    if (n < 0 || nrhs < 0 || lda < imax(1, n) || ldb < imax(1, order == ZR_COL ? n : nrhs)) return ZR_EINVAL;
    const int info = zr_potrf(order, uplo, n, a, lda);
    if (info == 0) zr_potrs(order, uplo, n, nrhs, a, lda, b, ldb);
    return info;
}

/* ------------------------------------------------------------------------ */
/* Closed form of zr_getrf: unblocked right-looking kij LU.
 * Every element receives a_ij <- fl(a_ij - fl(l_ik * u_kj)) for k ascending
 * (gemm.zig:169-189 with alpha = -1, trsm.zig:202-270 / :406-487), the pivot
 * is the first argmax (iamax.zig:32) and the column is scaled by fl(1/p)
 * when |p| >= sfmin, else divided (getrf2.zig:65-87). */
int zr_getrf_unblocked(int order, int m, int n, double* a, int lda, int32_t* ipiv) {
    if (m < 0 || n < 0 || lda < imax(1, order == ZR_COL ? m : n)) return ZR_EINVAL;
    if (m == 0 || n == 0) return 0;
    const double sfmin = zr_lamch('S');
    int info = 0;
    const int mn = imin(m, n);
    for (int k = 0; k < mn; ++k) {
        int p = k;
        double mx = fabs(a[idx(order, k, k, lda)]);
        for (int i = k + 1; i < m; ++i) {
            const double v = fabs(a[idx(order, i, k, lda)]);
            if (v > mx) { mx = v; p = i; }
        }
        ipiv[k] = p + 1;
        const double piv = a[idx(order, p, k, lda)];
        if (piv != 0.0) {
            if (p != k)
                for (int j = 0; j < n; ++j) {
                    double t = a[idx(order, k, j, lda)];
                    a[idx(order, k, j, lda)] = a[idx(order, p, j, lda)];
                    a[idx(order, p, j, lda)] = t;
                }
            const double d = a[idx(order, k, k, lda)];
            if (fabs(d) >= sfmin) {
                const double r = 1.0 / d;
                if (r != 1.0)
                    for (int i = k + 1; i < m; ++i) a[idx(order, i, k, lda)] = a[idx(order, i, k, lda)] * r;
            } else {
                for (int i = k + 1; i < m; ++i) a[idx(order, i, k, lda)] = a[idx(order, i, k, lda)] / d;
            }
        } else if (info == 0) {
            info = k + 1;
        }
        for (int j = k + 1; j < n; ++j) {
            const double u = a[idx(order, k, j, lda)];
            for (int i = k + 1; i < m; ++i)
                a[idx(order, i, j, lda)] = a[idx(order, i, j, lda)] - a[idx(order, i, k, lda)] * u;
        }
    }
    return info;
}

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

/* Column-major lower Cholesky, unblocked right-looking: l_jj = sqrt(a_jj -
 * sum_k l_jk^2), l_ij = (a_ij - sum_k l_ik l_jk) * fl(1/l_jj), k ascending. */
int zr_potrf_lower_unblocked(int n, double* a, int lda) {
    if (n < 0 || lda < imax(1, n)) return ZR_EINVAL;
    const i64 ld = lda;
    for (int j = 0; j < n; ++j) {
        const double ajj = a[j + j * ld];
        if (ajj <= 0.0 || isnan(ajj)) return j + 1;
        const double ljj = sqrt(ajj);
        a[j + j * ld] = ljj;
        const double r = 1.0 / ljj;
        for (int i = j + 1; i < n; ++i) a[i + j * ld] = a[i + j * ld] * r;
        for (int k = j + 1; k < n; ++k) {
            const double lkj = a[k + j * ld];
            for (int i = k; i < n; ++i) a[i + k * ld] = a[i + k * ld] - a[i + j * ld] * lkj;
        }
    }
    return 0;
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
