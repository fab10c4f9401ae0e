// extern "C" boundary of libzmlhip: the device-resident zmlhip_* API
// (include/zmlhip.h) and the synchronous host-pointer CBLAS/LAPACKE drop-ins
// (include/cblas.h, include/clapack.h). Argument checks restate zml's
// (each cited); errors never escape as C++ exceptions.
#include "zmlhip_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "../../include/cblas.h"
#include "../../include/clapack.h"
#include "../../include/zmlhip.h"
#include "mixed.hpp"

using namespace zmlhip;

namespace {

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const SyncError& e) {
    set_last_error(e.what());
    return ZMLHIP_ESYNC;
  } catch (const HipError& e) {
    set_last_error(e.what());
    return e.code == hipErrorOutOfMemory ? ZMLHIP_EALLOC : ZMLHIP_EHIP;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return ZMLHIP_EALLOC;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ZMLHIP_ENOTSUP;
  }
}

// guarded() for a call that queues work on `s`: it starts after every earlier
// call of this thread has finished with the shared workspaces (call_begin) and
// marks its own end on `s` (call_end).
template <class F>
int fenced(hipStream_t s, F&& f) {
  return guarded([&] {
    call_begin(s);
    const int rc = f();
    call_end(s);
    return rc;
  });
}

int trans_from_char(char c) {
  switch (c) {
    case 'N': case 'n': return kNoTrans;
    case 'T': case 't': return kTrans;
    case 'C': case 'c': return kConjTrans;
    default: return 0;
  }
}
int uplo_from_char(char c) {
  switch (c) {
    case 'U': case 'u': return kUpper;
    case 'L': case 'l': return kLower;
    default: return 0;
  }
}

// Elements spanned by a matrix of `rows` x `cols` in `layout` with leading dimension ld.
size_t span(int layout, int64_t rows, int64_t cols, int64_t ld) {
  if (rows <= 0 || cols <= 0) return 0;
  return layout == kColMajor ? (size_t)(ld * (cols - 1) + rows) : (size_t)(ld * (rows - 1) + cols);
}
size_t vspan(int64_t n, int64_t inc) {
  if (n <= 0) return 0;
  return (size_t)((n - 1) * (inc < 0 ? -inc : inc) + 1);
}

// Device copy of a host buffer, freed on scope exit.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) : bytes(b) {
    if (b) ZH_CHECK(hipMalloc(&p, b));
  }
  DevBuf(const void* host, size_t b, hipStream_t s) : DevBuf(b) {
    if (b && host) ZH_CHECK(hipMemcpyAsync(p, host, b, hipMemcpyHostToDevice, s));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void to_host(void* host, hipStream_t s) const {
    if (bytes && host) ZH_CHECK(hipMemcpyAsync(host, p, bytes, hipMemcpyDeviceToHost, s));
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// ---- validated device implementations (throw on HIP errors) ---------------------

// gemm.zig:14-65 mapping + :98-101 checks (T = double for dgemm, float for sgemm)
template <typename T>
int gemm_dev(int layout, int transa, int transb, int m, int n, int k, T alpha, const T* a, int lda, const T* b,
             int ldb, T beta, T* c, int ldc, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (!valid_trans(transa)) return -2;
  if (!valid_trans(transb)) return -3;
  if (layout == kRowMajor) {   // C^T = op(B)^T op(A)^T in the column-major view
    std::swap(transa, transb);
    std::swap(m, n);
    std::swap(a, b);
    std::swap(lda, ldb);
  }
  const int nrowa = is_notrans(transa) ? m : k, nrowb = is_notrans(transb) ? k : n;
  if (m < 0 || n < 0 || k < 0) return -4;
  if (lda < std::max(1, nrowa)) return layout == kRowMajor ? -11 : -9;
  if (ldb < std::max(1, nrowb)) return layout == kRowMajor ? -9 : -11;
  if (ldc < std::max(1, m)) return -14;
  if (m == 0 || n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return 0;
  gemm_cm<T>(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, 0);
  return 0;
}

// trsm.zig:18-63 mapping + :89-93 checks
template <typename T>
int trsm_dev(int layout, int side, int uplo, int transa, int diag, int m, int n, T alpha, const T* a, int lda,
             T* b, int ldb, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (!valid_side(side)) return -2;
  if (!valid_uplo(uplo)) return -3;
  if (!valid_trans(transa)) return -4;
  if (!valid_diag(diag)) return -5;
  if (layout == kRowMajor) {
    side = invert_side(side);
    uplo = invert_uplo(uplo);
    std::swap(m, n);
  }
  const int nrowa = side == kLeft ? m : n;
  if (m < 0 || n < 0) return -6;
  if (lda < std::max(1, nrowa)) return -10;
  if (ldb < std::max(1, m)) return -12;
  if (m == 0 || n == 0) return 0;
  trsm_cm<T>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, s);
  return 0;
}
int dtrsm_dev(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha,
              const double* a, int lda, double* b, int ldb, hipStream_t s) {
  return trsm_dev<double>(layout, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, s);
}

// syrk.zig:15-56 mapping + :80-84 checks (conjugate forms rejected)
template <typename T>
int syrk_dev(int layout, int uplo, int trans, int n, int k, T alpha, const T* a, int lda, T beta, T* c, int ldc,
             hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (!valid_uplo(uplo)) return -2;
  if (trans != kNoTrans && trans != kTrans) return -3;
  if (layout == kRowMajor) {
    uplo = invert_uplo(uplo);
    trans = invert_trans(trans);
  }
  const int nrowa = trans == kNoTrans ? n : k;
  if (n < 0 || k < 0) return -4;
  if (lda < std::max(1, nrowa)) return -8;
  if (ldc < std::max(1, n)) return -11;
  if (n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return 0;
  syrk_cm<T>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, s);
  return 0;
}

// trsv.zig:16-82 mapping and checks: row-major inverts uplo and trans
int dtrsv_dev(int layout, int uplo, int trans, int diag, int n, const double* a, int lda, double* x, int incx,
              hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (!valid_uplo(uplo)) return -2;
  if (!valid_trans(trans)) return -3;
  if (!valid_diag(diag)) return -4;
  if (n < 0) return -5;
  if (lda < std::max(1, n)) return -7;
  if (incx == 0) return -9;
  if (n == 0) return 0;
  if (layout == kRowMajor) {
    uplo = invert_uplo(uplo);
    trans = invert_trans(trans);
  }
  trsv_f64_cm(uplo, trans, diag, n, a, lda, x, incx, s);
  return 0;
}

// LAPACK dlange's norms (no zml counterpart). Row-major A is the column-major
// A^T: the one- and infinity-norms trade places.
bool valid_norm(char c) { return std::strchr("MmOo1IiFfEe", c) != nullptr && c != 0; }
int dlange_dev(int layout, char norm, int m, int n, const double* a, int lda, double* out, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (!valid_norm(norm)) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (lda < std::max(1, layout == kColMajor ? m : n)) return -6;
  if (m == 0 || n == 0) {
    ZH_CHECK(hipMemsetAsync(out, 0, sizeof(double), s));
    return 0;
  }
  if (layout == kRowMajor) {
    std::swap(m, n);
    if (norm == 'I' || norm == 'i') norm = 'O';
    else if (norm == 'O' || norm == 'o' || norm == '1') norm = 'I';
  }
  lange_f64_cm(norm, m, n, a, lda, out, s);
  return 0;
}

// gemv.zig:14-110
int dgemv_dev(int layout, int trans, int m, int n, double alpha, const double* a, int lda,
              const double* x, int incx, double beta, double* y, int incy, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (!valid_trans(trans)) return -2;
  if (layout == kRowMajor) {
    trans = invert_trans(trans);
    std::swap(m, n);
  }
  if (m < 0 || n < 0) return -3;
  if (lda < std::max(1, m)) return -7;
  if (incx == 0) return -9;
  if (incy == 0) return -12;
  gemv_f64_cm(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, s);
  return 0;
}

// laswp.zig:12-95 (1-based k1, k2 and ipiv)
int dlaswp_dev(int layout, int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx,
               hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (n < 0) return -2;
  if (layout == kRowMajor && lda < std::max(1, n)) return -4;
  if (incx == 0) return -8;
  if (n == 0) return 0;
  int i1, i2, inc, ix0;
  if (incx > 0) { ix0 = k1 - 1; i1 = k1 - 1; i2 = k2 - 1; inc = 1; }
  else { ix0 = k1 - 1 + (k1 - k2) * incx; i1 = k2 - 1; i2 = k1 - 1; inc = -1; }
  if (layout == kRowMajor) laswp_rm(n, a, lda, i1, i2, ipiv, inc, s, ix0, incx);
  else laswp_cm(n, a, lda, i1, i2, ipiv, inc, s, ix0, incx);
  return 0;
}

// getrf.zig:24-31; LAPACKE argument numbering (layout = 1)
template <typename T>
int getrf_dev(int layout, int m, int n, T* a, int lda, int* ipiv, int* info, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (lda < std::max(1, layout == kColMajor ? m : n)) return -5;
  ZH_CHECK(hipMemsetAsync(info, 0, sizeof(int), s));
  if (m == 0 || n == 0) return 0;
  if (layout == kRowMajor) {
    getrf_rm<T>(m, n, a, lda, ipiv, info, s);
  } else if (m == n) {
    transpose_square_inplace<T>(n, a, lda, s);
    getrf_rm<T>(m, n, a, lda, ipiv, info, s);
    transpose_square_inplace<T>(n, a, lda, s);
  } else {
    T* w = static_cast<T*>(workspace(sizeof(T) * (size_t)m * (size_t)n, 5));
    transpose_copy<T>(n, m, a, lda, w, n, s);
    getrf_rm<T>(m, n, w, n, ipiv, info, s);
    transpose_copy<T>(m, n, w, n, a, lda, s);
  }
  return 0;
}
int dgetrf_dev(int layout, int m, int n, double* a, int lda, int* ipiv, int* info, hipStream_t s) {
  return getrf_dev<double>(layout, m, n, a, lda, ipiv, info, s);
}

// getrs.zig:13-142 (trans 'N' / 'T' / 'C'; conjugation is a no-op for f64)
int dgetrs_dev(int layout, char trans_c, int n, int nrhs, const double* a, int lda, const int* ipiv,
               double* b, int ldb, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  const int trans = trans_from_char(trans_c);
  if (!trans) return -2;
  if (n < 0) return -3;
  if (nrhs < 0) return -4;
  if (lda < std::max(1, n)) return -6;
  if (ldb < std::max(1, layout == kColMajor ? n : nrhs)) return -9;
  if (n == 0 || nrhs == 0) return 0;
  if (nrhs == 1) {
    // one right-hand side: the vector interchanges and one persistent launch
    // per triangle (element (i, j) of the factors at a[i si + j sj])
    const int64_t si = layout == kRowMajor ? lda : 1, sj = layout == kRowMajor ? 1 : lda;
    const int64_t inc = layout == kRowMajor ? ldb : 1;
    const int32_t* ip = reinterpret_cast<const int32_t*>(ipiv);
    if (is_notrans(trans)) {
      laswp_vec(b, inc, 0, n - 1, ip, 1, s);
      trsv1_f64(n, a, si, sj, true, true, b, inc, s);     // L, unit
      trsv1_f64(n, a, si, sj, false, false, b, inc, s);   // U
    } else {
      trsv1_f64(n, a, sj, si, true, false, b, inc, s);    // U^T
      trsv1_f64(n, a, sj, si, false, true, b, inc, s);    // L^T, unit
      laswp_vec(b, inc, n - 1, 0, ip, -1, s);
    }
    return 0;
  }
  auto trsm = [&](int side, int uplo, int t, int diag) {
    return dtrsm_dev(layout, side, uplo, t, diag, n, nrhs, 1.0, a, lda, b, ldb, s);
  };
  auto swaps = [&](int inc) {
    const int i1 = inc > 0 ? 0 : n - 1, i2 = inc > 0 ? n - 1 : 0;
    if (layout == kRowMajor) laswp_rm(nrhs, b, ldb, i1, i2, ipiv, inc, s);
    else laswp_cm(nrhs, b, ldb, i1, i2, ipiv, inc, s);
  };
  if (is_notrans(trans)) {
    swaps(1);
    trsm(kLeft, kLower, kNoTrans, kUnit);
    trsm(kLeft, kUpper, kNoTrans, kNonUnit);
  } else {
    trsm(kLeft, kUpper, kTrans, kNonUnit);
    trsm(kLeft, kLower, kTrans, kUnit);
    swaps(-1);
  }
  return 0;
}

// potrf.zig:17-31
template <typename T>
int potrf_dev(int layout, char uplo_c, int n, T* a, int lda, int* info, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  int uplo = uplo_from_char(uplo_c);
  if (!uplo) return -2;
  if (n < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  ZH_CHECK(hipMemsetAsync(info, 0, sizeof(int), s));
  if (n == 0) return 0;
  if (layout == kRowMajor) uplo = invert_uplo(uplo);
  potrf_cm<T>(uplo, n, a, lda, info, s);
  return 0;
}
int dpotrf_dev(int layout, char uplo_c, int n, double* a, int lda, int* info, hipStream_t s) {
  return potrf_dev<double>(layout, uplo_c, n, a, lda, info, s);
}

// potrs.zig:15-117. zml checks ldb >= n in both layouts (potrs.zig:30); this
// follows LAPACKE (ldb >= nrhs for row-major) so that valid row-major calls pass.
int dpotrs_dev(int layout, char uplo_c, int n, int nrhs, const double* a, int lda, double* b,
               int ldb, hipStream_t s) {
  if (!valid_layout(layout)) return -1;
  const int uplo = uplo_from_char(uplo_c);
  if (!uplo) return -2;
  if (n < 0) return -3;
  if (nrhs < 0) return -4;
  if (lda < std::max(1, n)) return -6;
  if (ldb < std::max(1, layout == kColMajor ? n : nrhs)) return -8;
  if (n == 0 || nrhs == 0) return 0;
  if (nrhs == 1) {
    // one right-hand side: one persistent launch per triangle; the stored
    // triangle T has element (i, j) at a[i si + j sj]
    const int64_t si = layout == kRowMajor ? lda : 1, sj = layout == kRowMajor ? 1 : lda;
    const int64_t inc = layout == kRowMajor ? ldb : 1;
    if (uplo == kLower) {
      trsv1_f64(n, a, si, sj, true, false, b, inc, s);    // L
      trsv1_f64(n, a, sj, si, false, false, b, inc, s);   // L^T
    } else {
      trsv1_f64(n, a, sj, si, true, false, b, inc, s);    // U^T
      trsv1_f64(n, a, si, sj, false, false, b, inc, s);   // U
    }
    return 0;
  }
  auto trsm = [&](int u, int t) {
    return dtrsm_dev(layout, kLeft, u, t, kNonUnit, n, nrhs, 1.0, a, lda, b, ldb, s);
  };
  if (uplo == kUpper) {
    trsm(kUpper, kTrans);
    trsm(kUpper, kNoTrans);
  } else {
    trsm(kLower, kNoTrans);
    trsm(kLower, kTrans);
  }
  return 0;
}

}  // namespace

// =============================================================================
// Device-resident API (include/zmlhip.h)
// =============================================================================
extern "C" {

int zmlhip_version(void) { return 100; }

const char* zmlhip_last_error(void) { return last_error(); }

int zmlhip_profile_enable(int on) {
  profile_enable(on != 0);
  return 0;
}

int zmlhip_profile_collect(int kind, double* ms, long long* launches, double* flops, double* bytes) {
  return guarded([&] {
    profile_collect(kind, ms, launches, flops, bytes);
    return 0;
  });
}

int zmlhip_stream_synchronize(void* stream) {
  return guarded([&] {
    ZH_CHECK(hipStreamSynchronize(as_stream(stream)));
    return 0;
  });
}

int zmlhip_dgemm(int layout, int transa, int transb, int m, int n, int k, double alpha,
                 const double* a, int lda, const double* b, int ldb, double beta, double* c,
                 int ldc, void* stream) {
  return fenced(as_stream(stream), [&] {
    return gemm_dev<double>(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                            as_stream(stream));
  });
}

int zmlhip_dtrsm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha,
                 const double* a, int lda, double* b, int ldb, void* stream) {
  return fenced(as_stream(stream), [&] {
    return dtrsm_dev(layout, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb,
                     as_stream(stream));
  });
}

int zmlhip_dsyrk(int layout, int uplo, int trans, int n, int k, double alpha, const double* a,
                 int lda, double beta, double* c, int ldc, void* stream) {
  return fenced(as_stream(stream), [&] {
    return syrk_dev<double>(layout, uplo, trans, n, k, alpha, a, lda, beta, c, ldc, as_stream(stream));
  });
}

int zmlhip_idamax(int n, const double* x, int incx, int* result, void* stream) {
  return fenced(as_stream(stream), [&] {
    if (n <= 0) return -1;      // iamax.zig:19
    if (incx <= 0) return -3;
    iamax_f64(n, x, incx, result, as_stream(stream));
    return 0;
  });
}

int zmlhip_dscal(int n, double alpha, double* x, int incx, void* stream) {
  return fenced(as_stream(stream), [&] {
    if (n < 0) return -1;       // scal.zig:20
    if (incx <= 0) return -4;
    scal_f64(n, alpha, x, incx, as_stream(stream));
    return 0;
  });
}

int zmlhip_dswap(int n, double* x, int incx, double* y, int incy, void* stream) {
  return fenced(as_stream(stream), [&] {
    if (n < 0) return -1;       // swap.zig:21
    swap_f64(n, x, incx, y, incy, as_stream(stream));
    return 0;
  });
}

int zmlhip_dgemv(int layout, int transa, int m, int n, double alpha, const double* a, int lda,
                 const double* x, int incx, double beta, double* y, int incy, void* stream) {
  return fenced(as_stream(stream), [&] {
    return dgemv_dev(layout, transa, m, n, alpha, a, lda, x, incx, beta, y, incy, as_stream(stream));
  });
}

int zmlhip_dgemv_noalloc(int layout, int transa, int m, int n, double alpha, const double* a, int lda,
                         const double* x, int incx, double beta, double* y, int incy, void* stream) {
  // no call fence: this entry point touches none of the thread's workspaces,
  // streams or events (and creates none), only the device's static array
  return guarded([&] {
    if (!valid_layout(layout)) return -1;
    if (!valid_trans(transa)) return -2;
    int t = transa, mm = m, nn = n;
    if (layout == kRowMajor) {
      t = invert_trans(t);
      std::swap(mm, nn);
    }
    if (mm < 0 || nn < 0) return -3;
    if (lda < std::max(1, mm)) return -7;
    if (incx == 0) return -9;
    if (incy == 0) return -12;
    if (gemv_f64_cm_noalloc(t, mm, nn, alpha, a, lda, x, incx, beta, y, incy, as_stream(stream)) != 0) {
      set_last_error("dgemv_noalloc: more rows in the no-transpose view than the static scratch holds "
                     "(1048576)");
      return ZMLHIP_ENOTSUP;
    }
    return 0;
  });
}

int zmlhip_dtrsv(int layout, int uplo, int transa, int diag, int n, const double* a, int lda, double* x,
                 int incx, void* stream) {
  return fenced(as_stream(stream), [&] {
    return dtrsv_dev(layout, uplo, transa, diag, n, a, lda, x, incx, as_stream(stream));
  });
}

int zmlhip_dlange(int layout, char norm, int m, int n, const double* a, int lda, double* result, void* stream) {
  return fenced(as_stream(stream), [&] { return dlange_dev(layout, norm, m, n, a, lda, result, as_stream(stream)); });
}

int zmlhip_sgemm(int layout, int transa, int transb, int m, int n, int k, float alpha, const float* a, int lda,
                 const float* b, int ldb, float beta, float* c, int ldc, void* stream) {
  return fenced(as_stream(stream), [&] {
    return gemm_dev<float>(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, as_stream(stream));
  });
}

int zmlhip_strsm(int layout, int side, int uplo, int transa, int diag, int m, int n, float alpha, const float* a,
                 int lda, float* b, int ldb, void* stream) {
  return fenced(as_stream(stream), [&] {
    return trsm_dev<float>(layout, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, as_stream(stream));
  });
}

int zmlhip_ssyrk(int layout, int uplo, int trans, int n, int k, float alpha, const float* a, int lda, float beta,
                 float* c, int ldc, void* stream) {
  return fenced(as_stream(stream), [&] {
    return syrk_dev<float>(layout, uplo, trans, n, k, alpha, a, lda, beta, c, ldc, as_stream(stream));
  });
}

int zmlhip_sgetrf(int layout, int m, int n, float* a, int lda, int* ipiv, int* info, void* stream) {
  return fenced(as_stream(stream), [&] { return getrf_dev<float>(layout, m, n, a, lda, ipiv, info, as_stream(stream)); });
}

int zmlhip_spotrf(int layout, char uplo, int n, float* a, int lda, int* info, void* stream) {
  return fenced(as_stream(stream), [&] { return potrf_dev<float>(layout, uplo, n, a, lda, info, as_stream(stream)); });
}

int zmlhip_set_panel_spin_limit(long polls) {
  set_panel_spin_limit(polls == 0 ? 2000000L : polls);
  return 0;
}

int zmlhip_set_panel_one_launch_rows(long rows) {
  set_panel_one_launch_rows(rows);
  return 0;
}

int zmlhip_set_leaf_lookahead(int on) {
  const int was = leaf_lookahead() ? 1 : 0;
  set_leaf_lookahead(on);
  return was;
}

int zmlhip_set_gemm_lds_dma(int on) {
  const int was = gemm_dma_enabled() ? 1 : 0;
  set_gemm_dma(on);
  return was;
}

int zmlhip_dlaswp(int layout, int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx,
                  void* stream) {
  return fenced(as_stream(stream), [&] { return dlaswp_dev(layout, n, a, lda, k1, k2, ipiv, incx, as_stream(stream)); });
}

int zmlhip_dgetrf(int layout, int m, int n, double* a, int lda, int* ipiv, int* info, void* stream) {
  return fenced(as_stream(stream), [&] { return dgetrf_dev(layout, m, n, a, lda, ipiv, info, as_stream(stream)); });
}

int zmlhip_dgetrs(int layout, char trans, int n, int nrhs, const double* a, int lda,
                  const int* ipiv, double* b, int ldb, void* stream) {
  return fenced(as_stream(stream), [&] {
    return dgetrs_dev(layout, trans, n, nrhs, a, lda, ipiv, b, ldb, as_stream(stream));
  });
}

int zmlhip_dpotrf(int layout, char uplo, int n, double* a, int lda, int* info, void* stream) {
  return fenced(as_stream(stream), [&] { return dpotrf_dev(layout, uplo, n, a, lda, info, as_stream(stream)); });
}

int zmlhip_dpotrs(int layout, char uplo, int n, int nrhs, const double* a, int lda, double* b,
                  int ldb, void* stream) {
  return fenced(as_stream(stream), [&] {
    return dpotrs_dev(layout, uplo, n, nrhs, a, lda, b, ldb, as_stream(stream));
  });
}

int zmlhip_dgesv_mixed(int layout, int n, int nrhs, const double* a, int lda, const double* b, int ldb,
                       double* x, int ldx, const zmlhip_ir_opts* opts, zmlhip_ir_result* result,
                       void* stream) {
  return fenced(as_stream(stream), [&] {
    if (!valid_layout(layout)) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (lda < std::max(1, n)) return -5;
    if (ldb < std::max(1, layout == kColMajor ? n : nrhs)) return -7;
    if (ldx < std::max(1, layout == kColMajor ? n : nrhs)) return -9;
    if (opts && (opts->precision < 0 || opts->precision > 2)) return -10;
    if (!result) return -11;
    return dgesv_mixed(layout, n, nrhs, a, lda, b, ldb, x, ldx, opts, result, as_stream(stream));
  });
}

int zmlhip_dposv_mixed(int layout, char uplo, int n, int nrhs, const double* a, int lda, const double* b,
                       int ldb, double* x, int ldx, const zmlhip_ir_opts* opts, zmlhip_ir_result* result,
                       void* stream) {
  return fenced(as_stream(stream), [&] {
    if (!valid_layout(layout)) return -1;
    const int ul = uplo_from_char(uplo);
    if (!ul) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (lda < std::max(1, n)) return -6;
    if (ldb < std::max(1, layout == kColMajor ? n : nrhs)) return -8;
    if (ldx < std::max(1, layout == kColMajor ? n : nrhs)) return -10;
    if (opts && (opts->precision < 0 || opts->precision > 2)) return -11;
    if (!result) return -12;
    return dposv_mixed(layout, ul, n, nrhs, a, lda, b, ldb, x, ldx, opts, result, as_stream(stream));
  });
}

int zmlhip_batch_plan_create(zmlhip_batch_plan** plan, int count, const int* n, const int* nrhs,
                             const int* spd, const double* const* a, const double* const* b, double* const* x,
                             const zmlhip_ir_opts* opts, int autotune, void* stream) {
  return fenced(as_stream(stream), [&] {
    if (!plan) return -1;
    *plan = nullptr;
    if (count < 0) return -2;
    if (count > 0 && (!n || !nrhs || !a || !b || !x)) return -3;
    for (int i = 0; i < count; ++i) {
      if (n[i] < 1 || n[i] > 256) return -3;
      if (nrhs[i] < 1) return -4;
      if (!a[i] || !b[i] || !x[i]) return -6;
    }
    *plan = reinterpret_cast<zmlhip_batch_plan*>(
        batch_plan_create(count, n, nrhs, spd, a, b, x, opts, autotune, as_stream(stream)));
    return 0;
  });
}

int zmlhip_batch_plan_execute(zmlhip_batch_plan* plan, void* stream) {
  return fenced(as_stream(stream), [&] {
    if (!plan) return -1;
    batch_plan_execute(reinterpret_cast<BatchPlan*>(plan), as_stream(stream));
    return 0;
  });
}

int zmlhip_batch_plan_status(zmlhip_batch_plan* plan, int* status, void* stream) {
  return fenced(as_stream(stream), [&] {
    if (!plan) return -1;
    if (!status) return -2;
    batch_plan_status(reinterpret_cast<BatchPlan*>(plan), status, as_stream(stream));
    return 0;
  });
}

int zmlhip_batch_plan_shapes(const zmlhip_batch_plan* plan, int* tg, int* r, int* count) {
  return guarded([&] {
    if (!plan || !tg || !r || !count) return -1;
    batch_plan_variants(reinterpret_cast<const BatchPlan*>(plan), tg, r, count);
    return 0;
  });
}

int zmlhip_batch_plan_destroy(zmlhip_batch_plan* plan) {
  return guarded([&] {
    if (plan) batch_plan_destroy(reinterpret_cast<BatchPlan*>(plan));
    return 0;
  });
}

// =============================================================================
// Host-pointer CBLAS (include/cblas.h): synchronous, errors ignored as zml does
// =============================================================================

}  // extern "C"

namespace {
template <typename T>
void host_gemm(int order, int transa, int transb, int m, int n, int k, T alpha, const T* a, int lda, const T* b,
               int ldb, T beta, T* c, int ldc) {
  fenced(host_api_stream(), [&] {
    if (m <= 0 || n <= 0 || k < 0) return 0;
    const hipStream_t s = host_api_stream();
    const int ar = is_notrans(transa) ? m : k, ac = is_notrans(transa) ? k : m;
    const int br = is_notrans(transb) ? k : n, bc = is_notrans(transb) ? n : k;
    DevBuf da(a, sizeof(T) * span(order, ar, ac, lda), s);
    DevBuf db(b, sizeof(T) * span(order, br, bc, ldb), s);
    DevBuf dc(c, sizeof(T) * span(order, m, n, ldc), s);
    const int rc = gemm_dev<T>(order, transa, transb, m, n, k, alpha, da.as<T>(), lda, db.as<T>(), ldb, beta,
                               dc.as<T>(), ldc, s);
    if (rc == 0) dc.to_host(c, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}

template <typename T>
void host_trsm(int order, int side, int uplo, int transa, int diag, int m, int n, T alpha, const T* a, int lda,
               T* b, int ldb) {
  fenced(host_api_stream(), [&] {
    if (m <= 0 || n <= 0) return 0;
    const hipStream_t s = host_api_stream();
    const int na = side == CblasLeft ? m : n;
    DevBuf da(a, sizeof(T) * span(order, na, na, lda), s);
    DevBuf db(b, sizeof(T) * span(order, m, n, ldb), s);
    const int rc = trsm_dev<T>(order, side, uplo, transa, diag, m, n, alpha, da.as<T>(), lda, db.as<T>(), ldb, s);
    if (rc == 0) db.to_host(b, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}

template <typename T>
void host_syrk(int order, int uplo, int trans, int n, int k, T alpha, const T* a, int lda, T beta, T* c,
               int ldc) {
  fenced(host_api_stream(), [&] {
    if (n <= 0 || k < 0) return 0;
    const hipStream_t s = host_api_stream();
    const int ar = trans == CblasNoTrans ? n : k, ac = trans == CblasNoTrans ? k : n;
    DevBuf da(a, sizeof(T) * span(order, ar, ac, lda), s);
    DevBuf dc(c, sizeof(T) * span(order, n, n, ldc), s);
    const int rc = syrk_dev<T>(order, uplo, trans, n, k, alpha, da.as<T>(), lda, beta, dc.as<T>(), ldc, s);
    if (rc == 0) dc.to_host(c, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}
}  // namespace

extern "C" {

void cblas_dgemm(const CBLAS_ORDER order, const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                 const int m, const int n, const int k, const double alpha, const double* a,
                 const int lda, const double* b, const int ldb, const double beta, double* c,
                 const int ldc) {
  host_gemm<double>(order, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void cblas_sgemm(const CBLAS_ORDER order, const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                 const int m, const int n, const int k, const float alpha, const float* a, const int lda,
                 const float* b, const int ldb, const float beta, float* c, const int ldc) {
  host_gemm<float>(order, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void cblas_dtrsm(const CBLAS_ORDER order, const CBLAS_SIDE side, const CBLAS_UPLO uplo,
                 const CBLAS_TRANSPOSE transa, const CBLAS_DIAG diag, const int m, const int n,
                 const double alpha, const double* a, const int lda, double* b, const int ldb) {
  host_trsm<double>(order, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

void cblas_strsm(const CBLAS_ORDER order, const CBLAS_SIDE side, const CBLAS_UPLO uplo,
                 const CBLAS_TRANSPOSE transa, const CBLAS_DIAG diag, const int m, const int n,
                 const float alpha, const float* a, const int lda, float* b, const int ldb) {
  host_trsm<float>(order, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

void cblas_dsyrk(const CBLAS_ORDER order, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE trans,
                 const int n, const int k, const double alpha, const double* a, const int lda,
                 const double beta, double* c, const int ldc) {
  host_syrk<double>(order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}

void cblas_ssyrk(const CBLAS_ORDER order, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE trans, const int n,
                 const int k, const float alpha, const float* a, const int lda, const float beta, float* c,
                 const int ldc) {
  host_syrk<float>(order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}

void cblas_dtrsv(const CBLAS_ORDER order, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                 const CBLAS_DIAG diag, const int n, const double* a, const int lda, double* x, const int incx) {
  fenced(host_api_stream(), [&] {
    if (n <= 0 || incx == 0) return 0;
    const hipStream_t s = host_api_stream();
    DevBuf da(a, sizeof(double) * span(order, n, n, lda), s);
    DevBuf dx(x, sizeof(double) * vspan(n, incx), s);
    const int rc = dtrsv_dev(order, uplo, transa, diag, n, da.as<double>(), lda, dx.as<double>(), incx, s);
    if (rc == 0) dx.to_host(x, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}

unsigned int cblas_idamax(const int n, const double* x, const int incx) {
  unsigned int out = 0;   // blas.zig:4573-4578 returns 0 on invalid arguments
  fenced(host_api_stream(), [&] {
    if (n <= 0 || incx <= 0) return 0;
    const hipStream_t s = host_api_stream();
    DevBuf dx(x, sizeof(double) * vspan(n, incx), s);
    DevBuf dr(sizeof(int));
    iamax_f64(n, dx.as<double>(), incx, dr.as<int32_t>(), s);
    int r = 0;
    ZH_CHECK(hipMemcpyAsync(&r, dr.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    out = (unsigned int)r;
    return 0;
  });
  return out;
}

void cblas_dscal(const int n, const double alpha, double* x, const int incx) {
  fenced(host_api_stream(), [&] {
    if (n <= 0 || incx <= 0 || alpha == 1.0) return 0;
    const hipStream_t s = host_api_stream();
    DevBuf dx(x, sizeof(double) * vspan(n, incx), s);
    scal_f64(n, alpha, dx.as<double>(), incx, s);
    dx.to_host(x, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return 0;
  });
}

void cblas_dswap(const int n, double* x, const int incx, double* y, const int incy) {
  fenced(host_api_stream(), [&] {
    if (n <= 0) return 0;
    const hipStream_t s = host_api_stream();
    DevBuf dx(x, sizeof(double) * vspan(n, incx), s);
    DevBuf dy(y, sizeof(double) * vspan(n, incy), s);
    swap_f64(n, dx.as<double>(), incx, dy.as<double>(), incy, s);
    dx.to_host(x, s);
    dy.to_host(y, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return 0;
  });
}

void cblas_dgemv(const CBLAS_ORDER order, const CBLAS_TRANSPOSE transa, const int m, const int n,
                 const double alpha, const double* a, const int lda, const double* x, const int incx,
                 const double beta, double* y, const int incy) {
  fenced(host_api_stream(), [&] {
    if (m <= 0 || n <= 0 || incx == 0 || incy == 0) return 0;
    const hipStream_t s = host_api_stream();
    const bool nt = is_notrans(transa);
    const int lenx = nt ? n : m, leny = nt ? m : n;
    DevBuf da(a, sizeof(double) * span(order, m, n, lda), s);
    DevBuf dx(x, sizeof(double) * vspan(lenx, incx), s);
    DevBuf dy(y, sizeof(double) * vspan(leny, incy), s);
    const int rc = dgemv_dev(order, transa, m, n, alpha, da.as<double>(), lda, dx.as<double>(), incx,
                             beta, dy.as<double>(), incy, s);
    if (rc == 0) dy.to_host(y, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}

// =============================================================================
// Host-pointer LAPACKE (include/clapack.h): synchronous, LAPACKE return codes
// =============================================================================

double LAPACKE_dlamch(char cmach) {
  // lamch.zig:15-78 at T = f64 with rnd = 1, over Zig's std.math facts:
  // eps = floatEps / 2, sfmin = floatMin (1 / floatMax is smaller),
  // t = floatMantissaBits (52, the stored bits), emin = floatExponentMin
  // (-1022), emax = floatExponentMax (1023)
  const double eps = 0x1p-53;
  switch (cmach) {
    case 'E': case 'e': return eps;
    case 'S': case 's': return 0x1p-1022;
    case 'B': case 'b': return 2.0;
    case 'P': case 'p': return eps * 2.0;
    case 'N': case 'n': return 52.0;
    case 'R': case 'r': return 1.0;
    case 'M': case 'm': return -1022.0;
    case 'U': case 'u': return 0x1p-1022;
    case 'L': case 'l': return 1023.0;
    case 'O': case 'o': return 0x1.fffffffffffffp+1023;
    default: return 0.0;
  }
}

double LAPACKE_dlange(int layout, char norm, int m, int n, const double* a, int lda) {
  double out = -1.0;
  const int rc = fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (!valid_norm(norm)) return -2;
    if (m < 0 || n < 0) return -3;
    if (lda < std::max(1, layout == kColMajor ? m : n)) return -6;
    if (m == 0 || n == 0) { out = 0.0; return 0; }
    const hipStream_t s = host_api_stream();
    DevBuf da(a, sizeof(double) * span(layout, m, n, lda), s);
    DevBuf dr(sizeof(double));
    const int r = dlange_dev(layout, norm, m, n, da.as<double>(), lda, dr.as<double>(), s);
    if (r != 0) return r;
    ZH_CHECK(hipMemcpyAsync(&out, dr.p, sizeof(double), hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    return 0;
  });
  return rc == 0 ? out : (double)rc;
}

int LAPACKE_dlaswp(int layout, int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  return fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (n <= 0 || k2 < k1) return n < 0 ? -2 : 0;
    const hipStream_t s = host_api_stream();
    // rows touched: up to max over the used ipiv entries
    const int used = (k2 - k1) * std::abs(incx) + k1;
    int rows = k2;
    for (int e = 0; e < used; ++e) rows = std::max(rows, ipiv[e]);
    DevBuf da(a, sizeof(double) * span(layout, rows, n, lda), s);
    DevBuf dp(ipiv, sizeof(int) * (size_t)used, s);
    const int rc = dlaswp_dev(layout, n, da.as<double>(), lda, k1, k2, dp.as<int>(), incx, s);
    if (rc == 0) da.to_host(a, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}

}  // extern "C"

template <typename T>
static int host_getrf(int layout, int m, int n, T* a, int lda, int* ipiv) {
  return fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (lda < std::max(1, layout == kColMajor ? m : n)) return -5;
    if (m == 0 || n == 0) return 0;
    const hipStream_t s = host_api_stream();
    const int mn = std::min(m, n);
    DevBuf da(a, sizeof(T) * span(layout, m, n, lda), s);
    DevBuf dp(sizeof(int) * (size_t)mn);
    DevBuf di(sizeof(int));
    int rc = getrf_dev<T>(layout, m, n, da.as<T>(), lda, dp.as<int>(), di.as<int>(), s);
    if (rc != 0) return rc;
    int info = 0;
    da.to_host(a, s);
    dp.to_host(ipiv, s);
    ZH_CHECK(hipMemcpyAsync(&info, di.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    if (info == kErrSync) throw SyncError("getrf: a panel's cross-workgroup wait gave up");
    return info;
  });
}

extern "C" {

int LAPACKE_dgetrf(int layout, int m, int n, double* a, int lda, int* ipiv) {
  return host_getrf<double>(layout, m, n, a, lda, ipiv);
}

int LAPACKE_dgetrf2(int layout, int m, int n, double* a, int lda, int* ipiv) {
  // The recursive and the blocked LU produce the same factors and pivots
  // (zml's getrf calls getrf2 per panel); both run the blocked GPU driver.
  return host_getrf<double>(layout, m, n, a, lda, ipiv);
}

int LAPACKE_sgetrf(int layout, int m, int n, float* a, int lda, int* ipiv) {
  return host_getrf<float>(layout, m, n, a, lda, ipiv);
}

int LAPACKE_dgetrs(int layout, char trans, int n, int nrhs, const double* a, int lda, const int* ipiv,
                   double* b, int ldb) {
  return fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (n <= 0 || nrhs <= 0) return (n < 0) ? -3 : (nrhs < 0 ? -4 : 0);
    const hipStream_t s = host_api_stream();
    DevBuf da(a, sizeof(double) * span(layout, n, n, lda), s);
    DevBuf dp(ipiv, sizeof(int) * (size_t)n, s);
    DevBuf db(b, sizeof(double) * span(layout, n, nrhs, ldb), s);
    const int rc = dgetrs_dev(layout, trans, n, nrhs, da.as<double>(), lda, dp.as<int>(),
                              db.as<double>(), ldb, s);
    if (rc == 0) db.to_host(b, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}

int LAPACKE_dgesv(int layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb) {
  // gesv.zig:39-62: factor, then solve when info == 0
  return fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (lda < std::max(1, n)) return -5;
    if (ldb < std::max(1, layout == kColMajor ? n : nrhs)) return -8;
    if (n == 0) return 0;
    const hipStream_t s = host_api_stream();
    DevBuf da(a, sizeof(double) * span(layout, n, n, lda), s);
    DevBuf dp(sizeof(int) * (size_t)n);
    DevBuf di(sizeof(int));
    DevBuf db(b, sizeof(double) * span(layout, n, nrhs, ldb), s);
    int rc = dgetrf_dev(layout, n, n, da.as<double>(), lda, dp.as<int>(), di.as<int>(), s);
    if (rc != 0) return rc;
    int info = 0;
    ZH_CHECK(hipMemcpyAsync(&info, di.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    if (info == kErrSync) throw SyncError("gesv: a panel's cross-workgroup wait gave up");
    if (info == 0 && nrhs > 0) {
      rc = dgetrs_dev(layout, 'N', n, nrhs, da.as<double>(), lda, dp.as<int>(), db.as<double>(), ldb, s);
      if (rc != 0) return rc;
      db.to_host(b, s);
    }
    da.to_host(a, s);
    dp.to_host(ipiv, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return info;
  });
}

}  // extern "C"

template <typename T>
static int host_potrf(int layout, char uplo, int n, T* a, int lda) {
  return fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (!uplo_from_char(uplo)) return -2;
    if (n < 0) return -3;
    if (lda < std::max(1, n)) return -5;
    if (n == 0) return 0;
    const hipStream_t s = host_api_stream();
    DevBuf da(a, sizeof(T) * span(layout, n, n, lda), s);
    DevBuf di(sizeof(int));
    int rc = potrf_dev<T>(layout, uplo, n, da.as<T>(), lda, di.as<int>(), s);
    if (rc != 0) return rc;
    int info = 0;
    da.to_host(a, s);
    ZH_CHECK(hipMemcpyAsync(&info, di.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    return info;
  });
}

extern "C" {

int LAPACKE_dpotrf(int layout, char uplo, int n, double* a, int lda) {
  return host_potrf<double>(layout, uplo, n, a, lda);
}

int LAPACKE_dpotrf2(int layout, char uplo, int n, double* a, int lda) {
  // the recursive and the blocked Cholesky compute the same factor (zml's
  // potrf calls potrf2 per diagonal block); both run the blocked GPU driver
  return host_potrf<double>(layout, uplo, n, a, lda);
}

int LAPACKE_spotrf(int layout, char uplo, int n, float* a, int lda) {
  return host_potrf<float>(layout, uplo, n, a, lda);
}

int LAPACKE_dpotrs(int layout, char uplo, int n, int nrhs, const double* a, int lda, double* b, int ldb) {
  return fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (n <= 0 || nrhs <= 0) return (n < 0) ? -3 : (nrhs < 0 ? -4 : 0);
    const hipStream_t s = host_api_stream();
    DevBuf da(a, sizeof(double) * span(layout, n, n, lda), s);
    DevBuf db(b, sizeof(double) * span(layout, n, nrhs, ldb), s);
    const int rc = dpotrs_dev(layout, uplo, n, nrhs, da.as<double>(), lda, db.as<double>(), ldb, s);
    if (rc == 0) db.to_host(b, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return rc;
  });
}

int LAPACKE_dposv(int layout, char uplo, int n, int nrhs, double* a, int lda, double* b, int ldb) {
  // posv.zig:39-62: factor, then solve when info == 0
  return fenced(host_api_stream(), [&] {
    if (!valid_layout(layout)) return -1;
    if (!uplo_from_char(uplo)) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (lda < std::max(1, n)) return -6;
    if (ldb < std::max(1, layout == kColMajor ? n : nrhs)) return -8;
    if (n == 0) return 0;
    const hipStream_t s = host_api_stream();
    DevBuf da(a, sizeof(double) * span(layout, n, n, lda), s);
    DevBuf di(sizeof(int));
    DevBuf db(b, sizeof(double) * span(layout, n, nrhs, ldb), s);
    int rc = dpotrf_dev(layout, uplo, n, da.as<double>(), lda, di.as<int>(), s);
    if (rc != 0) return rc;
    int info = 0;
    ZH_CHECK(hipMemcpyAsync(&info, di.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    if (info == 0 && nrhs > 0) {
      rc = dpotrs_dev(layout, uplo, n, nrhs, da.as<double>(), lda, db.as<double>(), ldb, s);
      if (rc != 0) return rc;
      db.to_host(b, s);
    }
    da.to_host(a, s);
    ZH_CHECK(hipStreamSynchronize(s));
    return info;
  });
}

}  // extern "C"
