// Internal definitions shared by the HIP sources of libzmlhip.
//
// Conventions
//  * All device entry points take a hipStream_t and are asynchronous; the
//    host-pointer CBLAS/LAPACKE entry points (capi.cpp) are synchronous.
//  * Enumerations carry the CBLAS/LAPACKE numeric values zml maps to
//    (reference include/cblas.h:11-37).
//  * "cm" in a name means the routine works on the column-major view; a
//    row-major call is mapped onto it the way zml maps it (swap operands /
//    invert side and uplo), see each routine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

struct zmlhip_ir_opts;

namespace zmlhip {

enum Layout : int { kRowMajor = 101, kColMajor = 102 };
enum Trans : int { kNoTrans = 111, kTrans = 112, kConjTrans = 113, kConjNoTrans = 114 };
enum Uplo : int { kUpper = 121, kLower = 122 };
enum Diag : int { kNonUnit = 131, kUnit = 132 };
enum Side : int { kLeft = 141, kRight = 142 };

// Status codes returned by the device entry points (0 = success). LAPACK
// info values (> 0) and -i argument errors are returned as LAPACKE does.
enum Status : int {
  kOk = 0,
  kErrHip = -1000,       // a HIP runtime call failed
  kErrAlloc = -1001,     // device allocation failed
  kErrSync = -1002,      // a bounded device-side wait timed out
  kErrNotSupported = -1003,
};

inline bool is_notrans(int t) { return t == kNoTrans || t == kConjNoTrans; }
inline bool valid_trans(int t) { return t >= kNoTrans && t <= kConjNoTrans; }
inline bool valid_layout(int l) { return l == kRowMajor || l == kColMajor; }
inline bool valid_uplo(int u) { return u == kUpper || u == kLower; }
inline bool valid_diag(int d) { return d == kNonUnit || d == kUnit; }
inline bool valid_side(int s) { return s == kLeft || s == kRight; }
inline int invert_uplo(int u) { return u == kUpper ? kLower : kUpper; }
inline int invert_side(int s) { return s == kLeft ? kRight : kLeft; }
inline int invert_trans(int t) {
  switch (t) {
    case kNoTrans: return kTrans;
    case kTrans: return kNoTrans;
    case kConjTrans: return kConjNoTrans;
    default: return kConjTrans;
  }
}

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
};
// A bounded device-side wait gave up (kErrSync at the C boundary).
struct SyncError : std::runtime_error {
  explicit SyncError(const char* what) : std::runtime_error(what) {}
};

#define ZH_CHECK(expr)                                                                      \
  do {                                                                                      \
    hipError_t zh_err_ = (expr);                                                            \
    if (zh_err_ != hipSuccess) {                                                            \
      char zh_buf_[512];                                                                    \
      snprintf(zh_buf_, sizeof(zh_buf_), "%s failed: %s (%s:%d)", #expr,                    \
               hipGetErrorString(zh_err_), __FILE__, __LINE__);                             \
      throw ::zmlhip::HipError(zh_err_, zh_buf_);                                           \
    }                                                                                       \
  } while (0)

#define ZH_LAUNCH_CHECK() ZH_CHECK(hipGetLastError())

// Last error message (per thread), readable through zmlhip_last_error().
void set_last_error(const std::string& msg);
const char* last_error();

// Growable device workspace owned by the library, one per (thread, device).
// Handed-out pointers stay valid until the next ensure() with a larger size.
void* workspace(size_t bytes, int slot = 0);

// Stream used by the synchronous host-pointer entry points of this thread.
hipStream_t host_api_stream();

// Internal streams (default priority) and timing-free events of this thread
// on the current device, used to overlap panel factorization with updates.
hipStream_t aux_stream(int idx);
hipEvent_t aux_event(int idx);
// A library stream whose kernels leave the first `reserved` CUs free for
// latency-bound work on other streams (one per reservation, device and host thread).
hipStream_t stream_sparing_cus(int reserved = 8);

// Optional per-launch timing of the dominant kernels (kinds below), with HIP
// events recorded on the stream each kernel is launched on.
enum ProfKind : int { kProfGemmF64 = 1, kProfGemmLowp = 2, kProfPanel = 3 };
bool profiling();
void profile_enable(bool on);
void* profile_begin(int kind, hipStream_t s);
void profile_end(void* token, hipStream_t s, double flops, double bytes);
void profile_collect(int kind, double* ms, long long* launches, double* flops, double* bytes);

// Number of compute units of the current device.
int device_cu_count();

// Ordering of calls that share this thread's workspaces, internal streams and
// events: call_begin makes `s` wait for the end of the previous call (whatever
// stream it ran on), call_end marks the end of this one on `s`. Every device
// entry point brackets its work with them, so two calls queued on two streams
// never use the same scratch memory at the same time.
void call_begin(hipStream_t s);
void call_end(hipStream_t s);

// Spin limit of the LU panel's cross-workgroup waits (polls before giving up).
// A negative value makes every multi-workgroup panel give up at its first
// exchange, which tests use to exercise the error path with kernels that all
// exit normally.
long panel_spin_limit();
void set_panel_spin_limit(long polls);
// FP64 LU block columns with more rows than this take the one-launch panel.
long panel_one_launch_rows();
void set_panel_one_launch_rows(long rows);
// FP64 GEMMs whose shapes allow it move their operand slabs global -> LDS with
// the load unit (default); off, every FP64 GEMM stages them through registers.
// Both kernels give the same results bit for bit.
bool gemm_dma_enabled();
void set_gemm_dma(int on);
// Multi-workgroup LU leaves start the exchange of column k + 1 before the
// update of column k's tail (panel.hpp leaf_la_kernel; default), or finish
// each column first (leaf_kernel). Same values and pivots bit for bit.
bool leaf_lookahead();
void set_leaf_lookahead(int on);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---- kernels exposed between translation units -----------------------------

// Element-type generic forms (T = double or float) of the column-major
// kernels below; the _f64 names are their FP64 instances.
template <typename T>
void gemm_cm(int transa, int transb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda,
             const T* b, int64_t ldb, T beta, T* c, int64_t ldc, hipStream_t s, int tri);
template <typename T>
void scale_cm(int64_t m, int64_t n, T beta, T* c, int64_t ldc, hipStream_t s, int tri);
template <typename T>
void trsm_cm(int side, int uplo, int transa, int diag, int64_t m, int64_t n, T alpha, const T* a,
             int64_t lda, T* b, int64_t ldb, hipStream_t s);
template <typename T>
void syrk_cm(int uplo, int trans, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, T beta, T* c,
             int64_t ldc, hipStream_t s);
template <typename T>
void potrf_cm(int uplo, int64_t n, T* a, int64_t lda, int* info, hipStream_t s);

// Column-major FP64 GEMM: C = alpha * op(A) * op(B) + beta * C.
// tri: 0 full, kLower / kUpper restrict the update to that triangle of C
// (tiles outside it are skipped, diagonal tiles are masked), used by SYRK.
void gemm_f64_cm(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                 double* c, int64_t ldc, hipStream_t s, int tri = 0);

// Column-major triangular solve (zml trsm semantics, column-major view).
void trsm_f64_cm(int side, int uplo, int transa, int diag, int64_t m, int64_t n, double alpha,
                 const double* a, int64_t lda, double* b, int64_t ldb, hipStream_t s);

// Column-major SYRK: C = alpha * op(A) op(A)^T + beta * C on one triangle.
void syrk_f64_cm(int uplo, int trans, int64_t n, int64_t k, double alpha, const double* a,
                 int64_t lda, double beta, double* c, int64_t ldc, hipStream_t s);

// Scale an m x n column-major block: c = beta * c (beta == 0 writes zeros).
void scale_f64_cm(int64_t m, int64_t n, double beta, double* c, int64_t ldc, hipStream_t s,
                  int tri = 0);

// Row interchanges on a row-major matrix: for i = k1, k1+inc, ..., k2 (0-based,
// inclusive, inc = +1/-1) swap rows i and ipiv[ix]-1 (ipiv 1-based) over
// columns [0, ncols); ix starts at ix0 and advances by incx (default: ix = i).
void laswp_rm(int64_t ncols, double* a, int64_t lda, int k1, int k2, const int32_t* ipiv,
              int inc, hipStream_t s, int ix0 = -1, int incx = 0);
// Same for a column-major matrix (rows are strided).
void laswp_cm(int64_t ncols, double* a, int64_t lda, int k1, int k2, const int32_t* ipiv,
              int inc, hipStream_t s, int ix0 = -1, int incx = 0);

// Level-1 and Level-2 helpers (blas_f64.hip).
void iamax_f64(int64_t n, const double* x, int64_t incx, int32_t* out, hipStream_t s);
void scal_f64(int64_t n, double alpha, double* x, int64_t incx, hipStream_t s);
void swap_f64(int64_t n, double* x, int64_t incx, double* y, int64_t incy, hipStream_t s);
void gemv_f64_cm(int trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                 const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s);
// The same GEMV with its partial sums in a static device array of
// kGemvScratchDoubles doubles (no allocation); returns -1, with nothing
// queued, when a no-transpose product has more than that many rows.
constexpr size_t kGemvScratchDoubles = size_t(1) << 20;   // 8 MiB per device
int gemv_f64_cm_noalloc(int trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                        const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s);
// x <- T^-1 x for one right-hand side, T n x n triangular with element (i, j)
// at t[i * si + j * sj] (si or sj must be 1), x strided by incx: one
// persistent launch (blas_f64.hip). A bounded wait that gives up leaves NaN.
void trsv1_f64(int64_t n, const double* t, int64_t si, int64_t sj, bool lower, bool unit, double* x,
               int64_t incx, hipStream_t s);
// The interchanges k1..k2 (inc = +1) or k1 down to k2 (inc = -1) of ipiv
// (1-based, ipiv[k] for row k) applied to one vector of element stride ldx.
void laswp_vec(double* x, int64_t ldx, int k1, int k2, const int32_t* ipiv, int inc, hipStream_t s);
// x <- op(A)^-1 x, A n x n triangular (column-major view), x strided by incx.
void trsv_f64_cm(int uplo, int trans, int diag, int64_t n, const double* a, int64_t lda, double* x,
                 int64_t incx, hipStream_t s);
// LAPACK dlange norms ('M', 'O'/'1', 'I', 'F'/'E') of an m x n column-major
// matrix into *out (device); NaN entries propagate.
void lange_f64_cm(char norm, int64_t m, int64_t n, const double* a, int64_t lda, double* out,
                  hipStream_t s);

// LAPACK drivers (getrf_f64.hip, potrf_f64.hip).
void getrf_f64_rm(int64_t m, int64_t n, double* a, int64_t lda, int32_t* ipiv, int* info,
                  hipStream_t s);
template <typename T>
void getrf_rm(int64_t m, int64_t n, T* a, int64_t lda, int32_t* ipiv, int* info, hipStream_t s);
// The LU panel's give-up flag: cleared at the start of every factorization;
// report_leaf_error writes kErrSync into the device info when it is set.
void clear_leaf_error(hipStream_t s);
void report_leaf_error(int* info, hipStream_t s);
void potrf_f64_cm(int uplo, int64_t n, double* a, int64_t lda, int* info, hipStream_t s);

// In-place transpose of the leading n x n block of a matrix with leading
// dimension ld (switches between row- and column-major storage).
template <typename T>
void transpose_square_inplace(int64_t n, T* a, int64_t ld, hipStream_t s);
// Out-of-place: dst (cols x rows, ldd) = src^T (src rows x cols, lds; row-major view).
template <typename T>
void transpose_copy(int64_t rows, int64_t cols, const T* src, int64_t lds, T* dst, int64_t ldd,
                    hipStream_t s);

// Batched small solves (batched.hip)
struct BatchPlan;
BatchPlan* batch_plan_create(int count, const int* n, const int* nrhs, const int* spd,
                             const double* const* a, const double* const* b, double* const* x,
                             const zmlhip_ir_opts* opts, int autotune, hipStream_t s);
void batch_plan_execute(BatchPlan* P, hipStream_t s);
void batch_plan_status(BatchPlan* P, int* status, hipStream_t s);
void batch_plan_variants(const BatchPlan* P, int* tg, int* r, int* counts);
void batch_plan_destroy(BatchPlan* P);

}  // namespace zmlhip
