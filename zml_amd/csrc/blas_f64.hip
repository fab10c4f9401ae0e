// Level-1/2/3 helpers on the device: TRSM and SYRK (FP64 and FP32), IDAMAX,
// DSCAL, DSWAP, row interchanges (LASWP), transposes, GEMV, TRSV and the
// matrix norms of LANGE.
//
// TRSM is blocked: 32 x 32 diagonal blocks are solved by substitution (one
// thread per right-hand-side vector, the block staged in LDS), and the rest
// of the right-hand sides are updated with the MFMA GEMM. The arithmetic
// choices follow zml's trsm.zig where they are visible in a result: the
// left-side branches divide by the diagonal (trsm.zig:137-142,221-229,
// 298-305,364-371), the right-side branches multiply by its reciprocal
// (trsm.zig:458-486,538-566,575-600,653-678).
#include "zmlhip_internal.hpp"

#include <algorithm>
#include <mutex>

namespace zmlhip {
namespace {

constexpr int TB = 32;   // diagonal block size of the blocked TRSM

// Solves M x = b for each vector x of a 32-row block (left side: vectors are
// columns of B; right side: rows of B, with M the transpose of op(T)), in the
// column-sweep order of zml's no-transpose loops.
// forward: M lower (k ascending); else M upper (k descending).
template <typename T, bool LEFT>
__global__ void __launch_bounds__(256)
trsm_diag_kernel(int nb, const T* __restrict__ a, int64_t lda, int need_t, int forward,
                 int unit, T* __restrict__ b, int64_t ldb, int64_t nvec) {
  // separately rounded multiply and subtract, as zml's substitution loops
  // (trsm.zig, trsv.zig): exact cancellations stay exact
#pragma clang fp contract(off)
  __shared__ T M[TB][TB + 1];
  __shared__ T dinv[TB];
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int r = e % TB, c = e / TB;
    T v = T(0);
    if (r < nb && c < nb) v = need_t ? a[c + (int64_t)r * lda] : a[r + (int64_t)c * lda];
    M[r][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < TB) dinv[threadIdx.x] = threadIdx.x < nb ? T(1) / M[threadIdx.x][threadIdx.x] : T(1);
  __syncthreads();
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= nvec) return;
  T x[TB];
#pragma unroll
  for (int r = 0; r < TB; ++r) x[r] = r < nb ? (LEFT ? b[r + v * ldb] : b[v + (int64_t)r * ldb]) : T(0);
  if (forward) {
#pragma unroll
    for (int k = 0; k < TB; ++k) {
      if (k < nb) {
        if (!unit) x[k] = LEFT ? x[k] / M[k][k] : x[k] * dinv[k];
#pragma unroll
        for (int i = k + 1; i < TB; ++i) x[i] -= x[k] * M[i][k];
      }
    }
  } else {
#pragma unroll
    for (int k = TB - 1; k >= 0; --k) {
      if (k < nb) {
        if (!unit) x[k] = LEFT ? x[k] / M[k][k] : x[k] * dinv[k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] -= x[k] * M[i][k];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < TB; ++r)
    if (r < nb) {
      if (LEFT) b[r + v * ldb] = x[r];
      else b[v + (int64_t)r * ldb] = x[r];
    }
}

// ---- IDAMAX: first index of max |x| (iamax.zig:11-68; NaN never wins unless x[0] is NaN)
struct AmaxPair { double v; int64_t i; };

__device__ __forceinline__ void amax_combine(double& v, int64_t& i, double v2, int64_t i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

__device__ __forceinline__ void amax_block_reduce(double& v, int64_t& i) {
  for (int off = 32; off > 0; off >>= 1) {
    double v2 = __shfl_down(v, off);
    int64_t i2 = __shfl_down(i, off);
    amax_combine(v, i, v2, i2);
  }
  __shared__ double sv[16];
  __shared__ int64_t si[16];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { sv[w] = v; si[w] = i; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) amax_combine(v, i, sv[k], si[k]);
  }
}

__global__ void iamax_partial_kernel(int64_t n, const double* __restrict__ x, int64_t incx,
                                     AmaxPair* __restrict__ part) {
  double v = -1.0;
  int64_t idx = INT64_MAX;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double a = fabs(x[e * incx]);
    if (a != a) a = -1.0;
    amax_combine(v, idx, a, e);
  }
  amax_block_reduce(v, idx);
  if (threadIdx.x == 0) part[blockIdx.x] = AmaxPair{v, idx};
}

__global__ void iamax_final_kernel(int nparts, const AmaxPair* __restrict__ part,
                                   const double* __restrict__ x, int32_t* __restrict__ out) {
  double v = -1.0;
  int64_t idx = INT64_MAX;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) amax_combine(v, idx, part[k].v, part[k].i);
  amax_block_reduce(v, idx);
  if (threadIdx.x == 0) {
    const double x0 = x[0];
    *out = (x0 != x0 || idx == INT64_MAX) ? 0 : (int32_t)idx;
  }
}

__global__ void scal_kernel(int64_t n, double alpha, double* __restrict__ x, int64_t incx) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    x[e * incx] = x[e * incx] * alpha;
}

__global__ void swap_kernel(int64_t n, double* __restrict__ x, int64_t incx, double* __restrict__ y,
                            int64_t incy) {
  const int64_t kx = incx < 0 ? (-n + 1) * incx : 0, ky = incy < 0 ? (-n + 1) * incy : 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double* px = x + kx + e * incx;
    double* py = y + ky + e * incy;
    const double t = *px;
    *px = *py;
    *py = t;
  }
}

// Row-major rows are contiguous: thread = column, swaps applied in order.
// Rows i = k1, k1+inc, ..., k2 are exchanged with ipiv[ix] - 1 where ix starts
// at ix0 and advances by incx (laswp.zig:26-40).
__global__ void laswp_rm_kernel(int64_t ncols, double* __restrict__ a, int64_t lda, int k1, int k2,
                                const int32_t* __restrict__ ipiv, int inc, int ix0, int incx) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int i_end = k2 + inc;
  int ix = ix0;
  for (int i = k1; i != i_end; i += inc, ix += incx) {
    const int ip = ipiv[ix] - 1;
    if (ip != i) {
      double* p = a + (int64_t)i * lda + c;
      double* q = a + (int64_t)ip * lda + c;
      const double t = *p;
      *p = *q;
      *q = t;
    }
  }
}

__global__ void laswp_cm_kernel(int64_t ncols, double* __restrict__ a, int64_t lda, int k1, int k2,
                                const int32_t* __restrict__ ipiv, int inc, int ix0, int incx) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  double* col = a + c * lda;
  const int i_end = k2 + inc;
  int ix = ix0;
  for (int i = k1; i != i_end; i += inc, ix += incx) {
    const int ip = ipiv[ix] - 1;
    if (ip != i) {
      const double t = col[i];
      col[i] = col[ip];
      col[ip] = t;
    }
  }
}

constexpr int TT = 32;   // transpose tile

template <typename T>
__global__ void transpose_square_kernel(int64_t n, T* __restrict__ a, int64_t ld) {
  __shared__ T t1[TT][TT + 1], t2[TT][TT + 1];
  const int64_t I = blockIdx.y, J = blockIdx.x;
  if (I > J) return;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int r = ty; r < TT; r += 8) {
    const int64_t i1 = I * TT + r, j1 = J * TT + tx;
    if (i1 < n && j1 < n) t1[r][tx] = a[i1 * ld + j1];
    const int64_t i2 = J * TT + r, j2 = I * TT + tx;
    if (I != J && i2 < n && j2 < n) t2[r][tx] = a[i2 * ld + j2];
  }
  __syncthreads();
  for (int r = ty; r < TT; r += 8) {
    // block (J, I) <- transpose of block (I, J)
    const int64_t i2 = J * TT + r, j2 = I * TT + tx;
    if (i2 < n && j2 < n) a[i2 * ld + j2] = t1[tx][r];
    if (I != J) {
      const int64_t i1 = I * TT + r, j1 = J * TT + tx;
      if (i1 < n && j1 < n) a[i1 * ld + j1] = t2[tx][r];
    }
  }
}

// dst[c * ldd + r] = src[r * lds + c] for r < rows, c < cols.
template <typename T>
__global__ void transpose_copy_kernel(int64_t rows, int64_t cols, const T* __restrict__ src,
                                      int64_t lds, T* __restrict__ dst, int64_t ldd) {
  __shared__ T t[TT][TT + 1];
  const int64_t R = blockIdx.y, Cb = blockIdx.x;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int r = ty; r < TT; r += 8) {
    const int64_t i = R * TT + r, j = Cb * TT + tx;
    if (i < rows && j < cols) t[r][tx] = src[i * lds + j];
  }
  __syncthreads();
  for (int r = ty; r < TT; r += 8) {
    const int64_t j = Cb * TT + r, i = R * TT + tx;
    if (i < rows && j < cols) dst[j * ldd + i] = t[tx][r];
  }
}

inline int grid_for(int64_t n, int threads = 256, int cap = 65536) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + threads - 1) / threads, cap));
}

}  // namespace

// ---------------------------------------------------------------------------
template <typename T>
void trsm_cm(int side, int uplo, int transa, int diag, int64_t m, int64_t n, T alpha, const T* a,
             int64_t lda, T* b, int64_t ldb, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  if (alpha == T(0)) { scale_cm<T>(m, n, T(0), b, ldb, s, 0); return; }   // trsm.zig:99-113
  if (alpha != T(1)) scale_cm<T>(m, n, alpha, b, ldb, s, 0);
  const bool nt = is_notrans(transa);
  const bool unit = diag == kUnit;
  if (side == kLeft) {
    // op(A) is m x m; solve op(A) X = B block row by block row.
    const bool fwd = (uplo == kLower) == nt;   // op(A) lower -> forward
    const int64_t nblk = (m + TB - 1) / TB;
    for (int64_t bi = 0; bi < nblk; ++bi) {
      const int64_t blk = fwd ? bi : nblk - 1 - bi;
      const int64_t i0 = blk * TB, nb = std::min<int64_t>(TB, m - i0);
      hipLaunchKernelGGL((trsm_diag_kernel<T, true>), dim3(grid_for(n)), dim3(256), 0, s, (int)nb,
                         a + i0 + i0 * lda, lda, nt ? 0 : 1, fwd ? 1 : 0, unit ? 1 : 0,
                         b + i0, ldb, n);
      ZH_LAUNCH_CHECK();
      // update the remaining rows: B_rest -= op(A)_{rest, blk} X_blk
      const int64_t r0 = fwd ? i0 + nb : 0, rn = fwd ? m - (i0 + nb) : i0;
      if (rn > 0) {
        if (nt) gemm_cm<T>(kNoTrans, kNoTrans, rn, n, nb, T(-1), a + r0 + i0 * lda, lda, b + i0, ldb, T(1), b + r0, ldb, s, 0);
        else gemm_cm<T>(kTrans, kNoTrans, rn, n, nb, T(-1), a + i0 + r0 * lda, lda, b + i0, ldb, T(1), b + r0, ldb, s, 0);
      }
    }
  } else {
    // X op(A) = B with op(A) n x n, block column by block column.
    const bool fwd = (uplo == kUpper) == nt;   // op(A) upper -> X_j depends on X_<j
    const int64_t nblk = (n + TB - 1) / TB;
    for (int64_t bj = 0; bj < nblk; ++bj) {
      const int64_t blk = fwd ? bj : nblk - 1 - bj;
      const int64_t j0 = blk * TB, nb = std::min<int64_t>(TB, n - j0);
      // rows of B solve x M = b, i.e. M^T x^T = b^T with M = op(A_jj)
      hipLaunchKernelGGL((trsm_diag_kernel<T, false>), dim3(grid_for(m)), dim3(256), 0, s, (int)nb,
                         a + j0 + j0 * lda, lda, nt ? 1 : 0, fwd ? 1 : 0, unit ? 1 : 0,
                         b + j0 * ldb, ldb, m);
      ZH_LAUNCH_CHECK();
      const int64_t c0 = fwd ? j0 + nb : 0, cn = fwd ? n - (j0 + nb) : j0;
      if (cn > 0) {
        if (nt) gemm_cm<T>(kNoTrans, kNoTrans, m, cn, nb, T(-1), b + j0 * ldb, ldb, a + j0 + c0 * lda, lda, T(1), b + c0 * ldb, ldb, s, 0);
        else gemm_cm<T>(kNoTrans, kTrans, m, cn, nb, T(-1), b + j0 * ldb, ldb, a + c0 + j0 * lda, lda, T(1), b + c0 * ldb, ldb, s, 0);
      }
    }
  }
}
template void trsm_cm<double>(int, int, int, int, int64_t, int64_t, double, const double*, int64_t, double*,
                              int64_t, hipStream_t);
template void trsm_cm<float>(int, int, int, int, int64_t, int64_t, float, const float*, int64_t, float*,
                             int64_t, hipStream_t);

void trsm_f64_cm(int side, int uplo, int transa, int diag, int64_t m, int64_t n, double alpha,
                 const double* a, int64_t lda, double* b, int64_t ldb, hipStream_t s) {
  trsm_cm<double>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, s);
}

template <typename T>
void syrk_cm(int uplo, int trans, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, T beta, T* c,
             int64_t ldc, hipStream_t s) {
  if (n <= 0) return;
  const int ta = is_notrans(trans) ? kNoTrans : kTrans;
  const int tb = is_notrans(trans) ? kTrans : kNoTrans;
  gemm_cm<T>(ta, tb, n, n, k, alpha, a, lda, a, lda, beta, c, ldc, s, uplo);
}
template void syrk_cm<double>(int, int, int64_t, int64_t, double, const double*, int64_t, double, double*,
                              int64_t, hipStream_t);
template void syrk_cm<float>(int, int, int64_t, int64_t, float, const float*, int64_t, float, float*,
                             int64_t, hipStream_t);

void syrk_f64_cm(int uplo, int trans, int64_t n, int64_t k, double alpha, const double* a,
                 int64_t lda, double beta, double* c, int64_t ldc, hipStream_t s) {
  syrk_cm<double>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, s);
}

void laswp_rm(int64_t ncols, double* a, int64_t lda, int k1, int k2, const int32_t* ipiv, int inc,
              hipStream_t s, int ix0, int incx) {
  if (ncols <= 0) return;
  if (ix0 < 0) { ix0 = k1; incx = inc; }
  hipLaunchKernelGGL(laswp_rm_kernel, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, s,
                     ncols, a, lda, k1, k2, ipiv, inc, ix0, incx);
  ZH_LAUNCH_CHECK();
}

void laswp_cm(int64_t ncols, double* a, int64_t lda, int k1, int k2, const int32_t* ipiv, int inc,
              hipStream_t s, int ix0, int incx) {
  if (ncols <= 0) return;
  if (ix0 < 0) { ix0 = k1; incx = inc; }
  hipLaunchKernelGGL(laswp_cm_kernel, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, s,
                     ncols, a, lda, k1, k2, ipiv, inc, ix0, incx);
  ZH_LAUNCH_CHECK();
}

template <typename T>
void transpose_square_inplace(int64_t n, T* a, int64_t ld, hipStream_t s) {
  if (n <= 1) return;
  const unsigned t = (unsigned)((n + TT - 1) / TT);
  hipLaunchKernelGGL(transpose_square_kernel<T>, dim3(t, t), dim3(TT, 8), 0, s, n, a, ld);
  ZH_LAUNCH_CHECK();
}
template void transpose_square_inplace<double>(int64_t, double*, int64_t, hipStream_t);
template void transpose_square_inplace<float>(int64_t, float*, int64_t, hipStream_t);

template <typename T>
void transpose_copy(int64_t rows, int64_t cols, const T* src, int64_t lds, T* dst, int64_t ldd, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(transpose_copy_kernel<T>,
                     dim3((unsigned)((cols + TT - 1) / TT), (unsigned)((rows + TT - 1) / TT)),
                     dim3(TT, 8), 0, s, rows, cols, src, lds, dst, ldd);
  ZH_LAUNCH_CHECK();
}
template void transpose_copy<double>(int64_t, int64_t, const double*, int64_t, double*, int64_t, hipStream_t);
template void transpose_copy<float>(int64_t, int64_t, const float*, int64_t, float*, int64_t, hipStream_t);

// ---- Level 1 -----------------------------------------------------------------
void iamax_f64(int64_t n, const double* x, int64_t incx, int32_t* out, hipStream_t s) {
  const int parts = (int)std::min<int64_t>(256, (n + 255) / 256);
  AmaxPair* part = static_cast<AmaxPair*>(workspace(sizeof(AmaxPair) * 256, 7));
  hipLaunchKernelGGL(iamax_partial_kernel, dim3(parts), dim3(256), 0, s, n, x, incx, part);
  ZH_LAUNCH_CHECK();
  hipLaunchKernelGGL(iamax_final_kernel, dim3(1), dim3(256), 0, s, parts, part, x, out);
  ZH_LAUNCH_CHECK();
}

void scal_f64(int64_t n, double alpha, double* x, int64_t incx, hipStream_t s) {
  if (n <= 0 || alpha == 1.0) return;   // scal.zig:24
  hipLaunchKernelGGL(scal_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, alpha, x, incx);
  ZH_LAUNCH_CHECK();
}

void swap_f64(int64_t n, double* x, int64_t incx, double* y, int64_t incy, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(swap_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, x, incx, y, incy);
  ZH_LAUNCH_CHECK();
}

}  // namespace zmlhip

// ---- TRSV and LANGE (column-major view) ----------------------------------------
namespace zmlhip {
namespace {

__global__ void gather_kernel(int64_t n, const double* __restrict__ x, int64_t incx, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t kx = incx < 0 ? (-n + 1) * incx : 0;
  y[i] = x[kx + i * incx];
}
__global__ void scatter_kernel(int64_t n, const double* __restrict__ y, double* __restrict__ x, int64_t incx) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t kx = incx < 0 ? (-n + 1) * incx : 0;
  x[kx + i * incx] = y[i];
}

// max that keeps NaN (LAPACK's dlange propagates a NaN entry)
__device__ __forceinline__ double nmax(double a, double b) { return (b > a || b != b) ? b : a; }

// Block reduction of v (max or sum) into thread 0's return value (256 threads).
__device__ __forceinline__ double block_reduce(double v, bool sum) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    v = sum ? v + o : nmax(v, o);
  }
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = sw[0];
  for (int w = 1; w < 4; ++w) t = sum ? t + sw[w] : nmax(t, sw[w]);
  return t;
}

// part[block] = max |a_ij| (sq == 0) or sum (|a_ij| / *scale)^2 (sq == 1) over
// the elements the block visits (i fastest, so the reads are coalesced).
__global__ void lange_elem_kernel(int sq, int64_t m, int64_t n, const double* __restrict__ a, int64_t lda,
                                  const double* __restrict__ scale, double* __restrict__ part) {
  const double inv = sq ? 1.0 / *scale : 1.0;
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % m, j = e / m;
    const double v = fabs(a[i + j * lda]);
    if (sq) { const double q = v * inv; acc += q * q; }
    else acc = nmax(acc, v);
  }
  const double t = block_reduce(acc, sq != 0);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// colsum[j] = sum_i |a_ij| (one block per column, ascending partial sums per thread)
__global__ void lange_colsum_kernel(int64_t m, const double* __restrict__ a, int64_t lda, double* __restrict__ colsum) {
  const double* col = a + blockIdx.x * lda;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) acc += fabs(col[i]);
  const double t = block_reduce(acc, true);
  if (threadIdx.x == 0) colsum[blockIdx.x] = t;
}

// part[s * m + i] = sum over the columns of split s of |a_ij|
__global__ void lange_rowsum_partial(int64_t m, int64_t n, const double* __restrict__ a, int64_t lda,
                                     double* __restrict__ part, int64_t cps) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t sidx = blockIdx.y, j0 = sidx * cps, j1 = j0 + cps < n ? j0 + cps : n;
  if (i >= m) return;
  double acc = 0.0;
  for (int64_t j = j0; j < j1; ++j) acc += fabs(a[i + j * lda]);
  part[sidx * m + i] = acc;
}
__global__ void lange_rowsum_final(int64_t m, int nsplit, const double* __restrict__ part, double* __restrict__ rows) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  double acc = 0.0;
  for (int sp = 0; sp < nsplit; ++sp) acc += part[sp * m + i];
  rows[i] = acc;
}

// *out = max (or sum) of x[0..cnt); mode 2: *out = *scale * sqrt(sum) (the
// Frobenius norm from its scaled sum of squares; a zero, infinite or NaN scale
// is the result itself)
__global__ void lange_final_kernel(int mode, int64_t cnt, const double* __restrict__ x,
                                   const double* __restrict__ scale, double* __restrict__ out) {
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < cnt; e += blockDim.x) acc = mode == 0 ? nmax(acc, x[e]) : acc + x[e];
  const double t = block_reduce(acc, mode != 0);
  if (threadIdx.x == 0) {
    if (mode == 2) {
      const double sc = *scale;
      *out = (sc == 0.0 || !(sc < INFINITY)) ? sc : sc * sqrt(t);
    } else {
      *out = t;
    }
  }
}

}  // namespace

// ---- one right-hand side: persistent triangular solve and vector interchanges ----
namespace {

// x <- T^-1 x for one right-hand side, T n x n triangular with element (i, j)
// at t[i * si + j * sj]. One persistent launch: T is cut into 64-row blocks,
// workgroup g takes the block rows g, g + G, ... in solve order (top down for
// lower, bottom up for upper). For its block row it accumulates
//   s_i = sum over the blocks solved before it of T_ij x_j,
// taking each block's x as soon as its owner has published it (several ready
// blocks at a time), solves its diagonal block (one wave, x_j = (b_j - s_j) *
// fl(1 / T_jj), or without the scaling for a unit diagonal) and publishes x of its
// rows as tagged 64-bit words (tag << 32 | half of the double), read with
// device-scope relaxed atomics: a reader knows a word is current from its tag.
// The grid never exceeds the co-resident workgroups and a block row waits only
// on block rows earlier in the solve order, so every wait ends; a wait that
// exceeds its bound publishes NaN instead, which makes the result visibly wrong
// rather than silently so. ROW: sj == 1, a row's elements are contiguous (four
// threads share a row, 16 columns each); otherwise si == 1 and the lanes run
// down a column.
constexpr int T1B = 64;        // rows per block
constexpr int T1BATCH = 8;     // published blocks taken per pass
__device__ __forceinline__ void t1_put(uint64_t* p, uint32_t tag, double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  __hip_atomic_store(p, ((uint64_t)tag << 32) | (uint32_t)(b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, ((uint64_t)tag << 32) | (uint32_t)b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool t1_get(const uint64_t* p, uint32_t tag, double& v) {
  const uint64_t hi = __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t lo = __hip_atomic_load(const_cast<uint64_t*>(p + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  v = __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
  return (uint32_t)(hi >> 32) == tag && (uint32_t)(lo >> 32) == tag;
}

template <bool ROW>
__global__ void __launch_bounds__(256)
trsv1_kernel(int64_t n, const double* __restrict__ t, int64_t si, int64_t sj, int lower, int unit,
             double* __restrict__ x, int64_t incx, uint64_t* __restrict__ pub, uint32_t tag, long spin_limit) {
#pragma clang fp contract(off)
  __shared__ double D[T1B][T1B + 1];
  __shared__ double xs[T1BATCH][T1B];
  __shared__ double part[4][T1B];
  __shared__ int s_cnt, s_bad;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t nblk = (n + T1B - 1) / T1B;
  const int64_t xoff = incx < 0 ? (1 - n) * incx : 0;
  for (int64_t bi = blockIdx.x; bi < nblk; bi += gridDim.x) {
    const int64_t b = lower ? bi : nblk - 1 - bi;
    const int64_t r0 = b * T1B;
    const int rows = (int)(n - r0 < T1B ? n - r0 : T1B);
    // the diagonal block (its triangle), read while the contributions come in
    for (int e = tid; e < T1B * T1B; e += 256) {
      const int i = e / T1B, j = e % T1B;
      const bool in = i < rows && j < rows && (lower ? j <= i : j >= i);
      D[i][j] = in ? t[(r0 + i) * si + (r0 + j) * sj] : 0.0;
    }
    // ROW: thread -> row tid / 4, columns 16 (tid % 4) .. +15 of each block;
    // otherwise: thread -> row tid % 64, columns 16 (tid / 64) .. +15
    const int myrow = ROW ? tid >> 2 : lane, q = ROW ? tid & 3 : wv;
    const int64_t rowc = r0 + myrow < n ? r0 + myrow : n - 1;   // rows past n read row n - 1, unused
    double acc = 0.0;
    if (tid == 0) s_bad = 0;
    auto load16 = [&](int64_t jblk, double (&v)[16]) {
      const int64_t c0 = jblk * T1B + 16 * q;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int64_t col = c0 + k < n ? c0 + k : n - 1;
        v[k] = t[rowc * si + col * sj];
      }
    };
    for (int64_t jj = 0; jj < bi;) {
      // the factor's values of the next block do not depend on its x: they
      // are requested before waiting for it
      double vp[16];
      load16(lower ? jj : nblk - 1 - jj, vp);
      // wave 0 takes the next published blocks (the first one waited for)
      if (wv == 0) {
        // a block is looked at through one word (its last row's, written
        // last) until that is current, then read whole; the many waiting
        // workgroups then poll one word each instead of a block's 128
        int cnt = 0;
        long spins = 0;
        while (cnt < T1BATCH && jj + cnt < bi) {
          const int64_t jb = lower ? jj + cnt : nblk - 1 - (jj + cnt);
          const int64_t c0 = jb * T1B;
          const int64_t last = (c0 + T1B < n ? c0 + T1B : n) - 1;
          const uint64_t w = __hip_atomic_load(pub + 2 * last + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          bool ok = (uint32_t)(w >> 32) == tag;
          double v = 0.0;
          if (ok) ok = __all(c0 + lane >= n || t1_get(pub + 2 * (c0 + lane), tag, v));
          if (ok) {
            xs[cnt][lane] = c0 + lane < n ? v : 0.0;
            ++cnt;
            spins = 0;
            continue;
          }
          if (cnt > 0) break;                 // take what is there
          if (++spins > spin_limit) {         // give up: this block row becomes NaN
            if (lane == 0) s_bad = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(4);
        }
        if (lane == 0) s_cnt = cnt;
      }
      __syncthreads();
      const int cnt = s_cnt;
      if (s_bad) break;
#pragma unroll
      for (int k = 0; k < 16; ++k) acc += vp[k] * xs[0][16 * q + k];
      for (int c = 1; c < cnt; ++c) {
        double v[16];
        load16(lower ? jj + c : nblk - 1 - (jj + c), v);
#pragma unroll
        for (int k = 0; k < 16; ++k) acc += v[k] * xs[c][16 * q + k];
      }
      jj += cnt;
      __syncthreads();   // xs is refilled next pass
    }
    // the four partial sums of each row
    if (ROW) {
      acc += __shfl_xor(acc, 1);
      acc += __shfl_xor(acc, 2);
      if (q == 0) part[0][myrow] = acc;
    } else {
      part[q][myrow] = acc;
    }
    __syncthreads();
    if (wv == 0) {
      const bool bad = s_bad != 0;
      double ti = 0.0;
      if (lane < rows) {
        const double s = ROW ? part[0][lane] : part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
        ti = x[xoff + (r0 + lane) * incx] - s;
      }
      // substitution inside the block, one column at a time. The diagonal's
      // reciprocals are formed in parallel beforehand (x_k = t_k fl(1/d_k):
      // a division per step would sit on the serial chain), and each lane's
      // next element of the block is read from LDS one step ahead, so only the
      // broadcast and two operations per step are serial.
      const double rinv = (!unit && lane < rows) ? 1.0 / D[lane][lane] : 1.0;
      int kk = lower ? 0 : rows - 1;
      double dcur = D[lane][kk];
      for (int k = 0; k < rows; ++k) {
        const int kn = lower ? kk + 1 : kk - 1;
        const int kc = k + 1 < rows ? kn : kk;   // (an index inside the block either way)
        const double dnext = D[lane][kc];
        const uint64_t tb = (uint64_t)__double_as_longlong(ti), rb = (uint64_t)__double_as_longlong(rinv);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(tb >> 32), kk);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)tb, kk);
        const uint32_t rhi = (uint32_t)__builtin_amdgcn_readlane((int)(rb >> 32), kk);
        const uint32_t rlo = (uint32_t)__builtin_amdgcn_readlane((int)rb, kk);
        const double xk = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo)) *
                          __longlong_as_double((long long)(((uint64_t)rhi << 32) | rlo));
        if (lane == kk) ti = xk;
        const bool below = lower ? lane > kk : lane < kk;
        if (below && lane < rows) ti = ti - dcur * xk;
        kk = kn;
        dcur = dnext;
      }
      if (bad) ti = __builtin_nan("");
      if (lane < rows) {
        x[xoff + (r0 + lane) * incx] = ti;
        t1_put(pub + 2 * (r0 + lane), tag, ti);
      }
    }
    __syncthreads();   // D and part are reused by the next block row
  }
}

// The interchanges i <-> ipiv[ix] - 1 (k1..k2 in order, forward or backward)
// applied to one vector with element stride ldx, by one workgroup: 256
// interchanges at a time, the elements they touch gathered into LDS (an
// open-addressing table gives each distinct row a slot), swapped there in order
// by one lane, and scattered back.
constexpr int LV_CH = 256, LV_HT = 2048;
__global__ void __launch_bounds__(256)
laswp_vec_kernel(double* __restrict__ xv, int64_t ldx, int k1, int k2, const int32_t* __restrict__ ipiv,
                 int inc, int ix0, int incx) {
  __shared__ int hkey[LV_HT], hslot[LV_HT];
  __shared__ int rowof[2 * LV_CH], sa[LV_CH], sb[LV_CH];
  __shared__ double val[2 * LV_CH];
  __shared__ int s_n;
  const int tid = threadIdx.x;
  const int nsteps = inc > 0 ? k2 - k1 + 1 : k1 - k2 + 1;   // (k1..k2 stepping by inc)
  for (int c0 = 0; c0 < nsteps; c0 += LV_CH) {
    const int cn = nsteps - c0 < LV_CH ? nsteps - c0 : LV_CH;
    for (int h = tid; h < LV_HT; h += 256) hkey[h] = -1;
    if (tid == 0) s_n = 0;
    __syncthreads();
    int ia = -1, ib = -1;
    if (tid < cn) {
      const int st = c0 + tid;
      ia = inc > 0 ? k1 + st : k1 - st;
      ib = ipiv[ix0 + st * incx] - 1;
    }
    // every distinct row of the chunk enters the table
    for (int w = 0; w < 2; ++w) {
      const int r = w == 0 ? ia : ib;
      if (r >= 0) {
        uint32_t h = ((uint32_t)r * 2654435761u) & (LV_HT - 1);
        for (;;) {
          const int prev = atomicCAS(&hkey[h], -1, r);
          if (prev == -1 || prev == r) break;
          h = (h + 1) & (LV_HT - 1);
        }
      }
    }
    __syncthreads();
    for (int h = tid; h < LV_HT; h += 256) {
      const int key = hkey[h];
      if (key >= 0) {
        const int sl = atomicAdd(&s_n, 1);
        hslot[h] = sl;
        rowof[sl] = key;
        val[sl] = xv[(int64_t)key * ldx];
      }
    }
    __syncthreads();
    if (tid < cn) {
      uint32_t h = ((uint32_t)ia * 2654435761u) & (LV_HT - 1);
      while (hkey[h] != ia) h = (h + 1) & (LV_HT - 1);
      sa[tid] = hslot[h];
      h = ((uint32_t)ib * 2654435761u) & (LV_HT - 1);
      while (hkey[h] != ib) h = (h + 1) & (LV_HT - 1);
      sb[tid] = hslot[h];
    }
    __syncthreads();
    if (tid == 0) {
      for (int st = 0; st < cn; ++st) {
        const int a = sa[st], bsl = sb[st];
        if (a != bsl) {
          const double tmp = val[a];
          val[a] = val[bsl];
          val[bsl] = tmp;
        }
      }
    }
    __syncthreads();
    const int nn = s_n;
    for (int sl = tid; sl < nn; sl += 256) xv[(int64_t)rowof[sl] * ldx] = val[sl];
    __threadfence_block();
    __syncthreads();
  }
}

}  // namespace

struct Trsv1State {
  int max_resident = 0;
  uint32_t next_tag = 0;
  void* ws = nullptr;
  size_t bytes = 0;
};
static thread_local Trsv1State tl_trsv1[16];

void trsv1_f64(int64_t n, const double* t, int64_t si, int64_t sj, bool lower, bool unit, double* x,
               int64_t incx, hipStream_t s) {
  if (n <= 0) return;
  int dev = 0;
  ZH_CHECK(hipGetDevice(&dev));
  Trsv1State& st = tl_trsv1[dev & 15];
  const int64_t nblk = (n + T1B - 1) / T1B;
  const size_t bytes = sizeof(uint64_t) * 2 * (size_t)nblk * T1B;
  void* ws = workspace(bytes, 25);
  if (ws != st.ws || bytes > st.bytes || st.next_tag > 0x7fff0000u) {   // fresh words: every tag zero
    ZH_CHECK(hipMemsetAsync(ws, 0, bytes, s));
    st.ws = ws;
    st.bytes = bytes;
    st.next_tag = 0;
  }
  const bool row = sj == 1;
  if (st.max_resident == 0) {
    int a = 0, b = 0;
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, trsv1_kernel<true>, 256, 0));
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, trsv1_kernel<false>, 256, 0));
    st.max_resident = std::max(1, std::min(a, b)) * device_cu_count();
  }
  const int grid = (int)std::min<int64_t>(nblk, std::min(st.max_resident, 1024));
  const uint32_t tag = ++st.next_tag;
  const long limit = 1L << 22;
  if (row)
    hipLaunchKernelGGL(trsv1_kernel<true>, dim3(grid), dim3(256), 0, s, n, t, si, sj, lower ? 1 : 0,
                       unit ? 1 : 0, x, incx, static_cast<uint64_t*>(ws), tag, limit);
  else
    hipLaunchKernelGGL(trsv1_kernel<false>, dim3(grid), dim3(256), 0, s, n, t, si, sj, lower ? 1 : 0,
                       unit ? 1 : 0, x, incx, static_cast<uint64_t*>(ws), tag, limit);
  ZH_LAUNCH_CHECK();
}

void laswp_vec(double* x, int64_t ldx, int k1, int k2, const int32_t* ipiv, int inc, hipStream_t s) {
  // the interchange of row k uses ipiv[k], in both orders
  hipLaunchKernelGGL(laswp_vec_kernel, dim3(1), dim3(256), 0, s, x, ldx, k1, k2, ipiv, inc, k1, inc);
  ZH_LAUNCH_CHECK();
}

void trsv_f64_cm(int uplo, int trans, int diag, int64_t n, const double* a, int64_t lda, double* x,
                 int64_t incx, hipStream_t s) {
  if (n <= 0) return;
  // the vector as an n x 1 matrix: the left-side TRSM divides by the diagonal,
  // as trsv.zig does in every branch
  double* y = static_cast<double*>(workspace(sizeof(double) * (size_t)n, 19));
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, (const double*)x, incx, y);
  ZH_LAUNCH_CHECK();
  trsm_cm<double>(kLeft, uplo, trans, diag, n, 1, 1.0, a, lda, y, n, s);
  hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, (const double*)y, x, incx);
  ZH_LAUNCH_CHECK();
}

void lange_f64_cm(char norm, int64_t m, int64_t n, const double* a, int64_t lda, double* out, hipStream_t s) {
  constexpr int NPART = 256;
  const int64_t nsplit_max = 64;
  double* w = static_cast<double*>(workspace(sizeof(double) * (NPART + 8 + (size_t)std::max<int64_t>(m, n) *
                                                                 (nsplit_max + 1)), 20));
  double* part = w;             // NPART block results
  double* scal = w + NPART;     // Frobenius scale
  double* big = w + NPART + 8;  // column sums, or row-sum partials + row sums
  const unsigned eb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(NPART, (m * n + 255) / 256));
  switch (norm) {
    case 'M': case 'm':
      hipLaunchKernelGGL(lange_elem_kernel, dim3(eb), dim3(256), 0, s, 0, m, n, a, lda, (const double*)nullptr, part);
      hipLaunchKernelGGL(lange_final_kernel, dim3(1), dim3(256), 0, s, 0, (int64_t)eb, (const double*)part,
                         (const double*)nullptr, out);
      break;
    case 'F': case 'f': case 'E': case 'e':
      hipLaunchKernelGGL(lange_elem_kernel, dim3(eb), dim3(256), 0, s, 0, m, n, a, lda, (const double*)nullptr, part);
      hipLaunchKernelGGL(lange_final_kernel, dim3(1), dim3(256), 0, s, 0, (int64_t)eb, (const double*)part,
                         (const double*)nullptr, scal);
      hipLaunchKernelGGL(lange_elem_kernel, dim3(eb), dim3(256), 0, s, 1, m, n, a, lda, (const double*)scal, part);
      hipLaunchKernelGGL(lange_final_kernel, dim3(1), dim3(256), 0, s, 2, (int64_t)eb, (const double*)part,
                         (const double*)scal, out);
      break;
    case 'O': case 'o': case '1':
      hipLaunchKernelGGL(lange_colsum_kernel, dim3((unsigned)n), dim3(256), 0, s, m, a, lda, big);
      hipLaunchKernelGGL(lange_final_kernel, dim3(1), dim3(256), 0, s, 0, n, (const double*)big,
                         (const double*)nullptr, out);
      break;
    case 'I': case 'i': {
      const int64_t rb = (m + 255) / 256;
      const int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>({nsplit_max, n, (2048 + rb - 1) / rb}));
      const int64_t cps = (n + nsplit - 1) / nsplit;
      double* rows = big + nsplit * m;
      hipLaunchKernelGGL(lange_rowsum_partial, dim3((unsigned)rb, (unsigned)nsplit), dim3(256), 0, s, m, n, a, lda,
                         big, cps);
      hipLaunchKernelGGL(lange_rowsum_final, dim3((unsigned)rb), dim3(256), 0, s, m, (int)nsplit, (const double*)big,
                         rows);
      hipLaunchKernelGGL(lange_final_kernel, dim3(1), dim3(256), 0, s, 0, m, (const double*)rows,
                         (const double*)nullptr, out);
      break;
    }
    default:
      throw std::runtime_error("lange: unknown norm");
  }
  ZH_LAUNCH_CHECK();
}

}  // namespace zmlhip

// ---- GEMV (column-major view), used by cblas_dgemv and the FP64 residual ----
namespace zmlhip {
namespace {

constexpr int GV_SPLIT_ROWS = 256;

// y_part[s][i] = sum over columns j in split s of A[i + j lda] x_j (ascending j)
__global__ void gemv_n_partial(int64_t m, int64_t n, const double* __restrict__ a, int64_t lda,
                               const double* __restrict__ x, int64_t incx, double* __restrict__ part,
                               int64_t cols_per_split) {
  const int64_t i = blockIdx.x * (int64_t)GV_SPLIT_ROWS + threadIdx.x;
  const int64_t s = blockIdx.y;
  const int64_t j0 = s * cols_per_split, j1 = std::min<int64_t>(n, j0 + cols_per_split);
  if (i >= m) return;
  const int64_t kx = incx < 0 ? (-n + 1) * incx : 0;
  double acc = 0.0;
  for (int64_t j = j0; j < j1; ++j) acc += a[i + j * lda] * x[kx + j * incx];
  part[s * m + i] = acc;
}

__global__ void gemv_n_final(int64_t m, int nsplit, const double* __restrict__ part, double alpha,
                             double beta, double* __restrict__ y, int64_t incy) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  double acc = 0.0;
  for (int s = 0; s < nsplit; ++s) acc += part[s * m + i];
  const int64_t ky = incy < 0 ? (-m + 1) * incy : 0;
  double* py = y + ky + i * incy;
  const double v = alpha * acc;
  *py = beta == 0.0 ? v : beta * *py + v;
}

// y_j = alpha * sum_i A[i + j lda] x_i + beta y_j ; one workgroup per column.
__global__ void gemv_t_kernel(int64_t m, int64_t n, double alpha, const double* __restrict__ a,
                              int64_t lda, const double* __restrict__ x, int64_t incx, double beta,
                              double* __restrict__ y, int64_t incy) {
  const int64_t j = blockIdx.x;
  const int64_t kx = incx < 0 ? (-m + 1) * incx : 0;
  const double* col = a + j * lda;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) acc += col[i] * x[kx + i * incx];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double tot = sw[0] + sw[1] + sw[2] + sw[3];
    const int64_t ky = incy < 0 ? (-n + 1) * incy : 0;
    double* py = y + ky + j * incy;
    const double v = alpha * tot;
    *py = beta == 0.0 ? v : beta * *py + v;
  }
}

}  // namespace

// Column splits of the no-transpose GEMV: about 2048 workgroups' worth of
// partial sums (its scratch is nsplit x m doubles).
static int gemv_splits(int64_t m, int64_t n) {
  const int64_t rb = (m + GV_SPLIT_ROWS - 1) / GV_SPLIT_ROWS;
  const int nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(64, (2048 + rb - 1) / rb));
  return (int)std::min<int64_t>(nsplit, n);
}

// part: nsplit x m doubles of scratch (the library's workspace when null)
static void gemv_f64_cm_with(int trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                             const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s,
                             double* part, int nsplit) {
  if (m <= 0 || n <= 0 || (alpha == 0.0 && beta == 1.0)) return;
  const bool nt = is_notrans(trans);
  const int64_t leny = nt ? m : n;
  if (alpha == 0.0) {
    // y = beta * y over a strided vector: reuse the scale kernel on a 1 x leny view
    if (incy > 0) scale_f64_cm(1, leny, beta, y, incy, s);
    else scale_f64_cm(1, leny, beta, y + (-leny + 1) * incy, -incy, s);
    return;
  }
  if (nt) {
    const int64_t rb = (m + GV_SPLIT_ROWS - 1) / GV_SPLIT_ROWS;
    const int64_t cps = (n + nsplit - 1) / nsplit;
    if (!part) part = static_cast<double*>(workspace(sizeof(double) * nsplit * m, 6));
    hipLaunchKernelGGL(gemv_n_partial, dim3((unsigned)rb, nsplit), dim3(GV_SPLIT_ROWS), 0, s, m, n,
                       a, lda, x, incx, part, cps);
    ZH_LAUNCH_CHECK();
    hipLaunchKernelGGL(gemv_n_final, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, nsplit,
                       part, alpha, beta, y, incy);
    ZH_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)n), dim3(256), 0, s, m, n, alpha, a, lda, x,
                       incx, beta, y, incy);
    ZH_LAUNCH_CHECK();
  }
}

void gemv_f64_cm(int trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                 const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s) {
  gemv_f64_cm_with(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, s, nullptr, gemv_splits(m, n));
}

// ---- allocation-free GEMV ---------------------------------------------------------
// The partial sums of the no-transpose form live in this file-scope device
// array instead of a library workspace, so the entry point that uses it calls
// no allocator: each device holds its own copy for the life of the process.
// Its address is taken once per device (hipGetSymbolAddress) with that device
// current; the array's name itself is only a host-side stand-in. Every call on
// a device shares its array (calls that overlap on one device race); nothing in
// it has to start at zero (each call writes every partial sum it reads).
static __device__ __align__(256) double scratch[kGemvScratchDoubles];

double* gemv_scratch_address() {
  constexpr int kMaxDevices = 64;
  static double* addr[kMaxDevices] = {};
  static std::mutex mu;
  int dev = 0;
  ZH_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("gemv: device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    ZH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = static_cast<double*>(p);
  }
  return addr[dev];
}

int gemv_f64_cm_noalloc(int trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                        const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s) {
  if (m <= 0 || n <= 0 || (alpha == 0.0 && beta == 1.0)) return 0;
  int nsplit = gemv_splits(m, n);
  double* part = nullptr;
  if (is_notrans(trans) && alpha != 0.0) {
    // the splits the allocating path takes when they fit (same sums, same
    // order), fewer when nsplit x m would exceed the array
    if ((size_t)m > kGemvScratchDoubles) return -1;
    nsplit = (int)std::min<int64_t>(nsplit, (int64_t)(kGemvScratchDoubles / (size_t)m));
    part = gemv_scratch_address();
  }
  gemv_f64_cm_with(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, s, part, nsplit);
  return 0;
}

}  // namespace zmlhip
