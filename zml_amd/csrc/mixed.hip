// Mixed-precision LU solve with FP64 iterative refinement.
//
// No counterpart in zml (the reference has no mixed-precision, refinement or
// GMRES code); the FP64 interfaces it sits behind are dgesv's (gesv.zig).
//
//   1. Equilibrate: r_i, c_j powers of two with max_j |r_i a_ij| and
//      max_i |r_i a_ij c_j| in [1, 2) (exact in binary), A_hat = Dr A Dc,
//      rounded to an FP32 row-major working matrix W.
//   2. Factor P A_hat = L U in W with the blocked look-ahead LU of the FP64
//      path (same panel kernel, partial pivoting on the FP32 values): panels
//      and U12 in FP32, the trailing update in FP32, FP16 or BF16 on the
//      matrix cores with FP32 accumulation. For FP16/BF16 every block column
//      L21 is scaled per row and every block row U12 per column by powers of
//      two into [1, 2) before rounding, and the scales are undone in the GEMM
//      epilogue, so neither operand under- or overflows the 16-bit range.
//   3. x0 = Dc U^-1 L^-1 P Dr b, with blocked triangular solves through the
//      inverses of the diagonal blocks of L and U.
//   4. Refine in FP64: r = b - A x (the original A), d = solve(r), x += d,
//      until ||r||_inf <= tol * ||A||_inf ||x||_inf n u (u = 2^-53). When the
//      backward error stops falling by half per step, or grows, switch to
//      GMRES-IR (restarted FGMRES in FP64, the LU solve as right
//      preconditioner); when that does not reach the target either, solve
//      with the FP64 LU instead.
#include "zmlhip_internal.hpp"
#include "mixed.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/zmlhip.h"

namespace zmlhip {

void gemm_f64_cm(int, int, int64_t, int64_t, int64_t, double, const double*, int64_t,
                 const double*, int64_t, double, double*, int64_t, hipStream_t, int);

namespace {

constexpr int NB_MIX = 256;

// ---- equilibration ---------------------------------------------------------------
__device__ __forceinline__ double pow2_scale(double amax) {
  // 2^-e with e = floor(log2 amax): amax * scale lies in [1, 2)
  if (!(amax > 0.0) || !isfinite(amax)) return 1.0;
  return ldexp(1.0, -ilogb(amax));
}

__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

// rmax[i] = max_j |a_ij| (rs == null) or max over columns of |rs_i a_ij| into cmax[j].
__global__ void absmax_kernel(int layout, int64_t n, const double* __restrict__ a, int64_t lda,
                              const double* __restrict__ rscale, double* __restrict__ rmax,
                              double* __restrict__ cmax, int64_t chunk) {
  // element (i, j); threads walk the contiguous index, blocks.y split the other
  const int64_t fast = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (fast >= n) return;
  const int64_t s0 = blockIdx.y * chunk, s1 = std::min<int64_t>(n, s0 + chunk);
  double acc = 0.0;
  if (layout == kColMajor) {   // fast = i, slow = j
    const int64_t i = fast;
    for (int64_t j = s0; j < s1; ++j) {
      double v = fabs(a[i + j * lda]);
      if (rscale) {
        v *= rscale[i];
        if (v > cmax[j]) atomic_max_nonneg(&cmax[j], v);
      } else {
        acc = fmax(acc, v);
      }
    }
    if (!rscale) atomic_max_nonneg(&rmax[i], acc);
  } else {                     // fast = j, slow = i
    const int64_t j = fast;
    for (int64_t i = s0; i < s1; ++i) {
      double v = fabs(a[i * lda + j]);
      if (rscale) acc = fmax(acc, v * rscale[i]);
      else if (v > rmax[i]) atomic_max_nonneg(&rmax[i], v);
    }
    if (rscale) atomic_max_nonneg(&cmax[j], acc);
  }
}

__global__ void to_scale_kernel(int64_t n, const double* __restrict__ amax, double* __restrict__ scale) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) scale[i] = pow2_scale(amax[i]);
}

// W[i][j] = fl32(r_i a_ij c_j), W row-major (ld n); 32 x 32 tiles through LDS.
__global__ void equilibrate_kernel(int layout, int64_t n, const double* __restrict__ a, int64_t lda,
                                   const double* __restrict__ r, const double* __restrict__ c,
                                   float* __restrict__ W) {
  __shared__ float tile[32][33];
  const int64_t I = blockIdx.y * 32, J = blockIdx.x * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
  if (layout == kRowMajor) {
    for (int k = ty; k < 32; k += 8) {
      const int64_t i = I + k, j = J + tx;
      if (i < n && j < n) W[i * n + j] = (float)(r[i] * a[i * lda + j] * c[j]);
    }
  } else {
    for (int k = ty; k < 32; k += 8) {   // read column j = J + k, rows I + tx (contiguous)
      const int64_t i = I + tx, j = J + k;
      if (i < n && j < n) tile[tx][k] = (float)(r[i] * a[i + j * lda] * c[j]);
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
      const int64_t i = I + k, j = J + tx;
      if (i < n && j < n) W[i * n + j] = tile[k][tx];
    }
  }
}

// ---- FP32 / 16-bit operand preparation ---------------------------------------------
__device__ __forceinline__ float pow2_scale_f(float amax) {
  if (!(amax > 0.0f) || !isfinite(amax)) return 1.0f;
  return ldexpf(1.0f, -ilogbf(amax));
}

template <typename H> __device__ __forceinline__ H to_h(float v);
template <> __device__ __forceinline__ _Float16 to_h<_Float16>(float v) { return (_Float16)v; }
template <> __device__ __forceinline__ __bf16 to_h<__bf16>(float v) { return (__bf16)v; }

// Y (rows x ldy) = X_ik * s_i, s_i = 2^-floor(log2 max_k |X_ik|); inv[i] = 1 / s_i.
// One wave per row (k <= 512).
template <typename H>
__global__ void rows_to_lowp_kernel(int64_t rows, int kb, const float* __restrict__ X, int64_t ldx,
                                    H* __restrict__ Y, int64_t ldy, float* __restrict__ inv) {
  const int64_t row = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  const float* x = X + row * ldx;
  float mx = 0.0f;
  for (int k = lane; k < kb; k += 64) mx = fmaxf(mx, fabsf(x[k]));
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  const float s = pow2_scale_f(mx);
  H* y = Y + row * ldy;
  for (int k = lane; k < (int)ldy; k += 64) y[k] = to_h<H>(k < kb ? x[k] * s : 0.0f);
  if (lane == 0) inv[row] = 1.0f / s;
}

// YT (cols x ldy) = transpose of U (kb x cols) scaled per column; inv[j] = 1 / s_j.
// A workgroup handles 64 columns: the kb x 64 tile goes through LDS.
template <typename H>
__global__ void cols_to_lowp_t_kernel(int64_t cols, int kb, const float* __restrict__ U, int64_t ldu,
                                      H* __restrict__ YT, int64_t ldy, float* __restrict__ inv) {
  __shared__ float tile[256][65];
  __shared__ float cmax[4][64];
  const int64_t j0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4
  float mx = 0.0f;
  for (int k = ty; k < kb; k += 4) {
    const int64_t j = j0 + tx;
    const float v = j < cols ? U[k * ldu + j] : 0.0f;
    tile[k][tx] = v;
    mx = fmaxf(mx, fabsf(v));
  }
  cmax[ty][tx] = mx;
  __syncthreads();
  if (ty == 0) {
    const float m = fmaxf(fmaxf(cmax[0][tx], cmax[1][tx]), fmaxf(cmax[2][tx], cmax[3][tx]));
    cmax[0][tx] = pow2_scale_f(m);
    if (j0 + tx < cols) inv[j0 + tx] = 1.0f / cmax[0][tx];
  }
  __syncthreads();
  // write row j of YT: consecutive threads take consecutive k
  for (int jj = ty; jj < 64; jj += 4) {
    const int64_t j = j0 + jj;
    if (j >= cols) break;
    const float s = cmax[0][jj];
    for (int k = tx; k < (int)ldy; k += 64) YT[j * ldy + k] = to_h<H>(k < kb ? tile[k][jj] * s : 0.0f);
  }
}

// ---- FP32 triangular kernels ----------------------------------------------------------
// B (kb x cols, row-major) <- L^-1 B for a unit (or non-unit) lower L (kb <= 64):
// one thread per column, the 64 values in registers, L in LDS.
__global__ void __launch_bounds__(256)
trsm_lower_f32_kernel(int kb, const float* __restrict__ L, int64_t ldl, int unit,
                      float* __restrict__ B, int64_t ldb, int64_t cols) {
  __shared__ float Ls[64][65];
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int r = e / 64, c = e % 64;
    Ls[r][c] = (r < kb && c < kb) ? L[r * ldl + c] : 0.0f;
  }
  __syncthreads();
  const int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (col >= cols) return;
  float x[64];
#pragma unroll
  for (int r = 0; r < 64; ++r) x[r] = r < kb ? B[r * ldb + col] : 0.0f;
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    if (k < kb) {
      if (!unit) x[k] = x[k] / Ls[k][k];
#pragma unroll
      for (int i = k + 1; i < 64; ++i) x[i] -= x[k] * Ls[i][k];
    }
  }
#pragma unroll
  for (int r = 0; r < 64; ++r)
    if (r < kb) B[r * ldb + col] = x[r];
}

__global__ void fill_identity_f32_kernel(int64_t n, float* __restrict__ a, int64_t ld) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n * n) a[(e / n) * ld + e % n] = (e / n == e % n) ? 1.0f : 0.0f;
}

// U^T of a kb x kb diagonal block (row-major, ld) into dst (kb x kb).
__global__ void transpose_f32_kernel(int64_t kb, const float* __restrict__ src, int64_t lds,
                                     float* __restrict__ dst, int lower_only) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= kb * kb) return;
  const int64_t i = e / kb, j = e % kb;   // dst[i][j] = src[j][i]
  dst[i * kb + j] = (lower_only && j > i) ? 0.0f : src[j * lds + i];
}

// ---- blocked triangular solve helpers (vectors FP32) ---------------------------------
// t[r] = y[r] - sum_{c in [c0, c1)} W[(b0 + r) ld + c] z[c]  for r < nb  (one wave per row)
__global__ void rowdot_sub_kernel(int nb, const float* __restrict__ W, int64_t ld, int64_t b0,
                                  int64_t c0, int64_t c1, const float* __restrict__ z,
                                  const float* __restrict__ y, float* __restrict__ t) {
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= nb) return;
  const float* w = W + (b0 + r) * ld;
  double acc = 0.0;
  for (int64_t c = c0 + lane; c < c1; c += 64) acc += (double)w[c] * (double)z[c];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) t[r] = (float)((double)y[r] - acc);
}

// out[r] = sum_c M[r][c] t[c] (trans: M[c][r]) for an nb x nb block.
__global__ void block_matvec_kernel(int nb, const float* __restrict__ M, int trans,
                                    const float* __restrict__ t, float* __restrict__ out) {
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= nb) return;
  double acc = 0.0;
  for (int c = lane; c < nb; c += 64) acc += (double)(trans ? M[c * nb + r] : M[r * nb + c]) * (double)t[c];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) out[r] = (float)acc;
}

// y[i] = fl32(scale_i * v[perm[i]])
__global__ void gather_scale_kernel(int64_t n, const double* __restrict__ v, const int32_t* __restrict__ perm,
                                    const double* __restrict__ scale, float* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)(scale[perm[i]] * v[perm[i]]);
}

// out[i] = c_i w_i (mode 0) or out[i] += c_i w_i (mode 1), FP64
__global__ void scale_out_kernel(int64_t n, const float* __restrict__ w, const double* __restrict__ c,
                                 double* __restrict__ out, int accumulate) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = c[i] * (double)w[i];
  out[i] = accumulate ? out[i] + v : v;
}

// ---- FP64 vector kernels (refinement / GMRES) --------------------------------------------
__global__ void copy_f64_kernel(int64_t n, const double* __restrict__ x, int64_t incx, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i * incx];
}
__global__ void store_f64_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ y, int64_t incy) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i * incy] = x[i];
}
__global__ void axpy_f64_kernel(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
__global__ void scal_copy_f64_kernel(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}
// partial[b] = sum over the block of x_i y_i (mode 0) or max |x_i| (mode 1)
__global__ void reduce_f64_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                  int mode, double* __restrict__ partial) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc = mode == 0 ? acc + x[i] * y[i] : fmax(acc, fabs(x[i]));
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(acc, off);
    acc = mode == 0 ? acc + o : fmax(acc, o);
  }
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sw[0];
    for (int w = 1; w < 4; ++w) t = mode == 0 ? t + sw[w] : fmax(t, sw[w]);
    partial[blockIdx.x] = t;
  }
}
// row sums of |A| (FP64, either layout) -> s[i]
__global__ void rowsum_abs_kernel(int layout, int64_t n, const double* __restrict__ a, int64_t lda,
                                  double* __restrict__ s) {
  if (layout == kColMajor) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int64_t j = 0; j < n; ++j) acc += fabs(a[i + j * lda]);
    s[i] = acc;
  } else {
    const int64_t i = blockIdx.x;
    double acc = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) acc += fabs(a[i * lda + j]);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double sw[4];
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) s[i] = sw[0] + sw[1] + sw[2] + sw[3];
  }
}

inline unsigned blocks(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

// ---- LU operations (row-major FP32 working matrix) ---------------------------------------
struct MixedCtx {
  int prec = kPrecF16;
  _Float16* l16 = nullptr;   // rows x NB
  _Float16* u16 = nullptr;   // cols x NB (transposed)
  float* rs = nullptr;       // rows
  float* cs = nullptr;       // cols
  float* inv_all = nullptr;  // L11^-1 of every block column, NB x NB each
  float* tmp = nullptr;      // inverse scratch
};
thread_local MixedCtx tl_mix;

void panel_gemm_f32(int64_t rows, int64_t cols, int64_t kb, const float* L, int64_t ldl,
                    const float* U, int64_t ldu, float* C, int64_t ldc, hipStream_t s) {
  mp_gemm_f32(rows, cols, kb, -1.0f, L, ldl, U, ldu, 1.0f, nullptr, nullptr, C, ldc, s);
}

void trailing_gemm(int64_t rows, int64_t cols, int64_t kb, const float* L, int64_t ldl,
                   const float* U, int64_t ldu, float* C, int64_t ldc, hipStream_t s) {
  const MixedCtx& mx = tl_mix;
  if (mx.prec == kPrecF32) {
    void* tok = profile_begin(kProfGemmLowp, s);
    mp_gemm_f32(rows, cols, kb, -1.0f, L, ldl, U, ldu, 1.0f, nullptr, nullptr, C, ldc, s);
    profile_end(tok, s, 2.0 * rows * cols * kb, 4.0 * (rows * kb + kb * cols) + 8.0 * rows * cols);
    return;
  }
  const int64_t kp = (kb + 7) / 8 * 8;
  if (mx.prec == kPrecF16) {
    hipLaunchKernelGGL((rows_to_lowp_kernel<_Float16>), dim3(blocks(rows, 4)), dim3(256), 0, s, rows,
                       (int)kb, L, ldl, mx.l16, kp, mx.rs);
    hipLaunchKernelGGL((cols_to_lowp_t_kernel<_Float16>), dim3(blocks(cols, 64)), dim3(256), 0, s, cols,
                       (int)kb, U, ldu, mx.u16, kp, mx.cs);
    ZH_LAUNCH_CHECK();
    void* tok = profile_begin(kProfGemmLowp, s);
    mp_gemm_f16(rows, cols, kb, -1.0f, mx.l16, kp, mx.u16, kp, 1.0f, mx.rs, mx.cs, C, ldc, s);
    profile_end(tok, s, 2.0 * rows * cols * kb, 2.0 * (rows * kp + kp * cols) + 8.0 * rows * cols);
  } else {
    __bf16* l = reinterpret_cast<__bf16*>(mx.l16);
    __bf16* u = reinterpret_cast<__bf16*>(mx.u16);
    hipLaunchKernelGGL((rows_to_lowp_kernel<__bf16>), dim3(blocks(rows, 4)), dim3(256), 0, s, rows,
                       (int)kb, L, ldl, l, kp, mx.rs);
    hipLaunchKernelGGL((cols_to_lowp_t_kernel<__bf16>), dim3(blocks(cols, 64)), dim3(256), 0, s, cols,
                       (int)kb, U, ldu, u, kp, mx.cs);
    ZH_LAUNCH_CHECK();
    void* tok = profile_begin(kProfGemmLowp, s);
    mp_gemm_bf16(rows, cols, kb, -1.0f, l, kp, u, kp, 1.0f, mx.rs, mx.cs, C, ldc, s);
    profile_end(tok, s, 2.0 * rows * cols * kb, 2.0 * (rows * kp + kp * cols) + 8.0 * rows * cols);
  }
}

void trsm_f32(int64_t kb, int64_t cols, const float* L11, int64_t ld, float* B, int64_t ldb, hipStream_t s) {
  for (int64_t k0 = 0; k0 < kb; k0 += 64) {
    const int64_t b = std::min<int64_t>(64, kb - k0);
    hipLaunchKernelGGL(trsm_lower_f32_kernel, dim3(blocks(cols)), dim3(256), 0, s, (int)b,
                       L11 + k0 * ld + k0, ld, 1, B + k0 * ldb, ldb, cols);
    ZH_LAUNCH_CHECK();
    if (k0 + b < kb)
      mp_gemm_f32(kb - k0 - b, cols, b, -1.0f, L11 + (k0 + b) * ld + k0, ld, B + k0 * ldb, ldb, 1.0f,
                  nullptr, nullptr, B + (k0 + b) * ldb, ldb, s);
  }
}

// inv <- L^-1 (kb x kb, lower; unit or not) by forward substitution on the identity
void tri_inverse_f32(int64_t kb, const float* L, int64_t ld, float* inv, int unit, hipStream_t s) {
  hipLaunchKernelGGL(fill_identity_f32_kernel, dim3(blocks(kb * kb)), dim3(256), 0, s, kb, inv, kb);
  ZH_LAUNCH_CHECK();
  for (int64_t k0 = 0; k0 < kb; k0 += 64) {
    const int64_t b = std::min<int64_t>(64, kb - k0);
    // only columns < k0 + b of rows k0.. are non-zero so far: solve those
    hipLaunchKernelGGL(trsm_lower_f32_kernel, dim3(blocks(k0 + b)), dim3(256), 0, s, (int)b,
                       L + k0 * ld + k0, ld, unit, inv + k0 * kb, kb, k0 + b);
    ZH_LAUNCH_CHECK();
    if (k0 + b < kb)
      mp_gemm_f32(kb - k0 - b, k0 + b, b, -1.0f, L + (k0 + b) * ld + k0, ld, inv + k0 * kb, kb, 1.0f,
                  nullptr, nullptr, inv + (k0 + b) * kb, kb, s);
  }
}

void invert_f32(int64_t kb, const float* L11, int64_t ld, float* inv, hipStream_t s) {
  tri_inverse_f32(kb, L11, ld, inv, 1, s);
}

void apply_inv_f32(int64_t kb, int64_t cols, const float* inv, float* B, int64_t ldb, float* tmp,
                   hipStream_t s) {
  ZH_CHECK(hipMemcpy2DAsync(tmp, sizeof(float) * cols, B, sizeof(float) * ldb, sizeof(float) * cols, kb,
                            hipMemcpyDeviceToDevice, s));
  mp_gemm_f32(kb, cols, kb, 1.0f, inv, kb, tmp, cols, 0.0f, nullptr, nullptr, B, ldb, s);
}

// ---- solver state -------------------------------------------------------------------------------
struct Solver {
  int layout;
  int64_t n;
  const double* A;
  int64_t lda;
  hipStream_t s;
  float* W = nullptr;          // n x n FP32 factors (row-major)
  float* invL = nullptr;       // nblk x NB x NB
  float* invU = nullptr;       // nblk x NB x NB (inverse of U_bb^T)
  int32_t* ipiv = nullptr;
  int32_t* perm = nullptr;
  int* info = nullptr;
  double* r = nullptr;         // row scales
  double* c = nullptr;         // column scales
  double* rowsum = nullptr;
  float* y = nullptr;          // solve work (n each)
  float* z = nullptr;
  float* t = nullptr;
  float* t2 = nullptr;
  double* part = nullptr;      // reduction partials
  double anorm = 0.0;
  int64_t nblk = 0;
};

double reduce_host(const Solver& S, int64_t n, const double* x, const double* y, int mode) {
  const int nb = 256;
  hipLaunchKernelGGL(reduce_f64_kernel, dim3(nb), dim3(256), 0, S.s, n, x, y, mode, S.part);
  ZH_LAUNCH_CHECK();
  std::vector<double> h(nb);
  ZH_CHECK(hipMemcpyAsync(h.data(), S.part, sizeof(double) * nb, hipMemcpyDeviceToHost, S.s));
  ZH_CHECK(hipStreamSynchronize(S.s));
  double acc = 0.0;
  for (double v : h) acc = mode == 0 ? acc + v : std::max(acc, v);
  return acc;
}

// w (FP64) = Dc U^-1 L^-1 P Dr v; accumulate: w += ... instead
void lu_solve(const Solver& S, const double* v, double* w, int accumulate) {
  const int64_t n = S.n;
  hipLaunchKernelGGL(gather_scale_kernel, dim3(blocks(n)), dim3(256), 0, S.s, n, v, S.perm, S.r, S.y);
  ZH_LAUNCH_CHECK();
  // forward: z_b = invL_b (y_b - L[b, :b] z[:b])
  for (int64_t b = 0; b < S.nblk; ++b) {
    const int64_t b0 = b * NB_MIX, nb = std::min<int64_t>(NB_MIX, n - b0);
    hipLaunchKernelGGL(rowdot_sub_kernel, dim3(blocks(nb, 4)), dim3(256), 0, S.s, (int)nb, S.W, n, b0,
                       (int64_t)0, b0, S.z, S.y + b0, S.t);
    hipLaunchKernelGGL(block_matvec_kernel, dim3(blocks(nb, 4)), dim3(256), 0, S.s, (int)nb,
                       S.invL + b * NB_MIX * NB_MIX, 0, S.t, S.z + b0);
  }
  // backward: y_b = invU_b (z_b - U[b, b+1:] y[b+1:])
  for (int64_t b = S.nblk - 1; b >= 0; --b) {
    const int64_t b0 = b * NB_MIX, nb = std::min<int64_t>(NB_MIX, n - b0);
    hipLaunchKernelGGL(rowdot_sub_kernel, dim3(blocks(nb, 4)), dim3(256), 0, S.s, (int)nb, S.W, n, b0,
                       b0 + nb, n, S.y, S.z + b0, S.t);
    hipLaunchKernelGGL(block_matvec_kernel, dim3(blocks(nb, 4)), dim3(256), 0, S.s, (int)nb,
                       S.invU + b * NB_MIX * NB_MIX, 1, S.t, S.y + b0);
  }
  ZH_LAUNCH_CHECK();
  hipLaunchKernelGGL(scale_out_kernel, dim3(blocks(n)), dim3(256), 0, S.s, n, S.y, S.c, w, accumulate);
  ZH_LAUNCH_CHECK();
}

// r = b - A x (FP64, the caller's A)
void residual(const Solver& S, const double* b, const double* x, double* r) {
  ZH_CHECK(hipMemcpyAsync(r, b, sizeof(double) * S.n, hipMemcpyDeviceToDevice, S.s));
  const int trans = S.layout == kColMajor ? kNoTrans : kTrans;
  gemv_f64_cm(trans, S.n, S.n, -1.0, S.A, S.lda, x, 1, 1.0, r, 1, S.s);
}

double backward_error(const Solver& S, const double* r, const double* x) {
  const double rn = reduce_host(S, S.n, r, nullptr, 1);
  const double xn = reduce_host(S, S.n, x, nullptr, 1);
  return rn / (S.anorm * xn * (double)S.n * 0x1p-53);
}

// FGMRES(m) in FP64 with the LU solve as right preconditioner; updates x.
// Returns inner iterations used; *be gets the final backward error.
int gmres_ir(const Solver& S, const double* b, double* x, int restart, int max_it, double tol,
             double* be_out, double* work) {
  const int64_t n = S.n;
  const int m = restart;
  double* V = work;                    // (m + 1) x n
  double* Z = V + (size_t)(m + 1) * n; // m x n
  double* w = Z + (size_t)m * n;       // n
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1);
  int total = 0;
  double be = 0.0;
  for (int outer = 0; total < max_it; ++outer) {
    residual(S, b, x, w);
    be = backward_error(S, w, x);
    if (be <= tol) break;
    const double beta = std::sqrt(reduce_host(S, n, w, w, 0));
    if (!(beta > 0.0)) break;
    // ||r||_inf <= ||r||_2 = |g_{j+1}|: stop the cycle once that meets the target
    const double target = tol * S.anorm * reduce_host(S, n, x, nullptr, 1) * (double)n * 0x1p-53;
    hipLaunchKernelGGL(scal_copy_f64_kernel, dim3(blocks(n)), dim3(256), 0, S.s, n, 1.0 / beta, w, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m && total < max_it; ++j, ++total) {
      double* vj = V + (size_t)j * n;
      double* zj = Z + (size_t)j * n;
      lu_solve(S, vj, zj, 0);
      // w = A z_j
      const int trans = S.layout == kColMajor ? kNoTrans : kTrans;
      gemv_f64_cm(trans, n, n, 1.0, S.A, S.lda, zj, 1, 0.0, w, 1, S.s);
      for (int i = 0; i <= j; ++i) {   // modified Gram-Schmidt
        const double* vi = V + (size_t)i * n;
        const double h = reduce_host(S, n, w, vi, 0);
        H[i * m + j] = h;
        hipLaunchKernelGGL(axpy_f64_kernel, dim3(blocks(n)), dim3(256), 0, S.s, n, -h, vi, w);
      }
      const double hn = std::sqrt(reduce_host(S, n, w, w, 0));
      H[(j + 1) * m + j] = hn;
      if (hn > 0.0)
        hipLaunchKernelGGL(scal_copy_f64_kernel, dim3(blocks(n)), dim3(256), 0, S.s, n, 1.0 / hn, w,
                           V + (size_t)(j + 1) * n);
      for (int i = 0; i < j; ++i) {   // apply earlier rotations
        const double a = H[i * m + j], c2 = H[(i + 1) * m + j];
        H[i * m + j] = cs[i] * a + sn[i] * c2;
        H[(i + 1) * m + j] = -sn[i] * a + cs[i] * c2;
      }
      const double a = H[j * m + j], c2 = H[(j + 1) * m + j];
      const double rr = std::hypot(a, c2);
      cs[j] = rr > 0 ? a / rr : 1.0;
      sn[j] = rr > 0 ? c2 / rr : 0.0;
      H[j * m + j] = rr;
      H[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      if (std::fabs(g[j + 1]) <= target || hn == 0.0) { ++j; ++total; break; }
    }
    // solve H y = g (upper triangular j x j), x += Z y
    std::vector<double> yv(j);
    for (int i = j - 1; i >= 0; --i) {
      double acc = g[i];
      for (int k = i + 1; k < j; ++k) acc -= H[i * m + k] * yv[k];
      yv[i] = H[i * m + i] != 0.0 ? acc / H[i * m + i] : 0.0;
    }
    for (int i = 0; i < j; ++i)
      hipLaunchKernelGGL(axpy_f64_kernel, dim3(blocks(n)), dim3(256), 0, S.s, n, yv[i],
                         Z + (size_t)i * n, x);
    ZH_LAUNCH_CHECK();
  }
  residual(S, b, x, w);
  be = backward_error(S, w, x);
  *be_out = be;
  return total;
}

}  // namespace

// Factor, solve and refine; see the file comment. x (FP64, ld ldx) receives the
// solution of every right-hand side of b.
int dgesv_mixed(int layout, int64_t n, int64_t nrhs, const double* A, int64_t lda, const double* B,
                int64_t ldb, double* X, int64_t ldx, const zmlhip_ir_opts* o, zmlhip_ir_result* res,
                hipStream_t s) {
  zmlhip_ir_opts opts;
  if (o) opts = *o;
  else {
    std::memset(&opts, 0, sizeof(opts));
    opts.precision = kPrecF16;
  }
  if (opts.max_iter <= 0) opts.max_iter = 30;
  if (opts.gmres_restart <= 0) opts.gmres_restart = 30;
  if (opts.gmres_max_iter <= 0) opts.gmres_max_iter = 300;
  if (!(opts.tol > 0.0)) opts.tol = 1.0;
  std::memset(res, 0, sizeof(*res));
  if (n == 0 || nrhs == 0) return 0;

  Solver S;
  S.layout = layout;
  S.n = n;
  S.A = A;
  S.lda = lda;
  S.s = s;
  S.nblk = (n + NB_MIX - 1) / NB_MIX;
  // ---- workspaces (slots 8..): the FP32 factors dominate (4 n^2 bytes)
  S.W = static_cast<float*>(workspace(sizeof(float) * (size_t)n * n, 8));
  char* small = static_cast<char*>(workspace(
      sizeof(float) * 2 * (size_t)S.nblk * NB_MIX * NB_MIX + sizeof(int32_t) * 2 * n + 64 +
          sizeof(double) * 3 * n + sizeof(float) * 4 * (n + NB_MIX) + sizeof(double) * 512 + 4096, 9));
  S.invL = reinterpret_cast<float*>(small); small += sizeof(float) * S.nblk * NB_MIX * NB_MIX;
  S.invU = reinterpret_cast<float*>(small); small += sizeof(float) * S.nblk * NB_MIX * NB_MIX;
  S.ipiv = reinterpret_cast<int32_t*>(small); small += sizeof(int32_t) * n;
  S.perm = reinterpret_cast<int32_t*>(small); small += sizeof(int32_t) * n;
  S.info = reinterpret_cast<int*>(small); small += 64;
  S.r = reinterpret_cast<double*>(small); small += sizeof(double) * n;
  S.c = reinterpret_cast<double*>(small); small += sizeof(double) * n;
  S.rowsum = reinterpret_cast<double*>(small); small += sizeof(double) * n;
  S.y = reinterpret_cast<float*>(small); small += sizeof(float) * (n + NB_MIX);
  S.z = reinterpret_cast<float*>(small); small += sizeof(float) * (n + NB_MIX);
  S.t = reinterpret_cast<float*>(small); small += sizeof(float) * (n + NB_MIX);
  S.t2 = reinterpret_cast<float*>(small); small += sizeof(float) * (n + NB_MIX);
  S.part = reinterpret_cast<double*>(small);
  const int64_t kp = NB_MIX;
  char* lp = static_cast<char*>(workspace(sizeof(_Float16) * 2 * (size_t)(n + NB_MIX) * kp +
                                              sizeof(float) * 2 * (n + NB_MIX) + 4096, 10));
  tl_mix.prec = opts.precision;
  tl_mix.l16 = reinterpret_cast<_Float16*>(lp); lp += sizeof(_Float16) * (n + NB_MIX) * kp;
  tl_mix.u16 = reinterpret_cast<_Float16*>(lp); lp += sizeof(_Float16) * (n + NB_MIX) * kp;
  tl_mix.rs = reinterpret_cast<float*>(lp); lp += sizeof(float) * (n + NB_MIX);
  tl_mix.cs = reinterpret_cast<float*>(lp);

  // ---- 1. equilibration and the FP32 working matrix
  ZH_CHECK(hipMemsetAsync(S.r, 0, sizeof(double) * 2 * n, s));   // r, c hold the maxima first
  double* rmax = S.r;
  double* cmax = S.c;
  const int64_t split = 64, chunk = (n + split - 1) / split;
  hipLaunchKernelGGL(absmax_kernel, dim3(blocks(n), (unsigned)split), dim3(256), 0, s, layout, n, A, lda,
                     (const double*)nullptr, rmax, cmax, chunk);
  hipLaunchKernelGGL(to_scale_kernel, dim3(blocks(n)), dim3(256), 0, s, n, rmax, S.rowsum);   // rowsum <- r
  hipLaunchKernelGGL(absmax_kernel, dim3(blocks(n), (unsigned)split), dim3(256), 0, s, layout, n, A, lda,
                     (const double*)S.rowsum, rmax, cmax, chunk);
  ZH_CHECK(hipMemcpyAsync(S.r, S.rowsum, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(to_scale_kernel, dim3(blocks(n)), dim3(256), 0, s, n, cmax, S.c);
  ZH_LAUNCH_CHECK();
  hipLaunchKernelGGL(equilibrate_kernel, dim3(blocks(n, 32), blocks(n, 32)), dim3(32, 8), 0, s, layout, n,
                     A, lda, S.r, S.c, S.W);
  ZH_LAUNCH_CHECK();
  // ||A||_inf of the original matrix
  hipLaunchKernelGGL(rowsum_abs_kernel, dim3(layout == kColMajor ? blocks(n) : (unsigned)n), dim3(256), 0, s,
                     layout, n, A, lda, S.rowsum);
  ZH_LAUNCH_CHECK();
  S.anorm = reduce_host(S, n, S.rowsum, nullptr, 1);

  // ---- 2. factorization
  ZH_CHECK(hipMemsetAsync(S.info, 0, sizeof(int), s));
  struct Inv {
    static void store(int64_t kb, const float* L11, int64_t ld, float* inv, hipStream_t st) {
      invert_f32(kb, L11, ld, inv, st);
    }
  };
  const LuOps<float> ops{trailing_gemm, trsm_f32, Inv::store, apply_inv_f32, panel_gemm_f32};
  getrf_lookahead<float>(n, n, S.W, n, S.ipiv, S.info, NB_MIX, ops, s);
  int info = 0;
  ZH_CHECK(hipMemcpyAsync(&info, S.info, sizeof(int), hipMemcpyDeviceToHost, s));
  ZH_CHECK(hipStreamSynchronize(s));
  res->info = info;

  // inverses of the diagonal blocks of L and of U^T (for the triangular solves)
  float* scratch = static_cast<float*>(workspace(sizeof(float) * NB_MIX * NB_MIX, 11));
  for (int64_t b = 0; b < S.nblk; ++b) {
    const int64_t b0 = b * NB_MIX, nb = std::min<int64_t>(NB_MIX, n - b0);
    tri_inverse_f32(nb, S.W + b0 * n + b0, n, S.invL + b * NB_MIX * NB_MIX, 1, s);
    hipLaunchKernelGGL(transpose_f32_kernel, dim3(blocks(nb * nb)), dim3(256), 0, s, nb, S.W + b0 * n + b0, n,
                       scratch, 1);
    tri_inverse_f32(nb, scratch, nb, S.invU + b * NB_MIX * NB_MIX, 0, s);
  }
  ZH_LAUNCH_CHECK();
  // row permutation: perm[i] = original row of factored row i
  {
    std::vector<int32_t> ip(n), pm(n);
    ZH_CHECK(hipMemcpyAsync(ip.data(), S.ipiv, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i) pm[i] = (int32_t)i;
    for (int64_t i = 0; i < n; ++i) std::swap(pm[i], pm[ip[i] - 1]);
    ZH_CHECK(hipMemcpyAsync(S.perm, pm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  }

  // ---- 3./4. per right-hand side: solve and refine
  double* vec = static_cast<double*>(workspace(sizeof(double) * 4 * (size_t)n, 12));
  double* bj = vec;
  double* xj = vec + n;
  double* rj = vec + 2 * n;
  double* dj = vec + 3 * n;
  double* gw = nullptr;
  double worst = 0.0;
  for (int64_t col = 0; col < nrhs; ++col) {
    const int64_t bstride = layout == kColMajor ? 1 : ldb, xstride = layout == kColMajor ? 1 : ldx;
    const double* bcol = layout == kColMajor ? B + col * ldb : B + col;
    double* xcol = layout == kColMajor ? X + col * ldx : X + col;
    hipLaunchKernelGGL(copy_f64_kernel, dim3(blocks(n)), dim3(256), 0, s, n, bcol, bstride, bj);
    ZH_LAUNCH_CHECK();
    lu_solve(S, bj, xj, 0);
    double prev = INFINITY, be = INFINITY;
    int it = 0, mode = 0;   // 0 classic IR, 1 GMRES-IR, 2 FP64
    for (; it <= opts.max_iter; ++it) {
      residual(S, bj, xj, rj);
      be = backward_error(S, rj, xj);
      if (be <= opts.tol) break;
      const double rn = be;
      if (!(rn == rn) || rn > 2.0 * prev) {   // diverging: GMRES-IR from the plain LU solution
        lu_solve(S, bj, xj, 0);
        mode = 1;
        break;
      }
      if (it >= 2 && rn > 0.5 * prev) { mode = 1; break; }       // stagnating
      prev = rn;
      if (it == opts.max_iter) break;
      lu_solve(S, rj, xj, 1);
    }
    res->iterations = std::max(res->iterations, it);
    if (mode == 0 && !(be <= opts.tol)) mode = 1;
    if (mode == 1 && !opts.no_gmres) {
      if (!gw) gw = static_cast<double*>(workspace(sizeof(double) * (size_t)(2 * opts.gmres_restart + 3) * n, 13));
      double be2 = 0.0;
      const int gi = gmres_ir(S, bj, xj, opts.gmres_restart, opts.gmres_max_iter, opts.tol, &be2, gw);
      res->gmres_iterations = std::max(res->gmres_iterations, gi);
      res->fallback = std::max(res->fallback, 1);
      be = be2;
      if (!(be <= opts.tol)) mode = 2;
    }
    if (mode == 2 || !(be <= opts.tol)) {
      // FP64 factorization of a copy of A
      res->fallback = 2;
      double* A64 = static_cast<double*>(workspace(sizeof(double) * (size_t)n * n, 14));
      int32_t* ip64 = reinterpret_cast<int32_t*>(dj);
      int* inf64 = reinterpret_cast<int*>(rj);
      ZH_CHECK(hipMemcpy2DAsync(A64, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                                hipMemcpyDeviceToDevice, s));
      ZH_CHECK(hipMemsetAsync(inf64, 0, sizeof(int), s));
      // A64 read row-major is A (row-major input) or A^T (column-major input)
      getrf_f64_rm(n, n, A64, n, ip64, inf64, s);
      ZH_CHECK(hipMemcpyAsync(xj, bj, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      // x is an n x 1 row-major matrix (ld 1); a row-major (Left, uplo) solve is the
      // column-major (Right, inverted uplo) one on the same memory (trsm.zig:48-62)
      if (layout == kRowMajor) {   // P A = L U: x = U^-1 L^-1 P b
        laswp_rm(1, xj, 1, 0, (int)n - 1, ip64, 1, s);
        trsm_f64_cm(kRight, kUpper, kNoTrans, kUnit, 1, n, 1.0, A64, n, xj, 1, s);
        trsm_f64_cm(kRight, kLower, kNoTrans, kNonUnit, 1, n, 1.0, A64, n, xj, 1, s);
      } else {                     // P A^T = L U: A = U^T L^T P, x = P^T L^-T U^-T b
        trsm_f64_cm(kRight, kLower, kTrans, kNonUnit, 1, n, 1.0, A64, n, xj, 1, s);
        trsm_f64_cm(kRight, kUpper, kTrans, kUnit, 1, n, 1.0, A64, n, xj, 1, s);
        laswp_rm(1, xj, 1, (int)n - 1, 0, ip64, -1, s);
      }
      residual(S, bj, xj, rj);
      be = backward_error(S, rj, xj);
    }
    worst = std::max(worst, be);
    hipLaunchKernelGGL(store_f64_kernel, dim3(blocks(n)), dim3(256), 0, s, n, xj, xcol, xstride);
    ZH_LAUNCH_CHECK();
  }
  ZH_CHECK(hipStreamSynchronize(s));
  res->backward_error = worst;
  return info;
}

}  // namespace zmlhip
