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

constexpr int NB_MIX = 512;   // block column of the mixed factorization (two 256-wide halves)

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

// A "line" is a row of a row-major matrix or a column of a column-major one
// (its elements are contiguous); a "position" indexes along a line.
// out[line] = max_k |a[line ld + k]| * (scale ? scale[k] : 1); one wave per line.
// (sum: also the line's sum of |a|, in the same pass)
__global__ void line_absmax_kernel(int64_t n, const double* __restrict__ a, int64_t lda,
                                   const double* __restrict__ scale, double* __restrict__ out,
                                   double* __restrict__ sum = nullptr) {
  const int64_t line = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (line >= n) return;
  const double* p = a + line * lda;
  double acc = 0.0, tot = 0.0;
  for (int64_t k = lane; k < n; k += 64) {
    const double v = fabs(p[k]);
    acc = fmax(acc, scale ? v * scale[k] : v);
    tot += v;
  }
  for (int off = 32; off > 0; off >>= 1) {
    acc = fmax(acc, __shfl_xor(acc, off));
    tot += __shfl_xor(tot, off);
  }
  if (lane == 0) {
    out[line] = acc;
    if (sum) sum[line] = tot;
  }
}

// out[k] = max over lines of |a[line ld + k]| * (scale ? scale[line] : 1): one
// thread per position over a chunk of lines (blockIdx.y), one atomic per thread.
__global__ void cross_absmax_kernel(int64_t n, const double* __restrict__ a, int64_t lda,
                                    const double* __restrict__ scale, double* __restrict__ out,
                                    int64_t chunk) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t l0 = blockIdx.y * chunk, l1 = std::min<int64_t>(n, l0 + chunk);
  double acc = 0.0;
  for (int64_t l = l0; l < l1; ++l) acc = fmax(acc, scale ? fabs(a[l * lda + k]) * scale[l] : fabs(a[l * lda + k]));
  atomic_max_nonneg(&out[k], acc);
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

// YT (cols x ldy) = transpose of U (kb x cols, kb <= 512) scaled per column;
// inv[j] = 1 / s_j. A workgroup handles 32 columns: the kb x 32 tile goes
// through LDS.
template <typename H>
__global__ void __launch_bounds__(256) cols_to_lowp_t_kernel(int64_t cols, int kb, const float* __restrict__ U,
                                                             int64_t ldu, H* __restrict__ YT, int64_t ldy,
                                                             float* __restrict__ inv) {
  __shared__ float tile[512][33];
  __shared__ float cmax[8][32];
  const int64_t j0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  float mx = 0.0f;
  // 16 rows per thread in flight: clamped (in-bounds) loads issued together,
  // then masked
  const int64_t j = j0 + tx, jc = j < cols ? j : cols - 1;
  for (int k0 = ty; k0 < kb; k0 += 8 * 16) {
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = k0 + 8 * q, kc = k < kb ? k : kb - 1;
      v[q] = U[kc * ldu + jc];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      asm volatile("" : "+v"(v[q]));
      const int k = k0 + 8 * q;
      if (k < kb) {
        const float x = j < cols ? v[q] : 0.0f;
        tile[k][tx] = x;
        mx = fmaxf(mx, fabsf(x));
      }
    }
  }
  cmax[ty][tx] = mx;
  __syncthreads();
  if (ty == 0) {
    float m = cmax[0][tx];
#pragma unroll
    for (int r = 1; r < 8; ++r) m = fmaxf(m, cmax[r][tx]);
    cmax[0][tx] = pow2_scale_f(m);
    if (j0 + tx < cols) inv[j0 + tx] = 1.0f / cmax[0][tx];
  }
  __syncthreads();
  // row j of YT: consecutive threads take consecutive k
  for (int jj = ty; jj < 32; jj += 8) {
    const int64_t j = j0 + jj;
    if (j >= cols) break;
    const float sc = cmax[0][jj];
    for (int k = tx; k < (int)ldy; k += 32) YT[j * ldy + k] = to_h<H>(k < kb ? tile[k][jj] * sc : 0.0f);
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
  for (int r = 0; r < 64; ++r) {   // rows beyond kb: a valid row is read and masked (no branch per load)
    const float v = B[(r < kb ? r : kb - 1) * ldb + col];
    x[r] = r < kb ? v : 0.0f;
  }
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

// ---- optimistic panel: pivots confined to the diagonal block ----------------------------
// Partial-pivoting LU of a 256 x 256 FP32 block held in the registers of one
// workgroup: thread (ty, tx) of a 32 x 32 grid holds rows ty + 32 i and
// columns tx + 32 j (i, j < 8). Row interchanges are virtual: a chosen pivot
// row is marked done and the elimination simply skips done rows; the block is
// written back in pivot order at the end. ipiv[k] receives the
// block row that took position k (rowmap_to_ipiv_kernel turns that into the
// LAPACK form later, off the panel's critical path), and the same order goes
// out as a list of row moves (destination, source; global rows = block row +
// `base`), the form apply_perm_kernel reads: into `pairs`, and into the block
// column's list `bpairs` when given, after the moves already there if
// `bappend` (the block's rows follow those of the earlier blocks). Pivot =
// largest |a| among the rows not yet chosen, ties to the lowest row of the
// block. l = a * fl(1/p) when |p| >= FLT_MIN, else a / p (the rule of zml's
// getrf2.zig:65-87). A zero or NaN pivot sets *fail (which the block column's
// first block, bappend == 0, overwrites: 1 or 0).
constexpr int DB = 256;
constexpr int DIAG_TY = 32;   // thread rows of the diagonal-block LU

__device__ __forceinline__ int key_row(uint64_t k) {
  const uint32_t lo = (uint32_t)k;
  return lo >= 0x80000000u ? (int)(0xffffffffu - lo) : (int)(0x7fffffffu - lo);
}

// The largest int of the wave, in every lane: quad permutations, row rotations
// by 4 and 8, then the row broadcasts 15 and 31 (lane 63 ends with the
// maximum), read from lane 63.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_max_step(int f) {
  // lanes outside ROW_MASK see INT_MIN and keep their own value
  const int o = __builtin_amdgcn_update_dpp((int)0x80000000, f, CTRL, ROW_MASK, 0xf, false);
  return o > f ? o : f;
}
__device__ __forceinline__ int wave64_max(int f) {
  f = dpp_max_step<0xb1>(f);          // quad_perm [1,0,3,2]
  f = dpp_max_step<0x4e>(f);          // quad_perm [2,3,0,1]
  f = dpp_max_step<0x124>(f);         // row_ror:4
  f = dpp_max_step<0x128>(f);         // row_ror:8
  f = dpp_max_step<0x142, 0xa>(f);    // row_bcast:15 into rows 1 and 3
  f = dpp_max_step<0x143, 0xc>(f);    // row_bcast:31 into rows 2 and 3
  return __builtin_amdgcn_readlane(f, 63);
}

// The block is eliminated in chunks of 16 columns (half a column group). One
// wave, four block rows per lane, runs a chunk's 16 pivot steps in its
// registers: each row's rank, the wave maximum through DPP, the pivot row
// passed from the winning lane through LDS, the multipliers and the update of
// the chunk's later columns, with no barrier between steps (the wave's own
// 64 entries wait in LDS meanwhile). Then the chunk is applied to the later columns, step by
// step: the pivot rows' values there (U12) come from a forward substitution
// with the chunk's unit lower factor (one column per thread), and row i
// receives a[i][c] -= l_ik u_k[c] for k in step order while it has not been
// chosen (multiplier 0 afterwards). Every entry sees the same operations in
// the same order as in a column-by-column elimination, so the bits are the
// same.
template <int TY>
__global__ void __launch_bounds__(TY * 32) diag_getf2_kernel(float* __restrict__ P, int64_t ld,
                                                            int32_t* __restrict__ ipiv, int base,
                                                            int* __restrict__ fail, int32_t* __restrict__ pairs,
                                                            int32_t* __restrict__ bpairs, int bappend) {
  constexpr int TX = 32, RI = DB / TY, RJ = DB / TX, NT = TY * TX;
  constexpr int CW = 16;        // chunk width: half a column group
  constexpr int CS = CW + 4;    // chunk row stride in floats (rows stay 16-byte aligned)
  constexpr int NC = DB / CW;   // chunks
  constexpr int RL = DB / 64;   // block rows per lane of the panel wave: lane + 64 m
  static_assert(TY == 32 && NT == 1024 && TX == 2 * CW, "32 x 32 threads, 8 x 8 entries each");
  __shared__ __align__(16) float chunk[DB * CS];
  __shared__ __align__(16) float lz[DB * CS];   // the multiplier each row took at each step (0 once chosen)
  __shared__ __align__(16) float ubuf[CW * DB];   // row s: the chunk's pivot s on the later columns
  __shared__ float park[64 * RI * RJ];   // the panel wave's own entries during the steps
  __shared__ __align__(16) float prbuf[RL * CW];   // the winning lane's rows at a step
  __shared__ int prow[DB], rstep[DB], invp[DB];   // rstep[r]: the step that chose row r, or -1
  __shared__ int nmove[DB / 64], bfirst;
  __shared__ int bad;
  const int t = threadIdx.x, tx = t % TX, ty = t / TX, lane = t & 63, wv = t >> 6;
  float a[RI][RJ];
#pragma unroll
  for (int i = 0; i < RI; ++i)
#pragma unroll
    for (int j = 0; j < RJ; ++j) a[i][j] = P[(int64_t)(ty + TY * i) * ld + tx + TX * j];
  if (t < DB) rstep[t] = -1;
  if (t == 0) bad = 0;
#pragma unroll 1
  for (int ch = 0; ch < NC; ++ch) {
    const int k0 = CW * ch, jg = ch >> 1, h = ch & 1;
    const bool mine = (tx / CW) == h;   // this thread holds chunk columns (register group jg)
    const int cx = tx % CW;
    if (mine) {
#pragma unroll
      for (int j = 0; j < RJ; ++j)
        if (j == jg) {
#pragma unroll
          for (int i = 0; i < RI; ++i) chunk[(ty + TY * i) * CS + cx] = a[i][j];
        }
    }
    __syncthreads();
    if (wv == 0) {
      // ---- the chunk's pivot steps
#pragma unroll
      for (int i = 0; i < RI; ++i)
#pragma unroll
        for (int j = 0; j < RJ; ++j) park[(i * RJ + j) * 64 + lane] = a[i][j];
      float x[RL][CW];
      bool rd[RL];
#pragma unroll
      for (int m = 0; m < RL; ++m) {
        const int row = lane + 64 * m;
#pragma unroll
        for (int q = 0; q < CW / 4; ++q) {
          const float4 v = *reinterpret_cast<const float4*>(&chunk[row * CS + 4 * q]);
          x[m][4 * q] = v.x;
          x[m][4 * q + 1] = v.y;
          x[m][4 * q + 2] = v.z;
          x[m][4 * q + 3] = v.w;
        }
        rd[m] = rstep[row] >= 0;
      }
#pragma unroll
      for (int kq = 0; kq < CW; ++kq) {
        const int k = k0 + kq;
        // rank: the bits of |a| (ordered as |a|), then -1 for NaN (below
        // every number), -2 for a row already chosen
        int f[RL], fl = -2;
#pragma unroll
        for (int m = 0; m < RL; ++m) {
          const int vb = __float_as_int(x[m][kq]) & 0x7fffffff;
          f[m] = rd[m] ? -2 : (vb > 0x7f800000 ? -1 : vb);
          fl = f[m] > fl ? f[m] : fl;
        }
        const int fm = wave64_max(fl);
        // the lowest row of rank fm: the first m with a hit, its lowest lane
        // (some row is always left, so fm >= -1 and a hit exists)
        int pm = 0;
        uint64_t hit = 0;
#pragma unroll
        for (int m = RL - 1; m >= 0; --m) {
          const uint64_t bm = __ballot(f[m] == fm);
          if (bm) {
            hit = bm;
            pm = m;
          }
        }
        const int pl = hit ? (int)__builtin_ctzll(hit) : 0, p = pl + 64 * pm;
        // the pivot row from column kq on: lane pl puts its rows out, the wave
        // reads row pm back (the buffer is rewritten only after these reads)
        if (lane == pl) {
#pragma unroll
          for (int m = 0; m < RL; ++m)
#pragma unroll
            for (int q = 0; q < CW / 4; ++q)
              if (q >= kq / 4)
                *reinterpret_cast<float4*>(&prbuf[m * CW + 4 * q]) =
                    make_float4(x[m][4 * q], x[m][4 * q + 1], x[m][4 * q + 2], x[m][4 * q + 3]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float pr[CW];
#pragma unroll
        for (int q = 0; q < CW / 4; ++q) {
          if (q >= kq / 4) {
            const float4 v = *reinterpret_cast<const float4*>(&prbuf[pm * CW + 4 * q]);
            pr[4 * q] = v.x;
            pr[4 * q + 1] = v.y;
            pr[4 * q + 2] = v.z;
            pr[4 * q + 3] = v.w;
          }
        }
        const float pv = pr[kq];
        const bool ok = pv != 0.0f && pv == pv;
#pragma unroll
        for (int m = 0; m < RL; ++m) rd[m] = rd[m] || (m == pm && lane == pl);
        if (lane == 0) {
          prow[k] = p;
          rstep[p] = k;
          if (!ok) bad = 1;
        }
        // (recip is the same in every lane: the division runs only for a subnormal pivot)
        const bool recip = fabsf(pv) >= 0x1p-126f;
        const float rp = 1.0f / pv;
#pragma unroll
        for (int m = 0; m < RL; ++m) {
          float li;
          if (recip)
            li = (!rd[m] && ok) ? x[m][kq] * rp : 0.0f;
          else
            li = (!rd[m] && ok) ? x[m][kq] / pv : 0.0f;
          lz[(lane + 64 * m) * CS + kq] = li;
          if (!rd[m] && ok) x[m][kq] = li;
#pragma unroll
          for (int c = 0; c < CW; ++c)
            if (c > kq) x[m][c] -= li * pr[c];
        }
      }
#pragma unroll
      for (int m = 0; m < RL; ++m)
#pragma unroll
        for (int q = 0; q < CW / 4; ++q)
          *reinterpret_cast<float4*>(&chunk[(lane + 64 * m) * CS + 4 * q]) =
              make_float4(x[m][4 * q], x[m][4 * q + 1], x[m][4 * q + 2], x[m][4 * q + 3]);
#pragma unroll
      for (int i = 0; i < RI; ++i)
#pragma unroll
        for (int j = 0; j < RJ; ++j) a[i][j] = park[(i * RJ + j) * 64 + lane];
    }
    __syncthreads();
    // ---- the chunk applied to the later columns: the rest of group jg when
    // the chunk is its first half (tx >= CW), and the groups after jg
    if (mine) {
#pragma unroll
      for (int j = 0; j < RJ; ++j)
        if (j == jg) {
#pragma unroll
          for (int i = 0; i < RI; ++i) a[i][j] = chunk[(ty + TY * i) * CS + cx];
        }
    }
    const bool later_half = h == 0 && tx >= CW;
#pragma unroll
    for (int i = 0; i < RI; ++i) {
      const int s = rstep[ty + TY * i] - k0;
      if (s >= 0 && s < CW) {
#pragma unroll
        for (int j = 0; j < RJ; ++j)
          if (j > jg || (j == jg && later_half)) ubuf[s * DB + tx + TX * j] = a[i][j];
      }
    }
    __syncthreads();
    if (t < DB - k0 - CW) {
      // U12 column c: pivot row s minus the earlier pivots' multiples, in step order
      const int c = k0 + CW + t;
      float u[CW];
#pragma unroll
      for (int s = 0; s < CW; ++s) {
        const float* ls = &chunk[prow[k0 + s] * CS];
        float v = ubuf[s * DB + c];
#pragma unroll
        for (int q = 0; 4 * q < s; ++q) {
          const float4 l4 = *reinterpret_cast<const float4*>(ls + 4 * q);
          const float lv[4] = {l4.x, l4.y, l4.z, l4.w};
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (4 * q + e < s) v -= lv[e] * u[4 * q + e];
        }
        u[s] = v;
        ubuf[s * DB + c] = v;
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int q = 0; q < CW / 4; ++q) {
      float l4[RI][4];
#pragma unroll
      for (int i = 0; i < RI; ++i) {
        const float4 v = *reinterpret_cast<const float4*>(&lz[(ty + TY * i) * CS + 4 * q]);
        l4[i][0] = v.x;
        l4[i][1] = v.y;
        l4[i][2] = v.z;
        l4[i][3] = v.w;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
#pragma unroll
        for (int j = 0; j < RJ; ++j)
          if (j > jg || (j == jg && later_half)) {
            const float u = ubuf[(4 * q + e) * DB + tx + TX * j];
#pragma unroll
            for (int i = 0; i < RI; ++i) a[i][j] -= l4[i][e] * u;
          }
      }
    }
    __syncthreads();   // chunk, lz and ubuf are rewritten by the next chunk
  }
  // row r of the block goes to position invp[r]; position k moves in from
  // block row prow[k] when they differ (moves numbered in position order)
  const bool moved = t < DB && prow[t] != t;
  const uint64_t mb = __ballot(moved);
  if (t < DB) {
    invp[prow[t]] = t;
    ipiv[t] = prow[t];
    if (lane == 0) nmove[wv] = (int)__popcll(mb);
  }
  if (t == 0) {
    bfirst = bpairs && bappend ? bpairs[0] : 0;
    // the block column's first block starts the outcome word afresh
    if (bappend) {
      if (bad) atomicOr(fail, 1);
    } else {
      *fail = bad ? 1 : 0;
    }
  }
  __syncthreads();
  if (t < DB) {
    int idx = (int)__popcll(mb & ((1ull << lane) - 1)), total = 0;
#pragma unroll
    for (int v = 0; v < DB / 64; ++v) {
      idx += v < wv ? nmove[v] : 0;
      total += nmove[v];
    }
    if (moved) {
      pairs[1 + 2 * idx] = base + t;
      pairs[2 + 2 * idx] = base + prow[t];
      if (bpairs) {
        bpairs[1 + 2 * (bfirst + idx)] = base + t;
        bpairs[2 + 2 * (bfirst + idx)] = base + prow[t];
      }
    }
    if (t == 0) {
      pairs[0] = total;
      if (bpairs) bpairs[0] = bfirst + total;
    }
  }
#pragma unroll
  for (int i = 0; i < RI; ++i) {
    const int64_t row = invp[ty + TY * i];
#pragma unroll
    for (int j = 0; j < RJ; ++j) P[row * ld + tx + TX * j] = a[i][j];
  }
}

// ipiv[0..DB) holds a block's row order (position k took block row ipiv[k]);
// rewritten in place in the LAPACK form: the interchanges that, applied in
// order k = 0, 1, ..., produce that order (1-based, offset by `base`).
__global__ void __launch_bounds__(DB) rowmap_to_ipiv_kernel(int32_t* __restrict__ ipiv, int base) {
  __shared__ int prow[DB], at[DB], pos[DB];   // at: row at a position; pos: position of a row
  const int t = threadIdx.x;
  prow[t] = ipiv[t];
  at[t] = t;
  pos[t] = t;
  __syncthreads();
  if (t == 0) {
    for (int k = 0; k < DB; ++k) {
      const int p = prow[k], q = pos[p], r = at[k];
      ipiv[k] = base + q + 1;
      at[q] = r;
      pos[r] = q;
      at[k] = p;
      pos[p] = k;
    }
  }
}


// *fail |= 2 when some |x| of the rows x cols block is not <= tau (NaN included)
__global__ void bound_check_kernel(int64_t rows, int cols, const float* __restrict__ X, int64_t ld,
                                   float tau, int* __restrict__ fail) {
  bool over = false;
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float v = X[(e / cols) * ld + e % cols];
    over |= !(fabsf(v) <= tau);
  }
  if (__ballot(over) && (threadIdx.x & 63) == 0) atomicOr(fail, 2);
}

// ---- triangular solves with the FP32 factors -----------------------------------------
// One persistent launch per triangle. The factor is cut into TS x TS blocks;
// workgroup g owns block rows g, g + G, ... (forward: in increasing order,
// backward: decreasing). For block row b it accumulates
//   t_b = y_b - sum_j T_bj z_j      (j < b forward, j > b backward)
// taking each z_j as soon as its owner has published it, then
//   z_b = D_b t_b,  D_b = inverse of the diagonal block T_bb,
// and publishes z_b. Published values are 64-bit words (tag << 32 | float
// bits) written and read with device-scope relaxed atomics: a reader knows a
// value is current from its tag, so no cache-wide fence is needed. The grid
// never exceeds the number of co-resident workgroups, and every block row
// waits only on block rows earlier in its own order, so the waits always end;
// a spin limit turns a stall into an error flag instead of a hang.
// Layout of a workgroup: 4 waves x 32 rows; lane l covers columns l and l+64
// of each column block, so every row read is two coalesced 256-byte loads.
constexpr int TS = 128;
constexpr int TS_RPW = 32;

struct TriSolve {
  int64_t n, nblk;
  int upper;
  const float* W;            // row-major factor, ld ldw
  int64_t ldw;
  const float* inv;          // inverses of the NB_MIX diagonal blocks (row-major, ld = block size)
  const double* v;           // forward: y_i = fl32(rsc[perm[i]] v[perm[i]]) ...
  const int32_t* perm;
  const double* rsc;
  const float* yin;          // ... or y = yin when v is null
  float* zout;               // plain copy of the solution (optional)
  const double* csc;         // wout_i (+)= csc_i z_i (optional)
  double* wout;
  int accumulate;
  uint64_t* xch;             // n tagged words
  uint32_t tag;
  int* err;
  int64_t spin_limit;
};

__device__ __forceinline__ float wait_word(const uint64_t* p, uint32_t tag, int64_t limit, int* err) {
  for (int64_t it = 0;; ++it) {
    const uint64_t w = __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(w >> 32) == tag) return __uint_as_float((uint32_t)w);
    if (it >= limit) {
      atomicOr(err, 1);
      return 0.0f;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__global__ void __launch_bounds__(256) trisolve_f32_kernel(TriSolve a) {
  __shared__ float red[4][TS_RPW][65];
  __shared__ float tvec[TS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t bi = blockIdx.x; bi < a.nblk; bi += gridDim.x) {
    const int64_t b = a.upper ? a.nblk - 1 - bi : bi;
    const int64_t r0 = b * TS;
    const int rows = (int)std::min<int64_t>(TS, a.n - r0);
    const int wr0 = w * TS_RPW;   // first row of this wave within the block
    // D_b = the (o, o) sub-block of the inverse of its NB_MIX-block, fetched up
    // front: only the last z_j is on the critical path, D_b must not be
    const int64_t B = r0 / NB_MIX;
    const int64_t o = r0 - B * NB_MIX;
    const int64_t ldi = std::min<int64_t>(NB_MIX, a.n - B * NB_MIX);
    const float* D = a.inv + B * NB_MIX * NB_MIX + o * ldi + o;
    float d0[TS_RPW], d1[TS_RPW];
#pragma unroll
    for (int rr = 0; rr < TS_RPW; ++rr) {
      const bool rok = wr0 + rr < rows;
      const float* row = D + (wr0 + rr) * ldi;
      d0[rr] = rok && lane < rows ? row[lane] : 0.0f;
      d1[rr] = rok && lane + 64 < rows ? row[lane + 64] : 0.0f;
    }
    float acc[TS_RPW];
#pragma unroll
    for (int rr = 0; rr < TS_RPW; ++rr) acc[rr] = 0.0f;
    const int64_t cnt = a.upper ? a.nblk - 1 - b : b;
    for (int64_t q = 0; q < cnt; ++q) {
      const int64_t j = a.upper ? a.nblk - 1 - q : q;
      const int64_t c0 = j * TS;
      const int cols = (int)std::min<int64_t>(TS, a.n - c0);
      // the factor's values first (independent of z_j), then wait for z_j
      float l0[TS_RPW], l1[TS_RPW];
#pragma unroll
      for (int rr = 0; rr < TS_RPW; ++rr) {
        const bool rok = wr0 + rr < rows;
        const float* row = a.W + (r0 + wr0 + rr) * a.ldw + c0;
        l0[rr] = rok && lane < cols ? row[lane] : 0.0f;
        l1[rr] = rok && lane + 64 < cols ? row[lane + 64] : 0.0f;
      }
      const float z0 = lane < cols ? wait_word(a.xch + c0 + lane, a.tag, a.spin_limit, a.err) : 0.0f;
      const float z1 = lane + 64 < cols ? wait_word(a.xch + c0 + lane + 64, a.tag, a.spin_limit, a.err) : 0.0f;
#pragma unroll
      for (int rr = 0; rr < TS_RPW; ++rr) acc[rr] = fmaf(l1[rr], z1, fmaf(l0[rr], z0, acc[rr]));
    }
    // t = y - (row sums of the partials)
#pragma unroll
    for (int rr = 0; rr < TS_RPW; ++rr) red[w][rr][lane] = acc[rr];
    __syncthreads();
    if (lane < TS_RPW) {
      const int r = wr0 + lane;
      float s = 0.0f;
      for (int l = 0; l < 64; ++l) s += red[w][lane][l];
      float y = 0.0f;
      if (r < rows) {
        const int64_t i = r0 + r;
        if (a.v) {
          const int32_t p = a.perm[i];
          y = (float)(a.rsc[p] * a.v[p]);
        } else {
          y = a.yin[i];
        }
      }
      tvec[r] = r < rows ? y - s : 0.0f;
    }
    __syncthreads();
    // z_b = D_b t
    const float t0 = lane < rows ? tvec[lane] : 0.0f;
    const float t1 = lane + 64 < rows ? tvec[lane + 64] : 0.0f;
#pragma unroll
    for (int rr = 0; rr < TS_RPW; ++rr) acc[rr] = fmaf(d1[rr], t1, d0[rr] * t0);
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < TS_RPW; ++rr) red[w][rr][lane] = acc[rr];
    __syncthreads();
    if (lane < TS_RPW && wr0 + lane < rows) {
      float z = 0.0f;
      for (int l = 0; l < 64; ++l) z += red[w][lane][l];
      const int64_t i = r0 + wr0 + lane;
      __hip_atomic_store(a.xch + i, ((uint64_t)a.tag << 32) | __float_as_uint(z), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      if (a.zout) a.zout[i] = z;
      if (a.wout) {
        const double v = a.csc[i] * (double)z;
        a.wout[i] = a.accumulate ? a.wout[i] + v : v;
      }
    }
    __syncthreads();
  }
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
// max(a, b) that keeps a NaN from either side (fmax would drop it), so that a
// non-finite residual or solution is never read as a small backward error
__device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

// partial[b] = sum over the block of x_i y_i (mode 0) or max |x_i| (mode 1, NaN kept)
__global__ void reduce_f64_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                  int mode, double* __restrict__ partial) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc = mode == 0 ? acc + x[i] * y[i] : nan_max(acc, fabs(x[i]));
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(acc, off);
    acc = mode == 0 ? acc + o : nan_max(acc, o);
  }
  __shared__ double sw[4];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sw[0];
    for (int w = 1; w < 4; ++w) t = mode == 0 ? t + sw[w] : nan_max(t, sw[w]);
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
  float* invu_all = nullptr; // U11^-1 of every block column (filled by diagonal-block panels)
  std::vector<char> have_invu;
  float* inv_ready = nullptr; // L11^-1 the last diagonal-block panel already formed
  float* tmp = nullptr;      // inverse scratch
  const float* l16_of = nullptr;   // L21 whose rounded copy l16 holds (one block column's updates share it)
  const int32_t* block_pairs[2] = {nullptr, nullptr};   // composed interchanges by block-column parity
  int32_t* pending_pairs = nullptr;                      // the queued panel attempt's: its list
  int64_t pending_rows = 0;                              // ... and its row count
  int64_t block_pairs_j0[2] = {-1, -1};                 // ... of the block column starting here
  int64_t l16_rows = 0, l16_kb = 0;
  // optimistic panels
  bool optimistic = true;
  float tau = 16.0f;         // largest |l| accepted below the diagonal block
  float* pbak = nullptr;     // copy of the block column (rows x NB_MIX)
  float* ptmp = nullptr;     // A21 of the second half (rows x DB)
  float* scr3 = nullptr;     // DB x DB scratch blocks
  float* scr4 = nullptr;
  int* pfail = nullptr;
  int panels = 0, panels_fast = 0;
};
thread_local MixedCtx tl_mix;

void panel_gemm_f32(int64_t rows, int64_t cols, int64_t kb, const float* L, int64_t ldl,
                    const float* U, int64_t ldu, float* C, int64_t ldc, hipStream_t s) {
  mp_gemm_f32(rows, cols, kb, -1.0f, L, ldl, U, ldu, 1.0f, nullptr, nullptr, C, ldc, s);
}

void trailing_gemm(int64_t rows, int64_t cols, int64_t kb, const float* L, int64_t ldl,
                   const float* U, int64_t ldu, float* C, int64_t ldc, hipStream_t s) {
  MixedCtx& mx = tl_mix;
  if (mx.prec == kPrecF32) {
    void* tok = profile_begin(kProfGemmLowp, s);
    mp_gemm_f32(rows, cols, kb, -1.0f, L, ldl, U, ldu, 1.0f, nullptr, nullptr, C, ldc, s);
    profile_end(tok, s, 2.0 * rows * cols * kb, 4.0 * (rows * kb + kb * cols) + 8.0 * rows * cols);
    return;
  }
  const int64_t kp = (kb + 7) / 8 * 8;
  // the updates of one block column (next block, possibly in halves, then the
  // rest) run in order on this stream with the same L21: round it once
  const bool have_l = mx.l16_of == L && mx.l16_rows == rows && mx.l16_kb == kb;
  mx.l16_of = L;
  mx.l16_rows = rows;
  mx.l16_kb = kb;
  if (mx.prec == kPrecF16) {
    if (!have_l)
      hipLaunchKernelGGL((rows_to_lowp_kernel<_Float16>), dim3(blocks(rows, 4)), dim3(256), 0, s, rows,
                         (int)kb, L, ldl, mx.l16, kp, mx.rs);
    hipLaunchKernelGGL((cols_to_lowp_t_kernel<_Float16>), dim3(blocks(cols, 32)), dim3(256), 0, s, cols,
                       (int)kb, U, ldu, mx.u16, kp, mx.cs);
    ZH_LAUNCH_CHECK();
    void* tok = profile_begin(kProfGemmLowp, s);
    mp_gemm_f16(rows, cols, kb, -1.0f, mx.l16, kp, mx.u16, kp, 1.0f, mx.rs, mx.cs, C, ldc, s);
    profile_end(tok, s, 2.0 * rows * cols * kb, 2.0 * (rows * kp + kp * cols) + 8.0 * rows * cols);
  } else {
    __bf16* l = reinterpret_cast<__bf16*>(mx.l16);
    __bf16* u = reinterpret_cast<__bf16*>(mx.u16);
    if (!have_l)
      hipLaunchKernelGGL((rows_to_lowp_kernel<__bf16>), dim3(blocks(rows, 4)), dim3(256), 0, s, rows,
                         (int)kb, L, ldl, l, kp, mx.rs);
    hipLaunchKernelGGL((cols_to_lowp_t_kernel<__bf16>), dim3(blocks(cols, 32)), dim3(256), 0, s, cols,
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

// Inverse of a 64 x 64 lower triangular diagonal block (block b of the
// kb x kb matrix L, unit or not) into the same block of X (ld ldx): one wave,
// lane j solves L x = e_j by forward substitution with x in registers; the
// factor's entries are wave-uniform loads.
template <bool UNIT>
__global__ void __launch_bounds__(64) tri_inv64_kernel(const float* __restrict__ L, int64_t ld,
                                                       float* __restrict__ X, int64_t ldx) {
  const int j = threadIdx.x;
  const int64_t o = blockIdx.x * 64;
  const float* Lb = L + o * ld + o;
  float x[64];
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    const float* row = Lb + k * ld;
    float acc = k == j ? 1.0f : 0.0f;
#pragma unroll
    for (int i = 0; i < k; ++i) acc = fmaf(-row[i], x[i], acc);
    x[k] = UNIT ? acc : acc / row[k];
  }
  float* Xb = X + o * ldx + o;
#pragma unroll
  for (int k = 0; k < 64; ++k) Xb[k * ldx + j] = x[k];
}

// X (ld ldx) <- inverse of a 256 x 256 triangular matrix in one workgroup.
// !TRANS: L lower (ld ld), X = L^-1. TRANS: the input is an upper U; the
// kernel inverts L = U^T (reading U transposed) and writes X^T = U^-1.
// Phase A: wave w < 4 inverts the 64 x 64 diagonal block w (lane j solves for
// column j, the block staged in LDS). Phase B, column block J = 0..2 and row
// block I = J+1..3 in order: X_IJ = -X_II sum_{K=J}^{I-1} L_IK X_KJ, as 64 x 64
// products of LDS operands on the matrix cores (waves 0..3, one 32 x 32 tile
// each, 32x32x2 FP32 MFMA steps in k order).
// Global memory is read and written in whole blocks, numbered so that
// consecutive threads touch consecutive addresses in either orientation: the
// four diagonal blocks are loaded together, and each L block of phase B is
// loaded into registers while the previous product runs. The blocks above
// X's diagonal blocks are written as zeros.
constexpr int TI_LD = 65;
constexpr int TI_NT = 512;
typedef float f32x16 __attribute__((ext_vector_type(16)));
template <bool UNIT, bool TRANS>
__global__ void __launch_bounds__(TI_NT) tri_inv256_kernel(const float* __restrict__ L, int64_t ld,
                                                           float* __restrict__ X, int64_t ldx) {
  constexpr int NT = TI_NT;
  constexpr int SA = 4 * 4096 / NT;   // elements per thread of the four diagonal blocks
  constexpr int SB = 4096 / NT;       // of one 64 x 64 block
  __shared__ float smem[9 * 64 * TI_LD];
  float* Xd = smem;                       // 4 diagonal inverses, 64 x TI_LD each
  float* Xo = Xd + 4 * 64 * TI_LD;        // this column block's X_KJ, K > J (3)
  float* Ls = Xo + 3 * 64 * TI_LD;        // L_IK staging
  float* Tm = Ls + 64 * TI_LD;            // sum of products
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // element e of a 64 x 64 block: row er(e), column ec(e)
  auto er = [](int e) { return TRANS ? (e & 63) : (e >> 6); };
  auto ec = [](int e) { return TRANS ? (e >> 6) : (e & 63); };
  auto laddr = [&](int I, int K, int e) -> int64_t {
    const int64_t gr = I * 64 + er(e), gc = K * 64 + ec(e);
    return TRANS ? gc * ld + gr : gr * ld + gc;
  };
  auto xaddr = [&](int I, int J, int e) -> int64_t {
    const int64_t gr = I * 64 + er(e), gc = J * 64 + ec(e);
    return TRANS ? gc * ldx + gr : gr * ldx + gc;
  };
  // ---- phase A (Xo is free yet: it stages the four diagonal blocks, 4 fit in Xo + Ls)
  {
    float v[SA];
#pragma unroll
    for (int q = 0; q < SA; ++q) {
      const int f = t + NT * q, b = f >> 12, e = f & 4095;
      v[q] = L[laddr(b, b, e)];
    }
#pragma unroll
    for (int q = 0; q < SA; ++q) {
      const int f = t + NT * q, b = f >> 12, e = f & 4095;
      Xo[(b * 64 + er(e)) * TI_LD + ec(e)] = v[q];
    }
  }
  // the six blocks (I, J), I < J, of X are zero
  {
    constexpr int zI[6] = {0, 0, 0, 1, 1, 2}, zJ[6] = {1, 2, 3, 2, 3, 3};
#pragma unroll
    for (int q = 0; q < 6 * SB; ++q) {
      const int f = t + NT * q, b = f >> 12, e = f & 4095;
      X[xaddr(zI[b], zJ[b], e)] = 0.0f;
    }
  }
  float pf[SB];   // the next L block of phase B
  auto load_l = [&](int I, int K) {
#pragma unroll
    for (int q = 0; q < SB; ++q) pf[q] = L[laddr(I, K, t + NT * q)];
  };
  __syncthreads();
  if (w < 4) {
    const float* st = Xo + w * 64 * TI_LD;
    float x[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) {
      const float* row = st + k * TI_LD;
      float acc = k == lane ? 1.0f : 0.0f;
#pragma unroll
      for (int i = 0; i < k; ++i) acc = fmaf(-row[i], x[i], acc);
      x[k] = UNIT ? acc : acc / row[k];
    }
#pragma unroll
    for (int k = 0; k < 64; ++k) Xd[(w * 64 + k) * TI_LD + lane] = x[k];
  }
  load_l(1, 0);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SA; ++q) {
    const int f = t + NT * q, b = f >> 12, e = f & 4095;
    X[xaddr(b, b, e)] = Xd[(b * 64 + er(e)) * TI_LD + ec(e)];
  }
  // ---- phase B: wave w < 4 holds the 32 x 32 tile (mi, mj) of each product,
  // element r of lane l at row mi + mrow(r), column mj + l % 32
  const int mi = 32 * ((w & 3) >> 1), mj = 32 * (w & 1);
  auto mrow = [&](int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); };
  f32x16 mac;
  auto mma64 = [&](const float* Am, const float* Bm) {   // mac += Am Bm (64 x 64, ld TI_LD)
#pragma unroll 8
    for (int ks = 0; ks < 32; ++ks) {
      const float av = Am[(mi + (lane & 31)) * TI_LD + 2 * ks + (lane >> 5)];
      const float bv = Bm[(2 * ks + (lane >> 5)) * TI_LD + mj + (lane & 31)];
      mac = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, mac, 0, 0, 0);
    }
  };
  for (int J = 0; J < 3; ++J) {
    for (int I = J + 1; I < 4; ++I) {
#pragma unroll
      for (int r = 0; r < 16; ++r) mac[r] = 0.0f;
      for (int K = J; K < I; ++K) {
#pragma unroll
        for (int q = 0; q < SB; ++q) {
          const int e = t + NT * q;
          Ls[er(e) * TI_LD + ec(e)] = pf[q];
        }
        // the L block after this one: (I, K+1), else (I+1, J), else (J+2, J+1)
        int nI = I, nK = K + 1;
        if (nK == I) {
          nI = I + 1;
          nK = J;
          if (nI == 4) {
            nI = J + 2;
            nK = J + 1;
          }
        }
        if (nI < 4) load_l(nI, nK);
        __syncthreads();
        const float* Q = K == J ? Xd + J * 64 * TI_LD : Xo + (K - J - 1) * 64 * TI_LD;
        if (w < 4) mma64(Ls, Q);
        __syncthreads();
      }
      if (w < 4) {
#pragma unroll
        for (int r = 0; r < 16; ++r) Tm[(mi + mrow(r)) * TI_LD + mj + (lane & 31)] = mac[r];
      }
      __syncthreads();
      float* O = Xo + (I - J - 1) * 64 * TI_LD;
      if (w < 4) {
#pragma unroll
        for (int r = 0; r < 16; ++r) mac[r] = 0.0f;
        mma64(Xd + I * 64 * TI_LD, Tm);
#pragma unroll
        for (int r = 0; r < 16; ++r) O[(mi + mrow(r)) * TI_LD + mj + (lane & 31)] = -mac[r];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < SB; ++q) {
        const int e = t + NT * q;
        X[xaddr(I, J, e)] = O[er(e) * TI_LD + ec(e)];
      }
    }
  }
}

// X (ld ldx) <- inverse of the 256 x 256 lower triangular L (ld ld), unit or
// not (trans: of the upper U, through U^T); all of X's 256 x 256 is written.
void tri_inv256(const float* L, int64_t ld, float* X, int64_t ldx, int unit, hipStream_t s, bool trans = false) {
  if (trans) {
    if (unit) hipLaunchKernelGGL((tri_inv256_kernel<true, true>), dim3(1), dim3(TI_NT), 0, s, L, ld, X, ldx);
    else hipLaunchKernelGGL((tri_inv256_kernel<false, true>), dim3(1), dim3(TI_NT), 0, s, L, ld, X, ldx);
  } else {
    if (unit) hipLaunchKernelGGL((tri_inv256_kernel<true, false>), dim3(1), dim3(TI_NT), 0, s, L, ld, X, ldx);
    else hipLaunchKernelGGL((tri_inv256_kernel<false, false>), dim3(1), dim3(TI_NT), 0, s, L, ld, X, ldx);
  }
  ZH_LAUNCH_CHECK();
}

// inv <- L^-1 (kb x kb, ld kb, lower; unit or not). kb = 256 and 512 go
// through tri_inv256 (512: X21 = -X22 L21 X11 once more); other sizes through
// forward substitution on the identity.
void tri_inverse_f32(int64_t kb, const float* L, int64_t ld, float* inv, int unit, hipStream_t s) {
  hipLaunchKernelGGL(fill_identity_f32_kernel, dim3(blocks(kb * kb)), dim3(256), 0, s, kb, inv, kb);
  ZH_LAUNCH_CHECK();
  if (kb == 256) {
    tri_inv256(L, ld, inv, kb, unit, s);
    return;
  }
  if (kb == 512) {
    // X22 on a second stream while X11 and T = L21 X11 run on s
    float* T = static_cast<float*>(workspace(sizeof(float) * 256 * 256, 17));
    const hipStream_t s2 = aux_stream(2);
    const hipEvent_t ev_fork = aux_event(6), ev_join = aux_event(7);
    ZH_CHECK(hipEventRecord(ev_fork, s));
    ZH_CHECK(hipStreamWaitEvent(s2, ev_fork, 0));
    tri_inv256(L + 256 * ld + 256, ld, inv + 256 * kb + 256, kb, unit, s2);
    ZH_CHECK(hipEventRecord(ev_join, s2));
    tri_inv256(L, ld, inv, kb, unit, s);
    mp_gemm_f32(256, 256, 256, 1.0f, L + 256 * ld, ld, inv, kb, 0.0f, nullptr, nullptr, T, 256, s);
    ZH_CHECK(hipStreamWaitEvent(s, ev_join, 0));
    mp_gemm_f32(256, 256, 256, -1.0f, inv + 256 * kb + 256, kb, T, 256, 0.0f, nullptr, nullptr, inv + 256 * kb, kb,
                s);
    return;
  }
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

// B (kb x cols) <- inv B. Only the update stream calls it (U12 of the next block
// column and of the rest); a product of few output tiles (the next block
// column's) splits K so that enough workgroups share it.
void apply_inv_f32(int64_t kb, int64_t cols, const float* inv, float* B, int64_t ldb, float* tmp,
                   hipStream_t s) {
  ZH_CHECK(hipMemcpy2DAsync(tmp, sizeof(float) * cols, B, sizeof(float) * ldb, sizeof(float) * cols, kb,
                            hipMemcpyDeviceToDevice, s));
  const int64_t tiles = ((kb + 127) / 128) * ((cols + 127) / 128);
  if (tiles < 64 && kb >= 256) {
    float* part = static_cast<float*>(workspace(sizeof(float) * mp_gemm_f32_splitk_floats(kb, cols, kb), 18));
    mp_gemm_f32_splitk(kb, cols, kb, 1.0f, inv, kb, tmp, cols, 0.0f, B, ldb, part, s);
    return;
  }
  mp_gemm_f32(kb, cols, kb, 1.0f, inv, kb, tmp, cols, 0.0f, nullptr, nullptr, B, ldb, s);
}

// C = alpha A B + beta C (FP32, row-major) for the optimistic panels: a
// product of few output tiles (the 256 x 256 blocks, L21 of a short panel)
// splits K so that enough workgroups share it. Its partial sums use their own
// workspace slot, apart from the update stream's.
constexpr size_t kPanelSplitBytes = sizeof(float) * 8 * 128 * 128 * 128;   // 8 slices x 128 tiles
void panel_product_f32(int64_t m, int64_t n, int64_t K, float alpha, const float* A, int64_t lda, const float* B,
                       int64_t ldb, float beta, float* C, int64_t ldc, hipStream_t s) {
  const int64_t tiles = ((m + 127) / 128) * ((n + 127) / 128);
  if (tiles < 128 && K >= 128) {   // (the solvers reserve kPanelSplitBytes up front)
    float* part = static_cast<float*>(workspace(sizeof(float) * mp_gemm_f32_splitk_floats(m, n, K), 21));
    mp_gemm_f32_splitk(m, n, K, alpha, A, lda, B, ldb, beta, C, ldc, part, s);
    return;
  }
  mp_gemm_f32(m, n, K, alpha, A, lda, B, ldb, beta, nullptr, nullptr, C, ldc, s);
}

// One optimistic 256-wide block column: partial pivoting inside its diagonal
// block P (ld lda, rows x DB with rows >= DB), then L21 = A21 U11^-1 with A21
// taken from src21 (rows - DB x DB, ld src_ld; P's own rows below the block
// are overwritten). U11^-1 goes to uinv (ld uinv_ld). *pfail records a zero
// pivot (1) or an |l| above the threshold (2).
// linv (optional): the inverse of the block's unit lower factor is formed there
// (its strictly upper part already zero) on a second stream, joined by the
// event aux_event(5), while U11^-1 and L21 take this stream; the block's
// ipiv entries reach their LAPACK form there too (until then they hold the
// block's row order). Without linv that happens on this stream.
// The block's row moves go to `pairs`, and to the block column's list
// `bpairs` when given (appended if `bappend`).
void optimistic_block(float* P, int64_t lda, int64_t rows, const float* src21, int64_t src_ld, int32_t* ipiv,
                      int base, float* uinv, int64_t uinv_ld, hipStream_t s, int32_t* pairs,
                      float* linv = nullptr, int64_t linv_ld = 0, int32_t* bpairs = nullptr, int bappend = 0,
                      float* linv_init = nullptr) {
  MixedCtx& mx = tl_mix;
  hipLaunchKernelGGL(diag_getf2_kernel<DIAG_TY>, dim3(1), dim3(DIAG_TY * 32), 0, s, P, lda, ipiv, base, mx.pfail,
                     pairs, bpairs, bappend);
  ZH_LAUNCH_CHECK();
  if (linv_init) {   // the block column's whole L11^-1 buffer starts as the identity
    hipLaunchKernelGGL(fill_identity_f32_kernel, dim3(blocks(linv_ld * linv_ld)), dim3(256), 0, s, linv_ld,
                       linv_init, linv_ld);
    ZH_LAUNCH_CHECK();
  }
  if (linv) {
    const hipStream_t s2 = aux_stream(2);
    ZH_CHECK(hipEventRecord(aux_event(4), s));
    ZH_CHECK(hipStreamWaitEvent(s2, aux_event(4), 0));
    tri_inv256(P, lda, linv, linv_ld, 1, s2);
    hipLaunchKernelGGL(rowmap_to_ipiv_kernel, dim3(1), dim3(DB), 0, s2, ipiv, base);
    ZH_LAUNCH_CHECK();
    ZH_CHECK(hipEventRecord(aux_event(5), s2));
  } else {
    hipLaunchKernelGGL(rowmap_to_ipiv_kernel, dim3(1), dim3(DB), 0, s, ipiv, base);
    ZH_LAUNCH_CHECK();
  }
  // U11^-1 (upper, its lower blocks written as zeros) straight into its store
  tri_inv256(P, lda, uinv, uinv_ld, 0, s, true);
  if (rows > DB) {
    panel_product_f32(rows - DB, DB, DB, 1.0f, src21, src_ld, uinv, uinv_ld, 0.0f, P + DB * lda, lda, s);
    hipLaunchKernelGGL(bound_check_kernel, dim3(blocks((rows - DB) * DB / 16)), dim3(256), 0, s, rows - DB, DB,
                       (const float*)(P + DB * lda), lda, mx.tau, mx.pfail);
    ZH_LAUNCH_CHECK();
  }
}

// Optimistic block column of width 256 or 512 (two 256-wide halves: the
// first half's interchanges and U12 = L11^-1 A12 update the second half, which
// is then factored the same way, and its interchanges reach the first half's
// columns, as getrf2.zig:99-202 orders it). Pivots are taken inside the
// diagonal blocks; with every |l_ij| <= 1 no row below a block would have
// won a pivot search, so the result is the partial-pivoting one (up to
// rounding and ties); with |l_ij| <= tau > 1 it is threshold pivoting, whose
// growth per step is at most 1 + tau (the FP64 refinement then restores the
// accuracy, and GMRES-IR or the FP64 factorization take over if it cannot).
// try_panel_f32 queues the attempt; finish_panel_f32 waits for it and, when it
// failed, restores the block column, which the caller then factors with the
// full pivot search.
const int32_t* block_perm_f32(int64_t j0) {
  const MixedCtx& mx = tl_mix;
  const int par = (int)((j0 / NB_MIX) & 1);
  return mx.block_pairs_j0[par] == j0 ? mx.block_pairs[par] : nullptr;
}

bool try_panel_f32(int64_t m, float* A, int64_t lda, int64_t j0, int64_t jb, int32_t* ipiv,
                   hipStream_t s) {
  MixedCtx& mx = tl_mix;
  const int64_t rows = m - j0;
  mx.block_pairs_j0[(j0 / NB_MIX) & 1] = -1;
  if (!mx.optimistic || (jb != DB && jb != 2 * DB) || rows < jb) return false;
  ++mx.panels;
  mx.inv_ready = nullptr;
  mx.pending_rows = rows;
  float* P = A + j0 * lda + j0;
  float* ustore = mx.invu_all + (j0 / NB_MIX) * NB_MIX * NB_MIX;   // ld NB_MIX
  ZH_CHECK(hipMemcpy2DAsync(mx.pbak, sizeof(float) * jb, P, sizeof(float) * lda, sizeof(float) * jb, rows,
                            hipMemcpyDeviceToDevice, s));
  // L11^-1 (512 x 512, kept for the driver's U12): La^-1 and Lb^-1 beside the
  // halves' other work, the off-diagonal block at the end
  float* linv = jb == 2 * DB ? mx.inv_all + (j0 / NB_MIX) * NB_MIX * NB_MIX : nullptr;
  // row moves: each half's own list, and the block column's (both halves) for
  // the driver, in the buffer of the block column's parity
  int32_t* pr = perm_pairs_buffer(2);
  const int par = (int)((j0 / NB_MIX) & 1);
  int32_t* bpr = jb == NB_MIX ? perm_pairs_buffer(19 + par) : nullptr;
  optimistic_block(P, lda, rows, mx.pbak + DB * jb, jb, ipiv + j0, (int)j0, ustore, NB_MIX, s, pr, linv, jb, bpr, 0,
                   linv);
  if (jb == 2 * DB) {
    const int64_t c2 = j0 + DB;                 // second half: columns c2 .. c2 + DB
    const int64_t rows2 = rows - DB;
    apply_perm<float>(DB, A + c2, lda, pr, s);
    ZH_CHECK(hipStreamWaitEvent(s, aux_event(5), 0));   // La^-1
    // U12 of the first half = La^-1 A12, then A22 -= L21 U12 within the block column
    ZH_CHECK(hipMemcpy2DAsync(mx.scr4, sizeof(float) * DB, A + j0 * lda + c2, sizeof(float) * lda,
                              sizeof(float) * DB, DB, hipMemcpyDeviceToDevice, s));
    panel_product_f32(DB, DB, DB, 1.0f, linv, jb, mx.scr4, DB, 0.0f, A + j0 * lda + c2, lda, s);
    panel_product_f32(rows2, DB, DB, -1.0f, A + c2 * lda + j0, lda, A + j0 * lda + c2, lda, 1.0f, A + c2 * lda + c2,
                      lda, s);
    if (rows2 > DB)
      ZH_CHECK(hipMemcpy2DAsync(mx.ptmp, sizeof(float) * DB, A + (c2 + DB) * lda + c2, sizeof(float) * lda,
                                sizeof(float) * DB, rows2 - DB, hipMemcpyDeviceToDevice, s));
    optimistic_block(A + c2 * lda + c2, lda, rows2, mx.ptmp, DB, ipiv + c2, (int)c2,
                     ustore + DB * NB_MIX + DB, NB_MIX, s, pr, linv + DB * jb + DB, jb, bpr, 1);
    apply_perm<float>(DB, A + j0, lda, pr, s);
    // the off-diagonal block -Lb^-1 Lba La^-1
    ZH_CHECK(hipStreamWaitEvent(s, aux_event(5), 0));   // Lb^-1
    panel_product_f32(DB, DB, DB, 1.0f, A + c2 * lda + j0, lda, linv, jb, 0.0f, mx.scr3, DB, s);
    panel_product_f32(DB, DB, DB, -1.0f, linv + DB * jb + DB, jb, mx.scr3, DB, 0.0f, linv + DB * jb, jb, s);
  }
  mx.pending_pairs = bpr;
  return true;
}

// The outcome of the attempt try_panel_f32 queued for block column j0.
bool finish_panel_f32(int64_t j0, int64_t jb, float* A, int64_t lda, hipStream_t s) {
  MixedCtx& mx = tl_mix;
  float* P = A + j0 * lda + j0;
  const int64_t rows = mx.pending_rows;
  const int par = (int)((j0 / NB_MIX) & 1);
  int fail = 0;
  ZH_CHECK(hipMemcpyAsync(&fail, mx.pfail, sizeof(int), hipMemcpyDeviceToHost, s));
  ZH_CHECK(hipStreamSynchronize(s));
  if (fail == 0) {
    ++mx.panels_fast;
    // the solves read the 128 x 128 diagonal sub-blocks of U^-1: both halves are there
    if (jb == NB_MIX) {
      mx.have_invu[j0 / NB_MIX] = 1;
      mx.inv_ready = mx.inv_all + (j0 / NB_MIX) * NB_MIX * NB_MIX;
      mx.block_pairs[par] = mx.pending_pairs;
      mx.block_pairs_j0[par] = j0;
    }
    return true;
  }
  ZH_CHECK(hipMemcpy2DAsync(P, sizeof(float) * lda, mx.pbak, sizeof(float) * jb, sizeof(float) * jb, rows,
                            hipMemcpyDeviceToDevice, s));
  return false;
}

void invert_f32(int64_t kb, const float* L11, int64_t ld, float* inv, hipStream_t s) {
  tri_inverse_f32(kb, L11, ld, inv, 1, s);
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
  float* z = nullptr;          // forward-solve result (n)
  uint64_t* xch = nullptr;     // tagged exchange words of the triangular solves (n)
  int* serr = nullptr;         // triangular-solve stall flag
  mutable uint32_t tag = 0;    // last tag used in xch
  unsigned ts_grid = 1;        // workgroups of a triangular solve
  double* part = nullptr;      // reduction partials
  double anorm = 0.0;
  int64_t nblk = 0;
  // w (FP64) = M^-1 v (accumulate: w += M^-1 v), M the low-precision factorization
  void (*precond)(const Solver&, const double* v, double* w, int accumulate) = nullptr;
  // Cholesky only
  float* invLT = nullptr;      // transposed diagonal-block inverses of L
  float* ybuf = nullptr;       // (n + NB) x 16 right-hand sides, FP32 row-major
  double* pbuf = nullptr;      // split partial products of the FP64 residual (16 x n x splits)
};

double reduce_host(const Solver& S, int64_t n, const double* x, const double* y, int mode) {
  const int nb = 256;
  hipLaunchKernelGGL(reduce_f64_kernel, dim3(nb), dim3(256), 0, S.s, n, x, y, mode, S.part);
  ZH_LAUNCH_CHECK();
  std::vector<double> h(nb);
  ZH_CHECK(hipMemcpyAsync(h.data(), S.part, sizeof(double) * nb, hipMemcpyDeviceToHost, S.s));
  ZH_CHECK(hipStreamSynchronize(S.s));
  double acc = 0.0;
  for (double v : h) acc = mode == 0 ? acc + v : ((v > acc || v != v) ? v : acc);   // NaN kept
  return acc;
}

// w (FP64) = Dc U^-1 L^-1 P Dr v; accumulate: w += ... instead
void lu_solve(const Solver& S, const double* v, double* w, int accumulate) {
  TriSolve f{};
  f.n = S.n;
  f.nblk = (S.n + TS - 1) / TS;
  f.upper = 0;
  f.W = S.W;
  f.ldw = S.n;
  f.inv = S.invL;
  f.v = v;
  f.perm = S.perm;
  f.rsc = S.r;
  f.zout = S.z;
  f.xch = S.xch;
  f.tag = ++S.tag;
  f.err = S.serr;
  f.spin_limit = int64_t(1) << 22;
  hipLaunchKernelGGL(trisolve_f32_kernel, dim3(S.ts_grid), dim3(256), 0, S.s, f);
  ZH_LAUNCH_CHECK();
  TriSolve g = f;
  g.upper = 1;
  g.inv = S.invU;
  g.v = nullptr;
  g.yin = S.z;
  g.zout = nullptr;
  g.csc = S.c;
  g.wout = w;
  g.accumulate = accumulate;
  g.tag = ++S.tag;
  hipLaunchKernelGGL(trisolve_f32_kernel, dim3(S.ts_grid), dim3(256), 0, S.s, g);
  ZH_LAUNCH_CHECK();
}

// r = b - A x (FP64, the caller's A)
void residual(const Solver& S, const double* b, const double* x, double* r) {
  ZH_CHECK(hipMemcpyAsync(r, b, sizeof(double) * S.n, hipMemcpyDeviceToDevice, S.s));
  const int trans = S.layout == kColMajor ? kNoTrans : kTrans;
  gemv_f64_cm(trans, S.n, S.n, -1.0, S.A, S.lda, x, 1, 1.0, r, 1, S.s);
}

void check_solve(const Solver& S) {
  if (!S.serr) return;   // no persistent triangular solves in use
  int e = 0;
  ZH_CHECK(hipMemcpyAsync(&e, S.serr, sizeof(int), hipMemcpyDeviceToHost, S.s));
  ZH_CHECK(hipStreamSynchronize(S.s));
  if (e) throw SyncError("triangular solve stalled (spin limit reached)");
}

double backward_error(const Solver& S, const double* r, const double* x) {
  check_solve(S);
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
      S.precond(S, vj, zj, 0);
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

// ---- mixed-precision Cholesky ------------------------------------------------------------
// S (n x n, FP64, row-major) = the symmetric matrix whose `uplo` triangle the
// caller stores in a (layout, lda); 32 x 32 tiles of the lower triangle, each
// written to (I, J) and, transposed through LDS, to (J, I).
__global__ void symmetrize_kernel(int layout, int uplo, int64_t n, const double* __restrict__ a,
                                  int64_t lda, double* __restrict__ S) {
  __shared__ double tile[32][33];
  const int64_t I = blockIdx.y * 32, J = blockIdx.x * 32;
  if (J > I) return;
  const int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
  // element (i, j), i >= j, of the lower triangle, as stored by the caller
  const bool rowwise = (layout == kRowMajor) == (uplo == kLower);   // j contiguous in memory
  for (int k = ty; k < 32; k += 8) {
    const int ii = rowwise ? k : tx, jj = rowwise ? tx : k;
    const int64_t i = I + ii, j = J + jj;
    double v = 0.0;
    if (i < n && j < n && i >= j) {
      if (layout == kRowMajor) v = uplo == kLower ? a[i * lda + j] : a[j * lda + i];
      else v = uplo == kLower ? a[i + j * lda] : a[j + i * lda];
    }
    tile[ii][jj] = v;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t i = I + k, j = J + tx;   // S(i, j), i >= j
    if (i < n && j < n && i >= j) S[i * n + j] = tile[k][tx];
    const int64_t i2 = J + k, j2 = I + tx;   // S(i2, j2) = S(j2, i2), i2 <= j2
    if (i2 < n && j2 < n && j2 > i2) S[i2 * n + j2] = tile[tx][k];
  }
}

// d_i = 2^-floor(log2 sqrt(a_ii)) (1 for a non-positive or non-finite diagonal)
__global__ void diag_scale_kernel(int64_t n, const double* __restrict__ S, double* __restrict__ d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = S[i * n + i];
    d[i] = v > 0.0 ? pow2_scale(sqrt(v)) : 1.0;
  }
}

// dst (kb x kb, FP64, ld ldd) = src (FP32, ld lds) + shift I
// dst[i][j] = src[i][j] (trans: dst[j][i]) + shift on the diagonal, for a kb x kb block
__global__ void block_to_f64_kernel(int64_t kb, const float* __restrict__ src, int64_t lds,
                                    double* __restrict__ dst, int64_t ldd, double shift, int trans = 0) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= kb * kb) return;
  const int64_t i = e / kb, j = e % kb;
  dst[trans ? j * ldd + i : i * ldd + j] = (double)src[i * lds + j] + (i == j ? shift : 0.0);
}

// dst[i][j] = fl32(trans ? src[j][i] : src[i][j]) for a kb x kb block; lower_only
// writes j <= i only
__global__ void block_to_f32_kernel(int64_t kb, const double* __restrict__ src, int64_t lds,
                                    float* __restrict__ dst, int64_t ldd, int trans, int lower_only) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= kb * kb) return;
  const int64_t i = e / kb, j = e % kb;
  if (lower_only && j > i) return;
  dst[i * ldd + j] = (float)(trans ? src[j * lds + i] : src[i * lds + j]);
}


// dst (cols x rows, ld ldd) = src^T (rows x cols, ld lds), 32 x 32 tiles
__global__ void transpose_rect_f32_kernel(int64_t rows, int64_t cols, const float* __restrict__ src,
                                          int64_t lds, float* __restrict__ dst, int64_t ldd) {
  __shared__ float tile[32][33];
  const int64_t R = blockIdx.y * 32, Cc = blockIdx.x * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int k = ty; k < 32; k += 8) {
    const int64_t i = R + k, j = Cc + tx;
    tile[k][tx] = (i < rows && j < cols) ? src[i * lds + j] : 0.0f;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t j = Cc + k, i = R + tx;
    if (i < rows && j < cols) dst[j * ldd + i] = tile[tx][k];
  }
}

// Y (n x nr, FP32 row-major) = fl32(d_i R_iq), R column-major (ld ldr)
__global__ void pack_rows_kernel(int64_t n, int nr, const double* __restrict__ R, int64_t ldr,
                                 const double* __restrict__ d, float* __restrict__ Y) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n * nr) return;
  const int64_t i = e / nr;
  const int q = (int)(e % nr);
  Y[e] = (float)(d[i] * R[q * ldr + i]);
}
// D_iq (+)= d_i Y_iq
__global__ void unpack_rows_kernel(int64_t n, int nr, const float* __restrict__ Y,
                                   const double* __restrict__ d, double* __restrict__ D, int64_t ldd,
                                   int accumulate) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n * nr) return;
  const int64_t i = e / nr;
  const int q = (int)(e % nr);
  const double v = d[i] * (double)Y[e];
  double* p = D + q * ldd + i;
  *p = accumulate ? *p + v : v;
}

// D_iq += d_i Y_iq for the columns q set in mask
__global__ void unpack_add_masked_kernel(int64_t n, int nr, const float* __restrict__ Y,
                                         const double* __restrict__ d, double* __restrict__ D, int64_t ldd,
                                         unsigned mask) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n * nr) return;
  const int64_t i = e / nr;
  const int q = (int)(e % nr);
  if ((mask >> q) & 1u) D[q * ldd + i] += d[i] * (double)Y[e];
}

constexpr int MAXR = 16;   // right-hand sides per group of the Cholesky solve

// ---- triangular solves with up to 16 right-hand sides (persistent, FP32 MFMA) -------------
// L z = y (trans 0) or L^T z = y (trans 1) for the row-major lower factor L in W
// and Y (n x nr, FP32 row-major, ld ldy) overwritten by z. Block rows of T16
// are dealt to the workgroups in dependency order (as in trisolve_f32_kernel);
// wave w owns rows 32w .. 32w + 31 of its block and all 16 right-hand sides as
// two 16 x 16 accumulator tiles of v_mfma_f32_16x16x4f32. A finished block is
// published as n x 16 tagged words (tag << 32 | float), which a consumer
// workgroup stages once per column block in LDS. z_b = D_b (y_b - sum), D_b the
// T16 x T16 diagonal sub-block of the stored inverse of the enclosing NB_MIX
// block (inverse of L_BB for trans 0, its transpose for trans 1).
// Operand mapping of the MFMA (A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16 (i / 4), register i % 4): for trans 0 lane (i, kk)
// walks 4 consecutive columns c = 16 m + 4 kk + e of its row (one 16-byte load
// when the row start is aligned); for trans 1 it reads L[c][r0 + i] with
// c = 4 m + kk, 16 lanes covering 64 contiguous bytes of a row of L.
constexpr int T16 = 128;
typedef float f4v __attribute__((ext_vector_type(4)));

struct TriSolve16 {
  int64_t n, nblk;
  int trans, nr;
  const float* W;
  int64_t ldw;
  const float* inv;
  float* Y;
  int64_t ldy;
  uint64_t* xch;
  uint32_t tag;
  int* err;
  int64_t spin_limit;
};

__device__ __forceinline__ float wait_word16(const uint64_t* p, uint32_t tag, int64_t limit, int* err) {
  return wait_word(p, tag, limit, err);
}

template <bool VEC>
__device__ __forceinline__ f4v load_row4(const float* p, int valid) {
  if (VEC && valid >= 4) return *reinterpret_cast<const f4v*>(p);
  f4v v = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (e < valid) v[e] = p[e];
  return v;
}

template <bool VEC, bool TRANS>
__global__ void __launch_bounds__(256) trisolve16_kernel(TriSolve16 a) {
  __shared__ float zs[2][T16][17];
  __shared__ float ts[T16][17];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int li = lane & 15, kk = lane >> 4;
  const int wr = w * 32;
  for (int64_t bi = blockIdx.x; bi < a.nblk; bi += gridDim.x) {
    const int64_t b = TRANS ? a.nblk - 1 - bi : bi;
    const int64_t r0 = b * T16;
    const int rows = (int)std::min<int64_t>(T16, a.n - r0);
    // D_b rows wr + 16u + li, columns 16 m + 4 kk + e
    const int64_t B = r0 / NB_MIX;
    const int64_t o = r0 - B * NB_MIX;
    const int64_t ldi = std::min<int64_t>(NB_MIX, a.n - B * NB_MIX);
    const float* D = a.inv + B * NB_MIX * NB_MIX + o * ldi + o;
    const bool dvec = VEC && (ldi & 3) == 0;
    f4v dreg[2][8];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int r = wr + 16 * u + li;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int c = 16 * m + 4 * kk;
        const int valid = r < rows ? rows - c : 0;
        dreg[u][m] = dvec ? load_row4<true>(D + r * ldi + c, valid) : load_row4<false>(D + r * ldi + c, valid);
      }
    }
    f4v acc[2];
    acc[0] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    acc[1] = acc[0];
    const int64_t cnt = TRANS ? a.nblk - 1 - b : b;
    for (int64_t q = 0; q < cnt; ++q) {
      const int64_t j = TRANS ? a.nblk - 1 - q : q;
      const int64_t c0 = j * T16;
      const int cols = (int)std::min<int64_t>(T16, a.n - c0);
      const int buf = (int)(q & 1);
      // the factor's values first (independent of z_j)
      float lv[2][32];
      if (!TRANS) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int r = wr + 16 * u + li;
          const float* row = a.W + (r0 + r) * a.ldw + c0;
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const int c = 16 * m + 4 * kk;
            const f4v v = load_row4<VEC>(row + c, r < rows ? cols - c : 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) lv[u][4 * m + e] = v[e];
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int r = wr + 16 * u + li;
#pragma unroll
          for (int m = 0; m < 32; ++m) {
            const int c = 4 * m + kk;
            lv[u][m] = (r < rows && c < cols) ? a.W[(c0 + c) * a.ldw + r0 + r] : 0.0f;
          }
        }
      }
      // z_j into LDS: 8 words per thread, consecutive threads on consecutive words
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = t + 256 * u;
        const int c = e >> 4, qq = e & 15;
        zs[buf][c][qq] = c < cols ? wait_word16(a.xch + (c0 + c) * 16 + qq, a.tag, a.spin_limit, a.err) : 0.0f;
      }
      __syncthreads();
      if (!TRANS) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int e = 0; e < 4; ++e)
              acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(lv[u][4 * m + e], zs[buf][16 * m + 4 * kk + e][li],
                                                            acc[u], 0, 0, 0);
      } else {
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int m = 0; m < 32; ++m)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(lv[u][m], zs[buf][4 * m + kk][li], acc[u], 0, 0, 0);
      }
    }
    // t = y - sum (rows 16u + 4 kk + reg of this wave, right-hand side li)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int r = wr + 16 * u + 4 * kk + g;
        const float y = (r < rows && li < a.nr) ? a.Y[(r0 + r) * a.ldy + li] : 0.0f;
        ts[r][li] = r < rows ? y - acc[u][g] : 0.0f;
      }
    __syncthreads();
    // z_b = D_b t
    f4v z[2];
    z[0] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    z[1] = z[0];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          z[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(dreg[u][m][e], ts[16 * m + 4 * kk + e][li], z[u], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int r = wr + 16 * u + 4 * kk + g;
        if (r < rows) {
          const int64_t i = r0 + r;
          __hip_atomic_store(a.xch + i * 16 + li, ((uint64_t)a.tag << 32) | __float_as_uint(z[u][g]),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (li < a.nr) a.Y[i * a.ldy + li] = z[u][g];
        }
      }
    __syncthreads();
  }
}

// ---- FP64 residual products with up to 16 right-hand sides ---------------------------------
// P[s] (16 x n, ld n) = S[:, ks] X[ks, :] over the column range ks of split s,
// S row-major (ld lds), X column-major (ld ldx, nr <= 16 columns). A workgroup
// owns 64 rows (wave w: rows 16 w .. 16 w + 15) of one split and walks it in
// 64-column steps: the X slice is staged in LDS, S is read straight into the
// v_mfma_f64_16x16x4f64 A operand, lane (i, kk) taking columns 8 m + 2 kk + h
// (m = 0..7, h = 0..1) of its row as 16-byte loads, the next step's loads
// issued before the current step's products.
typedef double d4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <bool VEC>
__global__ void __launch_bounds__(256) skinny_mfma_f64_kernel(int64_t n, int nr, const double* __restrict__ S,
                                                              int64_t lds, const double* __restrict__ X,
                                                              int64_t ldx, int64_t kper, double* __restrict__ P) {
  __shared__ double xs[2][64][17];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, li = lane & 15, kk = lane >> 4;
  const int64_t row = blockIdx.x * (int64_t)64 + w * 16 + li;
  const int64_t kbeg = blockIdx.y * kper, kend = std::min<int64_t>(n, kbeg + kper);
  const double* srow = S + (row < n ? row : 0) * lds;
  auto load_s = [&](int64_t k0, d2v* sv) {
    const bool full = row < n && k0 + 64 <= kend;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int64_t c = k0 + 8 * m + 2 * kk;
      if (VEC && full) {
        sv[m] = *reinterpret_cast<const d2v*>(srow + c);
      } else {
        sv[m][0] = (row < n && c < kend) ? srow[c] : 0.0;
        sv[m][1] = (row < n && c + 1 < kend) ? srow[c + 1] : 0.0;
      }
    }
  };
  d4v acc = {0.0, 0.0, 0.0, 0.0};
  d2v cur[8], nxt[8];
  if (kbeg < kend) load_s(kbeg, nxt);
  int it = 0;
  for (int64_t k0 = kbeg; k0 < kend; k0 += 64, ++it) {
    const int buf = it & 1;
#pragma unroll
    for (int m = 0; m < 8; ++m) cur[m] = nxt[m];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + 256 * u;
      const int c = e & 63, q = e >> 6;
      xs[buf][c][q] = (k0 + c < kend && q < nr) ? X[q * ldx + k0 + c] : 0.0;
    }
    if (k0 + 64 < kend) load_s(k0 + 64, nxt);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(cur[m][h], xs[buf][8 * m + 2 * kk + h][li], acc, 0, 0, 0);
  }
  // D[i][j] in lane j + 16 (i % 4), register i / 4
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int64_t r = blockIdx.x * (int64_t)64 + w * 16 + 4 * g + kk;
    if (r < n && li < nr) P[(blockIdx.y * (int64_t)16 + li) * n + r] = acc[g];
  }
}

// Y[:, q] = beta Y[:, q] + alpha sum_s P[s][q] (splits summed in order)
__global__ void skinny_reduce_kernel(int64_t n, int nr, int splits, const double* __restrict__ P, double alpha,
                                     double beta, double* __restrict__ Y, int64_t ldy) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n * nr) return;
  const int64_t r = e % n;
  const int q = (int)(e / n);
  double acc = 0.0;
  for (int sp = 0; sp < splits; ++sp) acc += P[(sp * (int64_t)16 + q) * n + r];
  double* y = Y + q * ldy + r;
  *y = beta == 0.0 ? alpha * acc : beta * *y + alpha * acc;
}

// column split of the residual product: enough workgroups to cover the CUs a few times
struct SkinnySplit {
  int64_t kper;
  int splits;
};
SkinnySplit skinny_split(int64_t n) {
  const int64_t rb = (n + 63) / 64;
  int64_t want = std::max<int64_t>(1, (2048 + rb - 1) / rb);
  want = std::min<int64_t>(want, (n + 63) / 64);
  int64_t kper = ((n + want - 1) / want + 63) / 64 * 64;
  return SkinnySplit{kper, (int)((n + kper - 1) / kper)};
}

// Y (n x nr, column-major, ld ldy) = beta Y + alpha S X with P as scratch
// (16 x n x splits doubles)
void skinny_gemm_f64(const Solver& S, int nr, double alpha, const double* A, int64_t lda, const double* X,
                     int64_t ldx, double beta, double* Y, int64_t ldy, double* P) {
  const int64_t n = S.n;
  const SkinnySplit sp = skinny_split(n);
  const dim3 grid((unsigned)((n + 63) / 64), (unsigned)sp.splits);
  const bool vec = (lda & 1) == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
  if (vec)
    hipLaunchKernelGGL(skinny_mfma_f64_kernel<true>, grid, dim3(256), 0, S.s, n, nr, A, lda, X, ldx, sp.kper, P);
  else
    hipLaunchKernelGGL(skinny_mfma_f64_kernel<false>, grid, dim3(256), 0, S.s, n, nr, A, lda, X, ldx, sp.kper, P);
  ZH_LAUNCH_CHECK();
  hipLaunchKernelGGL(skinny_reduce_kernel, dim3(blocks(n * nr)), dim3(256), 0, S.s, n, nr, sp.splits,
                     (const double*)P, alpha, beta, Y, ldy);
  ZH_LAUNCH_CHECK();
}

// out[q] = max_i |M[q ld + i]| (out zeroed by the caller). A NaN is kept: its
// bit pattern (sign cleared) orders above +inf in the integer maximum.
__global__ void colmax_kernel(int64_t n, const double* __restrict__ M, int64_t ld, double* __restrict__ out) {
  const int q = blockIdx.y;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc = nan_max(acc, fabs(M[q * ld + i]));
  for (int off = 32; off > 0; off >>= 1) acc = nan_max(acc, __shfl_xor(acc, off));
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&out[q], fabs(acc));
}

struct CholWork {
  double* d64 = nullptr;       // diagonal block, FP64
  int* info_blk = nullptr;     // potrf info of every diagonal block
  float* tmp = nullptr;        // A21 copy, (n + NB) x NB
  float* l21t[2] = {};         // L21^T for the FP32 update, NB x (n + NB), per block parity
  _Float16* l16[2] = {};       // 16-bit L21, (n + NB) x NB, per block parity
  float* rs[2] = {};           // its row scales
};

// Lower Cholesky of W + shift I (W row-major, the lower triangle): diagonal
// blocks in FP64 (potrf and their inverses), L21 = A21 L11^-T as an FP32 GEMM
// with that inverse, the trailing lower triangle updated in FP32 / FP16 / BF16.
// Look-ahead of depth one: the panel stream factors block b+1 (and rounds its
// L21 to 16 bits) while the main stream, which spares a few CUs for it,
// applies block b beyond block b+1.
void chol_factor(Solver& S, const CholWork& cw, int prec, double shift) {
  const int64_t n = S.n;
  const hipStream_t caller = S.s;
  const hipStream_t sm = stream_sparing_cus(), sp = aux_stream(0);
  hipEvent_t ev_in = aux_event(0), ev_panel = aux_event(1), ev_next = aux_event(2), ev_out = aux_event(3);
  ZH_CHECK(hipMemsetAsync(cw.info_blk, 0, sizeof(int) * S.nblk, caller));
  ZH_CHECK(hipEventRecord(ev_in, caller));
  ZH_CHECK(hipStreamWaitEvent(sm, ev_in, 0));
  ZH_CHECK(hipStreamWaitEvent(sp, ev_in, 0));
  const int64_t kp = NB_MIX;
  // block column b: diagonal factor, inverse, L21, and L21's update operands
  auto panel = [&](int64_t j0, int64_t b) {
    const int64_t jb = std::min<int64_t>(NB_MIX, n - j0), m2 = n - j0 - jb;
    const int par = (int)(b & 1);
    float* Wd = S.W + j0 * n + j0;
    // the block as column-major lower (its transposed copy): that factorization's
    // triangular solves multiply by reciprocals instead of dividing
    hipLaunchKernelGGL(block_to_f64_kernel, dim3(blocks(jb * jb)), dim3(256), 0, sp, jb, (const float*)Wd, n,
                       cw.d64, jb, shift, 1);
    ZH_LAUNCH_CHECK();
    potrf_f64_cm(kLower, jb, cw.d64, jb, cw.info_blk + b, sp);
    float* invL = S.invL + b * NB_MIX * NB_MIX;
    float* invLT = S.invLT + b * NB_MIX * NB_MIX;
    hipLaunchKernelGGL(block_to_f32_kernel, dim3(blocks(jb * jb)), dim3(256), 0, sp, jb, (const double*)cw.d64, jb,
                       Wd, n, 1, 1);
    ZH_LAUNCH_CHECK();
    // the FP32 inverse of the rounded factor (as the LU panels invert U11), and its transpose
    tri_inverse_f32(jb, Wd, n, invL, 0, sp);
    hipLaunchKernelGGL(transpose_rect_f32_kernel, dim3(blocks(jb, 32), blocks(jb, 32)), dim3(32, 8), 0, sp, jb, jb,
                       (const float*)invL, jb, invLT, jb);
    ZH_LAUNCH_CHECK();
    if (m2 > 0) {
      float* A21 = S.W + (j0 + jb) * n + j0;
      ZH_CHECK(hipMemcpy2DAsync(cw.tmp, sizeof(float) * jb, A21, sizeof(float) * n, sizeof(float) * jb, m2,
                                hipMemcpyDeviceToDevice, sp));
      panel_product_f32(m2, jb, jb, 1.0f, cw.tmp, jb, invLT, jb, 0.0f, A21, n, sp);
      if (prec == kPrecF32) {
        hipLaunchKernelGGL(transpose_rect_f32_kernel, dim3(blocks(jb, 32), blocks(m2, 32)), dim3(32, 8), 0, sp, m2,
                           jb, (const float*)A21, n, cw.l21t[par], m2);
      } else if (prec == kPrecF16) {
        hipLaunchKernelGGL((rows_to_lowp_kernel<_Float16>), dim3(blocks(m2, 4)), dim3(256), 0, sp, m2, (int)jb,
                           (const float*)A21, n, cw.l16[par], kp, cw.rs[par]);
      } else {
        hipLaunchKernelGGL((rows_to_lowp_kernel<__bf16>), dim3(blocks(m2, 4)), dim3(256), 0, sp, m2, (int)jb,
                           (const float*)A21, n, reinterpret_cast<__bf16*>(cw.l16[par]), kp, cw.rs[par]);
      }
      ZH_LAUNCH_CHECK();
    }
    ZH_CHECK(hipEventRecord(ev_panel, sp));
  };
  // C (rows r0.., cols c0.., lower when `lower`) -= L21[r] L21[c]^T of block b
  auto update = [&](int64_t j0, int64_t b, int64_t r0, int64_t c0, int64_t rows, int64_t cols, int lower) {
    const int64_t jb = std::min<int64_t>(NB_MIX, n - j0), base = j0 + jb;
    const int par = (int)(b & 1);
    float* C = S.W + r0 * n + c0;
    const int64_t ro = r0 - base, co = c0 - base;   // offsets into the block's L21 rows
    void* tok = profile_begin(kProfGemmLowp, sm);
    if (prec == kPrecF32) {
      mp_gemm_f32(rows, cols, jb, -1.0f, S.W + r0 * n + j0, n, cw.l21t[par] + co, n - base, 1.0f, nullptr, nullptr,
                  C, n, sm, lower);
    } else if (prec == kPrecF16) {
      mp_gemm_f16(rows, cols, jb, -1.0f, cw.l16[par] + ro * kp, kp, cw.l16[par] + co * kp, kp, 1.0f,
                  cw.rs[par] + ro, cw.rs[par] + co, C, n, sm, lower);
    } else {
      const __bf16* l = reinterpret_cast<const __bf16*>(cw.l16[par]);
      mp_gemm_bf16(rows, cols, jb, -1.0f, l + ro * kp, kp, l + co * kp, kp, 1.0f, cw.rs[par] + ro, cw.rs[par] + co,
                   C, n, sm, lower);
    }
    const double fl = lower ? (double)rows * (rows + 1) * jb : 2.0 * rows * cols * jb;
    const double by = lower ? 4.0 * (double)rows * (rows + 1) + 2.0 * rows * jb : 8.0 * rows * cols + 2.0 * jb * (rows + cols);
    profile_end(tok, sm, fl, by);
  };
  panel(0, 0);
  for (int64_t j0 = 0, b = 0; j0 < n; j0 += NB_MIX, ++b) {
    const int64_t jb = std::min<int64_t>(NB_MIX, n - j0);
    const int64_t c1 = j0 + jb, m2 = n - c1;
    if (m2 <= 0) break;
    const int64_t jb2 = std::min<int64_t>(NB_MIX, m2);
    ZH_CHECK(hipStreamWaitEvent(sm, ev_panel, 0));
    // the next block column first (all its rows), then hand it to the panel stream
    update(j0, b, c1, c1, m2, jb2, 0);
    ZH_CHECK(hipEventRecord(ev_next, sm));
    // the rest of the lower triangle, beyond the next block column
    if (m2 > jb2) update(j0, b, c1 + jb2, c1 + jb2, m2 - jb2, m2 - jb2, 1);
    ZH_CHECK(hipStreamWaitEvent(sp, ev_next, 0));
    panel(c1, b + 1);
  }
  ZH_CHECK(hipEventRecord(ev_out, sp));
  ZH_CHECK(hipStreamWaitEvent(caller, ev_out, 0));
  ZH_CHECK(hipEventRecord(ev_out, sm));
  ZH_CHECK(hipStreamWaitEvent(caller, ev_out, 0));
}

// Y (n x nr, FP32 row-major, ld nr) <- L^-T L^-1 Y with the factor in S.W: two
// persistent launches of trisolve16_kernel
void chol_trsm_block(const Solver& S, int nr, float* Y) {
  if (nr < 1 || nr > MAXR) throw std::logic_error("chol_trsm_block: 1..16 right-hand sides");
  TriSolve16 f{};
  f.n = S.n;
  f.nblk = (S.n + T16 - 1) / T16;
  f.trans = 0;
  f.nr = nr;
  f.W = S.W;
  f.ldw = S.n;
  f.inv = S.invL;
  f.Y = Y;
  f.ldy = nr;
  f.xch = S.xch;
  f.err = S.serr;
  f.spin_limit = int64_t(1) << 22;
  const bool vec = (S.n & 3) == 0;
  for (int dir = 0; dir < 2; ++dir) {
    f.trans = dir;
    f.inv = dir ? S.invLT : S.invL;
    f.tag = ++S.tag;
    if (vec && dir) hipLaunchKernelGGL((trisolve16_kernel<true, true>), dim3(S.ts_grid), dim3(256), 0, S.s, f);
    else if (vec) hipLaunchKernelGGL((trisolve16_kernel<true, false>), dim3(S.ts_grid), dim3(256), 0, S.s, f);
    else if (dir) hipLaunchKernelGGL((trisolve16_kernel<false, true>), dim3(S.ts_grid), dim3(256), 0, S.s, f);
    else hipLaunchKernelGGL((trisolve16_kernel<false, false>), dim3(S.ts_grid), dim3(256), 0, S.s, f);
    ZH_LAUNCH_CHECK();
  }
}

// D (n x nr, column-major, ld ldd) (+)= Dg L^-T L^-1 Dg R
void chol_solve_block(const Solver& S, int nr, const double* R, int64_t ldr, double* D, int64_t ldd,
                      int accumulate) {
  const int64_t n = S.n;
  hipLaunchKernelGGL(pack_rows_kernel, dim3(blocks(n * nr)), dim3(256), 0, S.s, n, nr, R, ldr,
                     (const double*)S.r, S.ybuf);
  ZH_LAUNCH_CHECK();
  chol_trsm_block(S, nr, S.ybuf);
  hipLaunchKernelGGL(unpack_rows_kernel, dim3(blocks(n * nr)), dim3(256), 0, S.s, n, nr, (const float*)S.ybuf,
                     (const double*)S.r, D, ldd, accumulate);
  ZH_LAUNCH_CHECK();
}

void chol_precond(const Solver& S, const double* v, double* w, int accumulate) {
  chol_solve_block(S, 1, v, S.n, w, S.n, accumulate);
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
  S.precond = lu_solve;
  // ---- workspaces (slots 8..): the FP32 factors dominate (4 n^2 bytes)
  S.W = static_cast<float*>(workspace(sizeof(float) * (size_t)n * n, 8));
  workspace(kPanelSplitBytes, 21);   // split-K partial sums of the panel products
  char* small = static_cast<char*>(workspace(
      sizeof(float) * 2 * (size_t)S.nblk * NB_MIX * NB_MIX + sizeof(int32_t) * 2 * n + 64 +
          sizeof(double) * 3 * n + sizeof(float) * (n + NB_MIX) + sizeof(uint64_t) * (n + NB_MIX) + 64 +
          sizeof(double) * 512 + 4096, 9));
  S.invL = reinterpret_cast<float*>(small); small += sizeof(float) * S.nblk * NB_MIX * NB_MIX;
  S.invU = reinterpret_cast<float*>(small); small += sizeof(float) * S.nblk * NB_MIX * NB_MIX;
  S.ipiv = reinterpret_cast<int32_t*>(small); small += sizeof(int32_t) * n;
  S.perm = reinterpret_cast<int32_t*>(small); small += sizeof(int32_t) * n;
  S.info = reinterpret_cast<int*>(small); small += 64;
  S.r = reinterpret_cast<double*>(small); small += sizeof(double) * n;
  S.c = reinterpret_cast<double*>(small); small += sizeof(double) * n;
  S.rowsum = reinterpret_cast<double*>(small); small += sizeof(double) * n;
  S.z = reinterpret_cast<float*>(small); small += sizeof(float) * (n + NB_MIX);
  S.xch = reinterpret_cast<uint64_t*>(small); small += sizeof(uint64_t) * (n + NB_MIX);
  S.serr = reinterpret_cast<int*>(small); small += 64;
  S.part = reinterpret_cast<double*>(small);
  // tags start again from 1 on cleared exchange words
  ZH_CHECK(hipMemsetAsync(S.xch, 0, sizeof(uint64_t) * (n + NB_MIX) + 64, s));
  {
    int dev = 0, cus = 0, per_cu = 0;
    ZH_CHECK(hipGetDevice(&dev));
    ZH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trisolve_f32_kernel, 256, 0));
    const int64_t nts = (n + TS - 1) / TS;
    S.ts_grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nts, (int64_t)cus * std::max(per_cu, 1)));
  }
  const int64_t kp = NB_MIX;
  char* lp = static_cast<char*>(workspace(sizeof(_Float16) * 2 * (size_t)(n + NB_MIX) * kp +
                                              sizeof(float) * 2 * (n + NB_MIX) + 4096, 10));
  tl_mix.prec = opts.precision;
  tl_mix.l16_of = nullptr;
  tl_mix.block_pairs_j0[0] = tl_mix.block_pairs_j0[1] = -1;
  tl_mix.l16 = reinterpret_cast<_Float16*>(lp); lp += sizeof(_Float16) * (n + NB_MIX) * kp;
  tl_mix.u16 = reinterpret_cast<_Float16*>(lp); lp += sizeof(_Float16) * (n + NB_MIX) * kp;
  tl_mix.rs = reinterpret_cast<float*>(lp); lp += sizeof(float) * (n + NB_MIX);
  tl_mix.cs = reinterpret_cast<float*>(lp);

  // ---- 1. equilibration and the FP32 working matrix
  // r_i = 2^-floor(log2 max_j |a_ij|), then c_j = 2^-floor(log2 max_i |r_i a_ij|)
  ZH_CHECK(hipMemsetAsync(S.r, 0, sizeof(double) * 2 * n, s));   // r, c hold the maxima first
  double* rmax = S.r;
  double* cmax = S.c;
  const int64_t split = std::max<int64_t>(1, std::min<int64_t>(256, n / 64)), chunk = (n + split - 1) / split;
  if (layout == kRowMajor) {
    // row maxima and ||A||_inf of the original matrix in one pass
    hipLaunchKernelGGL(line_absmax_kernel, dim3(blocks(n, 4)), dim3(256), 0, s, n, A, lda,
                       (const double*)nullptr, rmax, S.rowsum);
    ZH_LAUNCH_CHECK();
    S.anorm = reduce_host(S, n, S.rowsum, nullptr, 1);
  } else {
    hipLaunchKernelGGL(cross_absmax_kernel, dim3(blocks(n), (unsigned)split), dim3(256), 0, s, n, A, lda,
                       (const double*)nullptr, rmax, chunk);
  }
  hipLaunchKernelGGL(to_scale_kernel, dim3(blocks(n)), dim3(256), 0, s, n, rmax, S.rowsum);   // rowsum <- r
  if (layout == kRowMajor)
    hipLaunchKernelGGL(cross_absmax_kernel, dim3(blocks(n), (unsigned)split), dim3(256), 0, s, n, A, lda,
                       (const double*)S.rowsum, cmax, chunk);
  else
    hipLaunchKernelGGL(line_absmax_kernel, dim3(blocks(n, 4)), dim3(256), 0, s, n, A, lda,
                       (const double*)S.rowsum, cmax);
  ZH_CHECK(hipMemcpyAsync(S.r, S.rowsum, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(to_scale_kernel, dim3(blocks(n)), dim3(256), 0, s, n, cmax, S.c);
  ZH_LAUNCH_CHECK();
  hipLaunchKernelGGL(equilibrate_kernel, dim3(blocks(n, 32), blocks(n, 32)), dim3(32, 8), 0, s, layout, n,
                     A, lda, S.r, S.c, S.W);
  ZH_LAUNCH_CHECK();
  if (layout == kColMajor) {   // ||A||_inf of the original matrix
    hipLaunchKernelGGL(rowsum_abs_kernel, dim3(blocks(n)), dim3(256), 0, s, layout, n, A, lda, S.rowsum);
    ZH_LAUNCH_CHECK();
    S.anorm = reduce_host(S, n, S.rowsum, nullptr, 1);
  }

  // ---- 2. factorization
  ZH_CHECK(hipMemsetAsync(S.info, 0, sizeof(int), s));
  struct Inv {
    static void store(int64_t kb, const float* L11, int64_t ld, float* inv, hipStream_t st) {
      if (inv == tl_mix.inv_ready) return;   // formed by the diagonal-block panel
      invert_f32(kb, L11, ld, inv, st);
    }
  };
  {
    char* pw = static_cast<char*>(workspace(
        sizeof(float) * ((size_t)(n + NB_MIX) * (NB_MIX + DB) + 2 * DB * DB) + 256, 15));
    tl_mix.pbak = reinterpret_cast<float*>(pw); pw += sizeof(float) * (size_t)(n + NB_MIX) * NB_MIX;
    tl_mix.ptmp = reinterpret_cast<float*>(pw); pw += sizeof(float) * (size_t)(n + NB_MIX) * DB;
    tl_mix.scr3 = reinterpret_cast<float*>(pw); pw += sizeof(float) * DB * DB;
    tl_mix.scr4 = reinterpret_cast<float*>(pw); pw += sizeof(float) * DB * DB;
    tl_mix.pfail = reinterpret_cast<int*>(pw);
    tl_mix.panels = tl_mix.panels_fast = 0;
    tl_mix.inv_all = S.invL;
    tl_mix.invu_all = S.invU;
    tl_mix.have_invu.assign(S.nblk, 0);
    tl_mix.optimistic = opts.no_optimistic == 0;
    tl_mix.tau = opts.pivot_threshold > 0.0 ? (float)opts.pivot_threshold : 16.0f;
  }
  struct InvAt {
    static float* at(int64_t j0) { return tl_mix.inv_all + (j0 / NB_MIX) * NB_MIX * NB_MIX; }
  };
  const LuOps<float> ops{trailing_gemm, trsm_f32, Inv::store, apply_inv_f32, panel_gemm_f32, try_panel_f32,
                         InvAt::at, 8, block_perm_f32, finish_panel_f32};
  getrf_lookahead<float>(n, n, S.W, n, S.ipiv, S.info, NB_MIX, ops, s);
  int info = 0;
  ZH_CHECK(hipMemcpyAsync(&info, S.info, sizeof(int), hipMemcpyDeviceToHost, s));
  ZH_CHECK(hipStreamSynchronize(s));
  if (info == kErrSync) throw SyncError("dgesv_mixed: a panel's cross-workgroup wait gave up");
  res->info = info;
  int ret = 0;   // first zero pivot of an FP64 fallback factorization (A singular)
  res->panels = tl_mix.panels;
  res->panels_optimistic = tl_mix.panels_fast;

  // inverses of the diagonal blocks of L and of U, row-major (for the triangular solves)
  float* scratch = static_cast<float*>(workspace(sizeof(float) * 2 * NB_MIX * NB_MIX, 11));
  float* scratch2 = scratch + NB_MIX * NB_MIX;
  // (L11^-1 of every block and U11^-1 of the diagonal-block panels are already there)
  for (int64_t b = 0; b < S.nblk; ++b) {
    const int64_t b0 = b * NB_MIX, nb = std::min<int64_t>(NB_MIX, n - b0);
    if (tl_mix.have_invu[b]) continue;
    // inv(U) = (inv(U^T))^T, U^T being lower
    hipLaunchKernelGGL(transpose_f32_kernel, dim3(blocks(nb * nb)), dim3(256), 0, s, nb, S.W + b0 * n + b0, n,
                       scratch, 1);
    tri_inverse_f32(nb, scratch, nb, scratch2, 0, s);
    hipLaunchKernelGGL(transpose_f32_kernel, dim3(blocks(nb * nb)), dim3(256), 0, s, nb, scratch2, nb,
                       S.invU + b * NB_MIX * NB_MIX, 0);
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
      int h64 = 0;
      ZH_CHECK(hipMemcpyAsync(&h64, inf64, sizeof(int), hipMemcpyDeviceToHost, s));
      ZH_CHECK(hipStreamSynchronize(s));
      if (h64 == kErrSync) throw SyncError("dgesv_mixed: a panel's cross-workgroup wait gave up");
      if (h64 > 0 && ret == 0) ret = h64;
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
    if (be > worst || be != be) worst = be;   // a NaN backward error stays visible
    hipLaunchKernelGGL(store_f64_kernel, dim3(blocks(n)), dim3(256), 0, s, n, xj, xcol, xstride);
    ZH_LAUNCH_CHECK();
  }
  ZH_CHECK(hipStreamSynchronize(s));
  res->backward_error = worst;
  return ret;
}


// Mixed-precision Cholesky solve of an SPD system (no counterpart in zml; the
// FP64 interface it sits behind is dposv's, posv.zig:15-70):
//   1. S = the full symmetric FP64 matrix from the caller's triangle (used by
//      every residual), d_i = 2^-floor(log2 sqrt(s_ii)), W = fl32(D S D);
//   2. W = L L^T (chol_factor). A non-positive pivot of the low-precision
//      factorization is answered by factoring again with a diagonal shift
//      (W + sigma I, sigma growing by 16 from 2^-10 on the unit-scale diagonal);
//      the refinement then works with that preconditioner;
//   3. the right-hand sides in groups of 16: X = D L^-T L^-1 D B, then FP64
//      refinement of the whole group (R = B - S X as one FP64 GEMM) until
//      every column meets the target; a column whose backward error stops
//      halving, or grows, goes to GMRES-IR, and one that GMRES-IR cannot
//      finish is solved with the FP64 Cholesky factorization of S.
int dposv_mixed(int layout, int uplo, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                const double* B, int64_t ldb, double* X, int64_t ldx, const zmlhip_ir_opts* o,
                zmlhip_ir_result* res, hipStream_t s) {
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
  S.layout = kRowMajor;   // S.A is the full symmetric copy
  S.n = n;
  S.s = s;
  S.nblk = (n + NB_MIX - 1) / NB_MIX;
  S.precond = chol_precond;
  double* S64 = static_cast<double*>(workspace(sizeof(double) * (size_t)n * n, 16));
  S.A = S64;
  S.lda = n;
  S.W = static_cast<float*>(workspace(sizeof(float) * (size_t)n * n, 8));
  workspace(kPanelSplitBytes, 21);   // split-K partial sums of the panel products
  const size_t blk2 = (size_t)NB_MIX * NB_MIX;
  const SkinnySplit ssp = skinny_split(n);
  char* small = static_cast<char*>(workspace(
      sizeof(float) * 2 * S.nblk * blk2 + sizeof(double) * 2 * n + sizeof(double) * 512 +
          sizeof(float) * (n + NB_MIX) * MAXR + sizeof(uint64_t) * (n + NB_MIX) * MAXR +
          sizeof(double) * MAXR * n * ssp.splits + sizeof(double) * blk2 + sizeof(int) * (S.nblk + 64) +
          sizeof(double) * 2 * MAXR + 16 * 256, 9));
  auto carve = [&](size_t bytes) { char* p = small; small += (bytes + 255) / 256 * 256; return p; };
  S.invL = reinterpret_cast<float*>(carve(sizeof(float) * S.nblk * blk2));
  S.invLT = reinterpret_cast<float*>(carve(sizeof(float) * S.nblk * blk2));
  S.r = reinterpret_cast<double*>(carve(sizeof(double) * n));
  S.c = S.r;
  S.rowsum = reinterpret_cast<double*>(carve(sizeof(double) * n));
  S.part = reinterpret_cast<double*>(carve(sizeof(double) * 512));
  S.ybuf = reinterpret_cast<float*>(carve(sizeof(float) * (n + NB_MIX) * MAXR));
  S.xch = reinterpret_cast<uint64_t*>(carve(sizeof(uint64_t) * (n + NB_MIX) * MAXR));
  S.serr = reinterpret_cast<int*>(carve(64));
  S.pbuf = reinterpret_cast<double*>(carve(sizeof(double) * MAXR * n * ssp.splits));
  // tags start again from 1 on cleared exchange words
  ZH_CHECK(hipMemsetAsync(S.xch, 0, sizeof(uint64_t) * (n + NB_MIX) * MAXR, s));
  ZH_CHECK(hipMemsetAsync(S.serr, 0, 64, s));
  {
    int dev = 0, cus = 0, per_cu = 0;
    ZH_CHECK(hipGetDevice(&dev));
    ZH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    per_cu = 1 << 20;
    for (const void* k : {(const void*)trisolve16_kernel<true, false>, (const void*)trisolve16_kernel<true, true>,
                          (const void*)trisolve16_kernel<false, false>, (const void*)trisolve16_kernel<false, true>}) {
      int pc = 0;
      ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, k, 256, 0));
      per_cu = std::min(per_cu, pc);
    }
    const int64_t nts = (n + T16 - 1) / T16;
    S.ts_grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nts, (int64_t)cus * std::max(per_cu, 1)));
  }
  CholWork cw;
  cw.d64 = reinterpret_cast<double*>(carve(sizeof(double) * blk2));
  cw.info_blk = reinterpret_cast<int*>(carve(sizeof(int) * (S.nblk + 64)));
  double* cmax = reinterpret_cast<double*>(carve(sizeof(double) * 2 * MAXR));
  char* lp = static_cast<char*>(workspace(sizeof(float) * (size_t)(n + NB_MIX) * NB_MIX * 3 +
                                              sizeof(float) * 2 * (n + NB_MIX) + 4096, 10));
  cw.tmp = reinterpret_cast<float*>(lp); lp += sizeof(float) * (size_t)(n + NB_MIX) * NB_MIX;
  for (int q = 0; q < 2; ++q) {   // the FP32 update uses l21t, the 16-bit ones l16 + rs
    cw.l21t[q] = reinterpret_cast<float*>(lp);
    cw.l16[q] = reinterpret_cast<_Float16*>(lp);
    lp += sizeof(float) * (size_t)(n + NB_MIX) * NB_MIX;
    cw.rs[q] = reinterpret_cast<float*>(lp);
    lp += sizeof(float) * (n + NB_MIX);
  }

  // ---- 1. symmetric FP64 copy, norm, scaling, FP32 working matrix
  hipLaunchKernelGGL(symmetrize_kernel, dim3(blocks(n, 32), blocks(n, 32)), dim3(32, 8), 0, s, layout, uplo, n,
                     A, lda, S64);
  hipLaunchKernelGGL(rowsum_abs_kernel, dim3((unsigned)n), dim3(256), 0, s, (int)kRowMajor, n,
                     (const double*)S64, n, S.rowsum);
  ZH_LAUNCH_CHECK();
  S.anorm = reduce_host(S, n, S.rowsum, nullptr, 1);
  hipLaunchKernelGGL(diag_scale_kernel, dim3(blocks(n)), dim3(256), 0, s, n, (const double*)S64, S.r);
  ZH_LAUNCH_CHECK();

  // ---- 2. factorization (shifted again while a pivot fails)
  std::vector<int> ib(S.nblk);
  double shift = 0.0;
  for (int attempt = 0;; ++attempt) {
    hipLaunchKernelGGL(equilibrate_kernel, dim3(blocks(n, 32), blocks(n, 32)), dim3(32, 8), 0, s, (int)kRowMajor,
                       n, (const double*)S64, n, (const double*)S.r, (const double*)S.r, S.W);
    ZH_LAUNCH_CHECK();
    chol_factor(S, cw, opts.precision, shift);
    ZH_CHECK(hipMemcpyAsync(ib.data(), cw.info_blk, sizeof(int) * S.nblk, hipMemcpyDeviceToHost, s));
    ZH_CHECK(hipStreamSynchronize(s));
    int first = 0;
    for (int64_t b = 0; b < S.nblk && !first; ++b)
      if (ib[b]) first = (int)(b * NB_MIX + ib[b]);
    if (attempt == 0) res->info = first;
    if (!first) break;
    if (attempt >= 4) return res->info;   // not positive definite: LAPACK-style info > 0, X untouched
    shift = shift == 0.0 ? 0x1p-10 : shift * 16.0;
    res->fallback = std::max(res->fallback, 1);
  }
  ZH_LAUNCH_CHECK();

  // ---- 3./4. groups of up to 16 right-hand sides
  double* vec = static_cast<double*>(workspace(sizeof(double) * 4 * (size_t)n * MAXR, 12));
  double* Bc = vec;
  double* Xc = vec + (size_t)n * MAXR;
  double* Rc = vec + 2 * (size_t)n * MAXR;
  double* gw = nullptr;
  double worst = 0.0;
  int ret = 0;   // first non-positive pivot of an FP64 fallback factorization
  std::vector<double> hm(2 * MAXR);
  for (int64_t q0 = 0; q0 < nrhs; q0 += MAXR) {
    const int nr = (int)std::min<int64_t>(MAXR, nrhs - q0);
    // B, X column-major n x nr (ld n)
    for (int q = 0; q < nr; ++q) {
      const int64_t col = q0 + q;
      const int64_t bstride = layout == kColMajor ? 1 : ldb;
      const double* bcol = layout == kColMajor ? B + col * ldb : B + col;
      hipLaunchKernelGGL(copy_f64_kernel, dim3(blocks(n)), dim3(256), 0, s, n, bcol, bstride, Bc + q * n);
    }
    ZH_LAUNCH_CHECK();
    chol_solve_block(S, nr, Bc, n, Xc, n, 0);
    std::vector<double> prev(nr, INFINITY), be(nr, INFINITY);
    std::vector<int> state(nr, 0);   // 0 refining, 1 converged, 2 needs GMRES-IR
    int it = 0;
    for (;; ++it) {
      // R = B - S X (FP64), per-column backward errors
      ZH_CHECK(hipMemcpyAsync(Rc, Bc, sizeof(double) * n * nr, hipMemcpyDeviceToDevice, s));
      skinny_gemm_f64(S, nr, -1.0, S64, n, Xc, n, 1.0, Rc, n, S.pbuf);
      ZH_CHECK(hipMemsetAsync(cmax, 0, sizeof(double) * 2 * MAXR, s));
      hipLaunchKernelGGL(colmax_kernel, dim3(64, nr), dim3(256), 0, s, n, (const double*)Rc, n, cmax);
      hipLaunchKernelGGL(colmax_kernel, dim3(64, nr), dim3(256), 0, s, n, (const double*)Xc, n, cmax + MAXR);
      ZH_LAUNCH_CHECK();
      ZH_CHECK(hipMemcpyAsync(hm.data(), cmax, sizeof(double) * 2 * MAXR, hipMemcpyDeviceToHost, s));
      ZH_CHECK(hipStreamSynchronize(s));
      check_solve(S);
      bool active = false;
      for (int q = 0; q < nr; ++q) {
        if (state[q] != 0) continue;
        be[q] = hm[q] / (S.anorm * hm[MAXR + q] * (double)n * 0x1p-53);
        if (be[q] <= opts.tol) state[q] = 1;
        else if (!(be[q] == be[q]) || be[q] > 2.0 * prev[q] || (it >= 2 && be[q] > 0.5 * prev[q]) ||
                 it >= opts.max_iter)
          state[q] = 2;
        else active = true;
        prev[q] = be[q];
      }
      if (!active) break;
      // X += M^-1 R on the columns still refining (the whole group is solved at once)
      unsigned mask = 0;
      for (int q = 0; q < nr; ++q)
        if (state[q] == 0) mask |= 1u << q;
      hipLaunchKernelGGL(pack_rows_kernel, dim3(blocks(n * nr)), dim3(256), 0, s, n, nr, (const double*)Rc, n,
                         (const double*)S.r, S.ybuf);
      ZH_LAUNCH_CHECK();
      chol_trsm_block(S, nr, S.ybuf);
      hipLaunchKernelGGL(unpack_add_masked_kernel, dim3(blocks(n * nr)), dim3(256), 0, s, n, nr,
                         (const float*)S.ybuf, (const double*)S.r, Xc, n, mask);
      ZH_LAUNCH_CHECK();
    }
    res->iterations = std::max(res->iterations, it);
    for (int q = 0; q < nr; ++q) {
      double* xq = Xc + q * n;
      const double* bq = Bc + q * n;
      if (state[q] == 2 && !opts.no_gmres) {
        if (!gw) gw = static_cast<double*>(workspace(sizeof(double) * (size_t)(2 * opts.gmres_restart + 3) * n, 13));
        double be2 = 0.0;
        const int gi = gmres_ir(S, bq, xq, opts.gmres_restart, opts.gmres_max_iter, opts.tol, &be2, gw);
        res->gmres_iterations = std::max(res->gmres_iterations, gi);
        res->fallback = std::max(res->fallback, 1);
        be[q] = be2;
        if (be2 <= opts.tol) state[q] = 1;
      }
      if (state[q] != 1) {
        // FP64 Cholesky of S (symmetric: the column-major view is S itself)
        res->fallback = 2;
        double* A64 = static_cast<double*>(workspace(sizeof(double) * (size_t)n * n, 14));
        int* inf64 = cw.info_blk + S.nblk;
        ZH_CHECK(hipMemcpyAsync(A64, S64, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
        ZH_CHECK(hipMemsetAsync(inf64, 0, sizeof(int), s));
        potrf_f64_cm(kLower, n, A64, n, inf64, s);
        int h64 = 0;
        ZH_CHECK(hipMemcpyAsync(&h64, inf64, sizeof(int), hipMemcpyDeviceToHost, s));
        ZH_CHECK(hipStreamSynchronize(s));
        if (h64 > 0 && ret == 0) ret = h64;
        ZH_CHECK(hipMemcpyAsync(xq, bq, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        trsm_f64_cm(kLeft, kLower, kNoTrans, kNonUnit, n, 1, 1.0, A64, n, xq, n, s);
        trsm_f64_cm(kLeft, kLower, kTrans, kNonUnit, n, 1, 1.0, A64, n, xq, n, s);
        double* rq = Rc + q * n;
        residual(S, bq, xq, rq);
        be[q] = backward_error(S, rq, xq);
      }
      if (be[q] > worst || be[q] != be[q]) worst = be[q];
      const int64_t col = q0 + q;
      const int64_t xstride = layout == kColMajor ? 1 : ldx;
      double* xcol = layout == kColMajor ? X + col * ldx : X + col;
      hipLaunchKernelGGL(store_f64_kernel, dim3(blocks(n)), dim3(256), 0, s, n, (const double*)xq, xcol, xstride);
    }
    ZH_LAUNCH_CHECK();
  }
  ZH_CHECK(hipStreamSynchronize(s));
  res->backward_error = worst;
  return ret;
}

}  // namespace zmlhip
