// FP64 / FP32 Cholesky (zml lapack/potrf.zig:17-240, potrf2.zig:17-217), column-major view.
//
// zml's potrf is left-looking with NB = 64; on the GPU the same factor is
// computed right-looking (the per-element operation order of the lower,
// column-major case is the unblocked right-looking one either way, see the
// oracle's closed form), blocked on two levels: for each 512-column block
// column, in 128-column steps
//   factor the diagonal block in one workgroup (registers, separately rounded
//   multiply/subtract, l_ij = a_ij * fl(1/l_jj) as zml's right-side trsm
//   does, correctly rounded sqrt),
//   TRSM the rows below (lower) / columns right of it (upper),
//   update the rest of the block column only;
// then SYRK the trailing triangle with K = 512 on the MFMA GEMM, the next
// block column first so that its factorization overlaps the rest.
// Row-major lower is column-major upper of the same memory (and vice versa),
// which is how row-major calls are mapped.
//
// info (device int, zero on entry) receives the first non-positive or NaN
// pivot (1-based) as zml returns it (potrf2.zig:37-46, potrf.zig:100-103).
// Columns at and after a failing pivot are left unspecified.
#include "zmlhip_internal.hpp"

#include <algorithm>

namespace zmlhip {
namespace {

constexpr int DB = 128;   // diagonal block of the driver

// Unblocked Cholesky of an nb x nb block (nb <= DB), register resident.
// lower: A = L L^T on the lower triangle; else A = U^T U on the upper one.
// Thread (tx, ty) of a 16 x 16 grid holds the elements (tx + 16 ii, ty + 16 cc),
// ii >= cc, of the lower-orientation block in registers. Column j travels
// through LDS unscaled (double buffered, one barrier per column); every thread
// takes l_jj = sqrt(a_jj), r = fl(1 / l_jj) and l_ij = fl(a_ij * r) from it, so
// the trailing update a_ic - l_ij * l_cj uses the same separately rounded
// values, in the same right-looking order, as the unblocked algorithm.
__device__ __forceinline__ double sqrt_rn(double x) { return sqrt(x); }   // IEEE, correctly rounded
__device__ __forceinline__ float sqrt_rn(float x) { return sqrtf(x); }

template <typename T>
__global__ void __launch_bounds__(256)
potrf_diag_kernel(int nb, T* __restrict__ a, int64_t lda, int lower, int* __restrict__ info,
                  int64_t info_base) {
#pragma clang fp contract(off)
  __shared__ T colbuf[2][DB];
  if (*info != 0) return;   // an earlier block already failed
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  T v[8][8];
  // every address of the nb x nb block is readable: all 36 loads are issued
  // unconditionally (clamped indices; the empty asm keeps the compiler from
  // sinking each one under its mask and waiting for it), then masked
#pragma unroll
  for (int ii = 0; ii < 8; ++ii)
#pragma unroll
    for (int cc = 0; cc <= ii; ++cc) {
      const int r = tx + 16 * ii, c = ty + 16 * cc;
      const int64_t rc = r < nb ? r : nb - 1, cl = c < nb ? c : nb - 1;
      v[ii][cc] = a[lower ? rc + cl * lda : cl + rc * lda];
    }
#pragma unroll
  for (int ii = 0; ii < 8; ++ii)
#pragma unroll
    for (int cc = 0; cc <= ii; ++cc) {
      asm volatile("" : "+v"(v[ii][cc]));
      const int r = tx + 16 * ii, c = ty + 16 * cc;
      if (!(r < nb && c <= r)) v[ii][cc] = T(0);
    }
  // column 0 (ty == 0, cc == 0)
  if (ty == 0) {
#pragma unroll
    for (int ii = 0; ii < 8; ++ii)
      if (tx + 16 * ii < nb) colbuf[0][tx + 16 * ii] = v[ii][0];
  }
  __syncthreads();
  int fail = 0;
  for (int j = 0; j < nb; ++j) {
    const T* cur = colbuf[j & 1];
    const T d = cur[j];
    if (!(d > T(0))) { fail = j + 1; break; }   // non-positive or NaN pivot
    const T ljj = sqrt_rn(d);
    const T rj = T(1) / ljj;
    T lr[8], lc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      lr[k] = cur[tx + 16 * k] * rj;
      lc[k] = cur[ty + 16 * k] * rj;
    }
    const int jc = j >> 4, jt = j & 15;
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) {
      if (cc == jc && ty == jt) {   // this thread holds part of column j: finalize it
#pragma unroll
        for (int ii = cc; ii < 8; ++ii) {
          const int r = tx + 16 * ii;
          if (r == j) v[ii][cc] = ljj;
          else if (r > j) v[ii][cc] = lr[ii];
        }
      }
      if (16 * cc + 15 > j) {   // some column of this group lies right of j
        const int c = ty + 16 * cc;
#pragma unroll
        for (int ii = cc; ii < 8; ++ii) {
          const int r = tx + 16 * ii;
          if (c > j && c <= r && r < nb) v[ii][cc] = v[ii][cc] - lr[ii] * lc[cc];
        }
      }
    }
    // publish column j + 1, now fully updated
    const int jn = j + 1;
    if (jn < nb && ty == (jn & 15)) {
#pragma unroll
      for (int cc = 0; cc < 8; ++cc)
        if (cc == (jn >> 4)) {
#pragma unroll
          for (int ii = cc; ii < 8; ++ii) {
            const int r = tx + 16 * ii;
            if (r >= jn && r < nb) colbuf[jn & 1][r] = v[ii][cc];
          }
        }
    }
    __syncthreads();
  }
  const int done = fail ? fail - 1 : nb;   // columns that were completed
#pragma unroll
  for (int ii = 0; ii < 8; ++ii)
#pragma unroll
    for (int cc = 0; cc <= ii; ++cc) {
      const int r = tx + 16 * ii, c = ty + 16 * cc;
      if (r < nb && c <= r && c < done) a[lower ? r + (int64_t)c * lda : c + (int64_t)r * lda] = v[ii][cc];
    }
  if (fail && t == 0) *info = (int)(info_base + fail);
}

}  // namespace

template <typename T>
void potrf_cm(int uplo, int64_t n, T* a, int64_t lda, int* info, hipStream_t s) {
  if (n <= 0) return;
  const bool lower = uplo == kLower;
  constexpr int64_t OB = 512;   // block column of the trailing updates (their K)
  auto at = [&](int64_t i, int64_t j) { return a + i + j * lda; };
  // factor block column [j0, j0 + jb) in DB-wide steps whose updates stay inside it
  auto factor_block = [&](int64_t j0, int64_t jb, hipStream_t st) {
    for (int64_t k0 = j0; k0 < j0 + jb; k0 += DB) {
      const int64_t kb = std::min<int64_t>(DB, j0 + jb - k0);
      T* d = at(k0, k0);
      hipLaunchKernelGGL(potrf_diag_kernel<T>, dim3(1), dim3(256), 0, st, (int)kb, d, lda, lower ? 1 : 0, info, k0);
      ZH_LAUNCH_CHECK();
      const int64_t r = n - k0 - kb;           // rows below (lower) / columns right (upper)
      if (r <= 0) break;
      const int64_t w = j0 + jb - (k0 + kb);   // columns of the block column still to update
      if (lower) {
        // A21 <- A21 L11^-T ; the block column's part of A22 <- A22 - A21 A21^T
        T* a21 = at(k0 + kb, k0);
        trsm_cm<T>(kRight, kLower, kTrans, kNonUnit, r, kb, T(1), d, lda, a21, lda, st);
        if (w > 0)
          gemm_cm<T>(kNoTrans, kTrans, r, w, kb, T(-1), a21, lda, a21, lda, T(1), at(k0 + kb, k0 + kb), lda, st,
                     kLower);
      } else {
        // A12 <- U11^-T A12 ; the block row's part of A22 <- A22 - A12^T A12
        T* a12 = at(k0, k0 + kb);
        trsm_cm<T>(kLeft, kUpper, kTrans, kNonUnit, kb, r, T(1), d, lda, a12, lda, st);
        if (w > 0)
          gemm_cm<T>(kTrans, kNoTrans, w, r, kb, T(-1), a12, lda, a12, lda, T(1), at(k0 + kb, k0 + kb), lda, st,
                     kUpper);
      }
    }
  };
  if (n <= OB) {
    factor_block(0, n, s);
    return;
  }
  // Look-ahead of depth one: after block column j0 is factored, the caller's
  // stream first updates the next block column and hands it to a panel stream,
  // then applies the rest of the trailing update while the panel stream
  // factors the next block column. Every update of a region stays ordered on
  // the caller's stream; the panel stream touches only its own block column.
  const hipStream_t sp = aux_stream(0);
  const hipEvent_t ev_in = aux_event(0), ev_panel = aux_event(1), ev_next = aux_event(2);
  ZH_CHECK(hipEventRecord(ev_in, s));
  ZH_CHECK(hipStreamWaitEvent(sp, ev_in, 0));
  factor_block(0, std::min<int64_t>(OB, n), sp);
  ZH_CHECK(hipEventRecord(ev_panel, sp));
  for (int64_t j0 = 0; j0 < n; j0 += OB) {
    const int64_t jb = std::min<int64_t>(OB, n - j0);
    const int64_t r = n - j0 - jb;
    if (r <= 0) break;
    const int64_t jb2 = std::min<int64_t>(OB, r), r2 = r - jb2;
    const int64_t c1 = j0 + jb, c2 = c1 + jb2;
    ZH_CHECK(hipStreamWaitEvent(s, ev_panel, 0));
    if (lower) {
      const T* l21 = at(c1, j0);
      gemm_cm<T>(kNoTrans, kTrans, r, jb2, jb, T(-1), l21, lda, l21, lda, T(1), at(c1, c1), lda, s, kLower);
      ZH_CHECK(hipEventRecord(ev_next, s));
      if (r2 > 0) syrk_cm<T>(kLower, kNoTrans, r2, jb, T(-1), at(c2, j0), lda, T(1), at(c2, c2), lda, s);
    } else {
      const T* u12 = at(j0, c1);
      gemm_cm<T>(kTrans, kNoTrans, jb2, r, jb, T(-1), u12, lda, u12, lda, T(1), at(c1, c1), lda, s, kUpper);
      ZH_CHECK(hipEventRecord(ev_next, s));
      if (r2 > 0) syrk_cm<T>(kUpper, kTrans, r2, jb, T(-1), at(j0, c2), lda, T(1), at(c2, c2), lda, s);
    }
    ZH_CHECK(hipStreamWaitEvent(sp, ev_next, 0));
    factor_block(c1, jb2, sp);
    ZH_CHECK(hipEventRecord(ev_panel, sp));
  }
  ZH_CHECK(hipStreamWaitEvent(s, ev_panel, 0));
}
template void potrf_cm<double>(int, int64_t, double*, int64_t, int*, hipStream_t);
template void potrf_cm<float>(int, int64_t, float*, int64_t, int*, hipStream_t);

void potrf_f64_cm(int uplo, int64_t n, double* a, int64_t lda, int* info, hipStream_t s) {
  potrf_cm<double>(uplo, n, a, lda, info, s);
}

}  // namespace zmlhip
