// FP64 Cholesky (zml lapack/potrf.zig:17-240, potrf2.zig:17-217), column-major view.
//
// zml's potrf is left-looking with NB = 64; on the GPU the same factor is
// computed right-looking (the per-element operation order of the lower,
// column-major case is the unblocked right-looking one either way, see the
// oracle's closed form): for each 128-column block
//   factor the diagonal block in one workgroup (LDS, separately rounded
//   multiply/subtract, l_ij = a_ij * fl(1/l_jj) as zml's right-side trsm
//   does, correctly rounded sqrt),
//   TRSM the block column below (lower) / block row right of it (upper),
//   SYRK the trailing triangle on the MFMA GEMM.
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

// Unblocked Cholesky of an nb x nb block (nb <= DB) held in LDS.
// lower: A = L L^T on the lower triangle; else A = U^T U on the upper one.
__global__ void __launch_bounds__(256)
potrf_diag_kernel(int nb, double* __restrict__ a, int64_t lda, int lower, int* __restrict__ info,
                  int64_t info_base) {
#pragma clang fp contract(off)
  __shared__ double S[DB * (DB + 1)];
  __shared__ int s_fail;
  if (*info != 0) return;   // an earlier block already failed
  const int t = threadIdx.x;
  const int ls = DB + 1;
  // S holds the block in "lower" orientation: S[r*ls + c] = L-view element (r, c), r >= c
  for (int e = t; e < nb * nb; e += blockDim.x) {
    const int r = e % nb, c = e / nb;
    if (r >= c) S[r * ls + c] = lower ? a[r + (int64_t)c * lda] : a[c + (int64_t)r * lda];
  }
  if (t == 0) s_fail = 0;
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (t == 0) {
      const double ajj = S[j * ls + j];
      if (ajj <= 0.0 || ajj != ajj) s_fail = j + 1;
      else S[j * ls + j] = sqrt(ajj);
    }
    __syncthreads();
    if (s_fail) break;
    const double r = 1.0 / S[j * ls + j];
    for (int i = j + 1 + t; i < nb; i += blockDim.x) S[i * ls + j] = S[i * ls + j] * r;
    __syncthreads();
    // trailing lower triangle: S[i][c] -= S[i][j] * S[c][j], j < c <= i; a 16 x 16
    // thread grid walks it (rows by ty, columns by tx) without index division
    {
      const int tx = t & 15, ty = t >> 4;
      for (int i = j + 1 + ty; i < nb; i += 16) {
        const double lij = S[i * ls + j];
        for (int c = j + 1 + tx; c <= i; c += 16) S[i * ls + c] = S[i * ls + c] - lij * S[c * ls + j];
      }
    }
    __syncthreads();
  }
  const int fail = s_fail;
  const int done = fail ? fail - 1 : nb;   // columns that were completed
  for (int e = t; e < nb * nb; e += blockDim.x) {
    const int r = e % nb, c = e / nb;
    if (r >= c && c < done) {
      if (lower) a[r + (int64_t)c * lda] = S[r * ls + c];
      else a[c + (int64_t)r * lda] = S[r * ls + c];
    }
  }
  if (fail && t == 0) *info = (int)(info_base + fail);
}

}  // namespace

void potrf_f64_cm(int uplo, int64_t n, double* a, int64_t lda, int* info, hipStream_t s) {
  const bool lower = uplo == kLower;
  for (int64_t j0 = 0; j0 < n; j0 += DB) {
    const int64_t jb = std::min<int64_t>(DB, n - j0);
    double* d = a + j0 + j0 * lda;
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, s, (int)jb, d, lda, lower ? 1 : 0,
                       info, j0);
    ZH_LAUNCH_CHECK();
    const int64_t r = n - j0 - jb;
    if (r <= 0) break;
    if (lower) {
      // A21 <- A21 L11^-T ; A22 <- A22 - A21 A21^T
      double* a21 = a + (j0 + jb) + j0 * lda;
      trsm_f64_cm(kRight, kLower, kTrans, kNonUnit, r, jb, 1.0, d, lda, a21, lda, s);
      syrk_f64_cm(kLower, kNoTrans, r, jb, -1.0, a21, lda, 1.0, a + (j0 + jb) + (j0 + jb) * lda, lda, s);
    } else {
      // A12 <- U11^-T A12 ; A22 <- A22 - A12^T A12
      double* a12 = a + j0 + (j0 + jb) * lda;
      trsm_f64_cm(kLeft, kUpper, kTrans, kNonUnit, jb, r, 1.0, d, lda, a12, lda, s);
      syrk_f64_cm(kUpper, kTrans, r, jb, -1.0, a12, lda, 1.0, a + (j0 + jb) + (j0 + jb) * lda, lda, s);
    }
  }
}

}  // namespace zmlhip
