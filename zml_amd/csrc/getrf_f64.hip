// FP64 LU with partial pivoting (zml lapack/getrf.zig:14-152, getrf2.zig:14-210).
//
// Storage: the factorization works on a row-major matrix (rows contiguous) so
// that row interchanges move contiguous memory. Column-major callers are
// transposed in place (square) or through a workspace copy around it.
//
// Structure (right-looking, blocked):
//   for each block column of NB = 256:
//     factor the (m - j) x 256 panel in leaves of 64 columns:
//       leaf: one persistent kernel over all panel rows (below)
//       leaf swaps -> other panel columns, TRSM + GEMM inside the panel
//     apply the panel's interchanges left and right of it (LASWP)
//     U12 = L11^-1 A12 (TRSM), A22 -= L21 U12 (MFMA GEMM)
//
// Leaf kernel: G workgroups of 256 threads, thread = one panel row held in
// registers (64 doubles). Per column k: workgroup argmax -> every workgroup
// publishes (|a|, row) and its candidate row, then a flag barrier; each
// workgroup then picks the global winner itself (largest |a|, lowest row on
// ties: iamax.zig:32 uses a strict '>'), swaps, scales by fl(1/p) when
// |p| >= sfmin and divides otherwise (getrf2.zig:65-87), and applies the
// rank-1 update with separately rounded multiply and subtract (no FMA),
// i.e. the very operation sequence of zml's unblocked closed form.
#include "zmlhip_internal.hpp"
#include "panel.hpp"

#include <algorithm>
#include <climits>

namespace zmlhip {

constexpr int NB_F64 = 256;     // block column width of the FP64 driver

// Row-major wrappers over the column-major kernels (zml's own mappings:
// gemm.zig:48-65, trsm.zig:48-62).
static void gemm_rm_nn(int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                       const double* b, int64_t ldb, double beta, double* c, int64_t ldc,
                       hipStream_t s) {
  gemm_f64_cm(kNoTrans, kNoTrans, n, m, k, alpha, b, ldb, a, lda, beta, c, ldc, s);
}
static void trsm_rm_llnu(int64_t m, int64_t n, const double* a, int64_t lda, double* b, int64_t ldb,
                         hipStream_t s) {
  trsm_f64_cm(kRight, kUpper, kNoTrans, kUnit, n, m, 1.0, a, lda, b, ldb, s);
}

struct LeafState {
  void* ws = nullptr;
  int launches = 0;
  int max_resident = 0;
};
static thread_local LeafState tl_leaf[16];

static LeafState& leaf_state(hipStream_t s, void** ws_out) {
  int dev = 0;
  ZH_CHECK(hipGetDevice(&dev));
  LeafState& st = tl_leaf[dev & 15];
  void* ws = workspace(panel::ws_bytes(), 3);
  if (ws != st.ws || st.launches > 60000000) {   // fresh workspace: all tags zero
    ZH_CHECK(hipMemsetAsync(ws, 0, panel::ws_bytes(), s));
    st.ws = ws;
    st.launches = 0;
  }
  if (st.max_resident == 0) {
    int per_cu = 0;
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, panel::leaf_kernel<double, true>, panel::pt_of<double>(), 0));
    st.max_resident = std::max(1, per_cu) * device_cu_count();
  }
  *ws_out = ws;
  return st;
}

// Factor one leaf (rows [r0, m), columns [c0, c0 + w)) of a row-major matrix.
template <typename T>
void leaf_factor(int64_t m, T* A, int64_t lda, int64_t r0, int64_t c0, int w, int32_t* ipiv,
                 int* info, hipStream_t s) {
  void* ws = nullptr;
  LeafState& st = leaf_state(s, &ws);
  const int64_t M = m - r0;
  constexpr int PT = panel::pt_of<T>();
  const int G = (int)((M + PT - 1) / PT);
  if (G > panel::GMAX || G > st.max_resident)
    throw std::runtime_error("getrf: the matrix is too tall for one resident panel grid");
  const panel::WS w_ = panel::carve(ws);
  if (G == 1) {
    hipLaunchKernelGGL((panel::leaf_kernel<T, false>), dim3(1), dim3(PT), 0, s, M, w,
                       A + r0 * lda + c0, lda, ipiv + r0, r0, info, c0, w_, 0u, 0L);
  } else {
    // tags of this launch: base + 1 .. base + 64, above every earlier launch's
    const uint32_t base = (uint32_t)st.launches * (uint32_t)(panel::W + 1);
    st.launches++;
    hipLaunchKernelGGL((panel::leaf_kernel<T, true>), dim3(G), dim3(PT), 0, s, M, w,
                       A + r0 * lda + c0, lda, ipiv + r0, r0, info, c0, w_, base, 2000000L);
  }
  ZH_LAUNCH_CHECK();
}
template void leaf_factor<double>(int64_t, double*, int64_t, int64_t, int64_t, int, int32_t*, int*, hipStream_t);
template void leaf_factor<float>(int64_t, float*, int64_t, int64_t, int64_t, int, int32_t*, int*, hipStream_t);

// Interchanges ipiv[k1..k2] (ascending) applied to `ncols` columns starting
// at `a` of a row-major matrix, through the composed permutation.
int32_t* compose_perm(int k1, int k2, const int32_t* ipiv, hipStream_t s) {
  if (k2 - k1 + 1 > panel::PERM_MAX / 2) throw std::runtime_error("compose_perm: block too wide");
  int32_t* pairs = static_cast<int32_t*>(workspace(sizeof(int32_t) * (2 * panel::PERM_MAX + 4), 4));
  hipLaunchKernelGGL(panel::compose_perm_kernel, dim3(1), dim3(64), 0, s, k1, k2, ipiv, pairs);
  ZH_LAUNCH_CHECK();
  return pairs;
}
template <typename T>
void apply_perm(int64_t ncols, T* a, int64_t lda, const int32_t* pairs, hipStream_t s) {
  if (ncols <= 0) return;
  hipLaunchKernelGGL((panel::apply_perm_kernel<T>), dim3((unsigned)((ncols + 15) / 16)), dim3(256), 0,
                     s, ncols, a, lda, pairs);
  ZH_LAUNCH_CHECK();
}
template void apply_perm<double>(int64_t, double*, int64_t, const int32_t*, hipStream_t);
template void apply_perm<float>(int64_t, float*, int64_t, const int32_t*, hipStream_t);

// Row-major FP64 getrf on device memory. ipiv (1-based, min(m,n) entries) and
// info live in device memory; info must be zero on entry.
void getrf_f64_rm(int64_t m, int64_t n, double* A, int64_t lda, int32_t* ipiv, int* info,
                  hipStream_t s) {
  const int64_t mn = std::min(m, n);
  for (int64_t j0 = 0; j0 < mn; j0 += NB_F64) {
    const int64_t jb = std::min<int64_t>(NB_F64, mn - j0);
    // ---- panel, in leaves
    for (int64_t k0 = j0; k0 < j0 + jb; k0 += panel::W) {
      const int kb = (int)std::min<int64_t>(panel::W, j0 + jb - k0);
      leaf_factor<double>(m, A, lda, k0, k0, kb, ipiv, info, s);
      // interchanges of this leaf on the other panel columns
      const int64_t rc = j0 + jb - (k0 + kb);
      if (k0 > j0 || rc > 0) {
        const int32_t* pr = compose_perm((int)k0, (int)(k0 + kb - 1), ipiv, s);
        apply_perm<double>(k0 - j0, A + j0, lda, pr, s);
        apply_perm<double>(rc, A + k0 + kb, lda, pr, s);
      }
      if (rc > 0) {
        trsm_rm_llnu(kb, rc, A + k0 * lda + k0, lda, A + k0 * lda + k0 + kb, lda, s);
        if (m - k0 - kb > 0)
          gemm_rm_nn(m - k0 - kb, rc, kb, -1.0, A + (k0 + kb) * lda + k0, lda, A + k0 * lda + k0 + kb,
                     lda, 1.0, A + (k0 + kb) * lda + k0 + kb, lda, s);
      }
    }
    // ---- interchanges outside the panel (getrf.zig:82-104)
    const int64_t nr = n - j0 - jb;
    if (j0 > 0 || nr > 0) {
      const int32_t* pr = compose_perm((int)j0, (int)(j0 + jb - 1), ipiv, s);
      apply_perm<double>(j0, A, lda, pr, s);
      apply_perm<double>(nr, A + j0 + jb, lda, pr, s);
    }
    if (nr > 0) {
      trsm_rm_llnu(jb, nr, A + j0 * lda + j0, lda, A + j0 * lda + j0 + jb, lda, s);
      if (m - j0 - jb > 0)
        gemm_rm_nn(m - j0 - jb, nr, jb, -1.0, A + (j0 + jb) * lda + j0, lda, A + j0 * lda + j0 + jb,
                   lda, 1.0, A + (j0 + jb) * lda + j0 + jb, lda, s);
    }
  }
}

int leaf_error_flag(hipStream_t s) {
  void* ws = workspace(panel::ws_bytes(), 3);
  const panel::WS w = panel::carve(ws);
  int h = 0;
  ZH_CHECK(hipMemcpyAsync(&h, w.err, sizeof(int), hipMemcpyDeviceToHost, s));
  ZH_CHECK(hipStreamSynchronize(s));
  return h;
}

}  // namespace zmlhip
