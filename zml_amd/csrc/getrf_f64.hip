// LU with partial pivoting, FP64 and FP32 (zml lapack/getrf.zig:14-152, getrf2.zig:14-210).
//
// Storage: the factorization works on a row-major matrix (rows contiguous) so
// that row interchanges move contiguous memory. Column-major callers are
// transposed in place (square) or through a workspace copy around it.
//
// Structure (right-looking, blocked):
//   for each block column of NB = 512:
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
#include "panel_fused.hpp"
#include "mixed.hpp"

#include <algorithm>
#include <climits>

namespace zmlhip {

constexpr int NB_F64 = 512;     // block column width of the FP64 driver

// Trailing-update kernels of a row-major LU: C -= L U and B <- L11^-1 B, each a
// column-major kernel on the transposed view (zml's own mappings: gemm.zig:48-65,
// trsm.zig:48-62).
// The wide U12 solve goes through the inverse of the unit lower L11: the
// panel stream forms L11^-1 right after factoring the block column, and the
// main stream then computes U12 = L11^-1 A12 as one GEMM on a copy of A12
// (LuOps is declared in mixed.hpp, shared with the mixed-precision driver).

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
    int per_cu_la = 0;
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, panel::leaf_kernel<double, true>, panel::pt_of<double>(), 0));
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu_la, panel::leaf_la_kernel<double>, panel::la_threads<double>(), 0));
    st.max_resident = std::max(1, std::min(per_cu, per_cu_la)) * device_cu_count();
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
    if (leaf_lookahead())
      hipLaunchKernelGGL((panel::leaf_la_kernel<T>), dim3(G), dim3(panel::la_threads<T>()), 0, s, M, w, A + r0 * lda + c0, lda,
                         ipiv + r0, r0, info, c0, w_, base, panel_spin_limit());
    else
      hipLaunchKernelGGL((panel::leaf_kernel<T, true>), dim3(G), dim3(PT), 0, s, M, w,
                         A + r0 * lda + c0, lda, ipiv + r0, r0, info, c0, w_, base, panel_spin_limit());
  }
  ZH_LAUNCH_CHECK();
}
template void leaf_factor<double>(int64_t, double*, int64_t, int64_t, int64_t, int, int32_t*, int*, hipStream_t);
template void leaf_factor<float>(int64_t, float*, int64_t, int64_t, int64_t, int, int32_t*, int*, hipStream_t);

// Interchanges ipiv[k1..k2] (ascending) applied to `ncols` columns starting
// at `a` of a row-major matrix, through the composed permutation. Each
// stream that composes permutations uses its own workspace slot.
int32_t* perm_pairs_buffer(int slot) {
  return static_cast<int32_t*>(workspace(sizeof(int32_t) * (2 * panel::PERM_MAX + 4), slot));
}
int32_t* compose_perm(int k1, int k2, const int32_t* ipiv, hipStream_t s, int slot) {
  if (k2 - k1 + 1 > panel::PERM_MAX / 2) throw std::runtime_error("compose_perm: block too wide");
  int32_t* pairs = perm_pairs_buffer(slot);
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

// Factor the block column [j0, j0 + jb) (rows j0..m-1) of a row-major matrix
// in leaves; interchanges are applied inside the block column only.
template <typename T>
void factor_panel(int64_t m, T* A, int64_t lda, int64_t j0, int64_t jb, int32_t* ipiv, int* info,
                  const LuOps<T>& ops, hipStream_t s, int perm_slot) {
  for (int64_t k0 = j0; k0 < j0 + jb; k0 += panel::W) {
    const int kb = (int)std::min<int64_t>(panel::W, j0 + jb - k0);
    leaf_factor<T>(m, A, lda, k0, k0, kb, ipiv, info, s);
    const int64_t rc = j0 + jb - (k0 + kb);
    if (k0 > j0 || rc > 0) {
      const int32_t* pr = compose_perm((int)k0, (int)(k0 + kb - 1), ipiv, s, perm_slot);
      apply_perm<T>(k0 - j0, A + j0, lda, pr, s);
      apply_perm<T>(rc, A + k0 + kb, lda, pr, s);
    }
    if (rc > 0) {
      ops.trsm(kb, rc, A + k0 * lda + k0, lda, A + k0 * lda + k0 + kb, lda, s);
      if (m - k0 - kb > 0)
        (ops.panel_gemm ? ops.panel_gemm : ops.gemm)(m - k0 - kb, rc, kb, A + (k0 + kb) * lda + k0,
                                                     lda, A + k0 * lda + k0 + kb, lda,
                                                     A + (k0 + kb) * lda + k0 + kb, lda, s);
    }
  }
}

// Row-major right-looking LU with look-ahead of depth one: while the main
// stream applies step b to the columns beyond block b+1, the panel stream
// already factors block b+1 (whose columns step b updated first). ipiv
// (1-based) and info live in device memory; info must be zero on entry. When a
// panel's bounded wait gave up, info ends as kErrSync (the factors are invalid).
template <typename T>
void getrf_lookahead(int64_t m, int64_t n, T* A, int64_t lda, int32_t* ipiv, int* info, int64_t nb,
                     const LuOps<T>& ops, hipStream_t s) {
  const int64_t mn = std::min(m, n);
  if (mn <= 0) return;
  clear_leaf_error(s);
  hipStream_t sp = aux_stream(0);
  hipEvent_t ev_in = aux_event(0), ev_panel = aux_event(1), ev_next = aux_event(2);
  const hipStream_t caller = s;
  if (ops.spare_cus > 0) {   // the trailing updates move to a CU-sparing stream
    s = stream_sparing_cus(ops.spare_cus);
    ZH_CHECK(hipEventRecord(ev_in, caller));
    ZH_CHECK(hipStreamWaitEvent(s, ev_in, 0));
  }
  // L11^-1 of consecutive block columns alternate between two buffers; the
  // copy of A12 is used by the main stream only
  T* inv = static_cast<T*>(workspace(sizeof(T) * 2 * nb * nb, 0));
  T* tmp = static_cast<T*>(workspace(sizeof(T) * nb * (size_t)std::max<int64_t>(n, 1), 1));
  // a panel in two parts: the optimistic attempt is queued first; its outcome
  // (and the full pivot search after a failed one) and the inverse follow
  auto one_launch = [&](int64_t j0) { return ops.panel && m - j0 > ops.panel_min_rows; };
  auto panel_begin = [&](int64_t j0, int64_t jb) {
    return ops.try_panel && ops.try_panel(m, A, lda, j0, jb, ipiv, sp);
  };
  auto panel_end = [&](int64_t j0, int64_t jb, int64_t blk, bool attempted) {
    const bool ok = attempted && (!ops.try_panel_finish || ops.try_panel_finish(j0, jb, A, lda, sp));
    if (!ok) {
      if (one_launch(j0)) ops.panel(m, A, lda, j0, jb, ipiv, info, sp);
      else factor_panel<T>(m, A, lda, j0, jb, ipiv, info, ops, sp, 2);
    }
    T* invp = ops.inv_store ? ops.inv_store(j0) : inv + (blk & 1) * nb * nb;
    if (j0 + jb < n || ops.inv_store) ops.invert(jb, A + j0 * lda + j0, lda, invp, sp);
    ZH_CHECK(hipEventRecord(ev_panel, sp));
  };
  ZH_CHECK(hipEventRecord(ev_in, s));
  ZH_CHECK(hipStreamWaitEvent(sp, ev_in, 0));
  {
    const int64_t jb0 = std::min(nb, mn);
    panel_end(0, jb0, 0, panel_begin(0, jb0));
  }
  for (int64_t j0 = 0, blk = 0; j0 < mn; j0 += nb, ++blk) {
    const int64_t jb = std::min(nb, mn - j0);
    const T* invb = ops.inv_store ? ops.inv_store(j0) : inv + (blk & 1) * nb * nb;
    ZH_CHECK(hipStreamWaitEvent(s, ev_panel, 0));
    const int32_t* pr = ops.block_perm && (!ops.panel || one_launch(j0)) ? ops.block_perm(j0) : nullptr;
    if (!pr) pr = compose_perm((int)j0, (int)(j0 + jb - 1), ipiv, s, 4);
    const int64_t c1 = j0 + jb;                       // first column right of the block
    const int64_t jb2 = std::min(nb, std::max<int64_t>(0, mn - c1));   // next block width
    const int64_t rows = m - c1;
    if (jb2 > 0) {
      // update the next block column first, then hand it to the panel stream
      apply_perm<T>(jb2, A + c1, lda, pr, s);
      ops.apply_inv(jb, jb2, invb, A + j0 * lda + c1, lda, tmp, s);
      if (rows > 0) ops.gemm(rows, jb2, jb, A + c1 * lda + j0, lda, A + j0 * lda + c1, lda, A + c1 * lda + c1, lda, s);
      ZH_CHECK(hipEventRecord(ev_next, s));
    }
    // a split panel's attempt goes ahead of this step's remaining work (the
    // host then queues that work while the panel runs)
    bool begun = false, attempted = false;
    if (jb2 > 0 && ops.try_panel_finish) {
      ZH_CHECK(hipStreamWaitEvent(sp, ev_next, 0));
      attempted = panel_begin(c1, jb2);
      begun = true;
    }
    // the rest: interchanges left of the block and beyond the next one, then
    // update (queued before the next panel's outcome, which waits on the host)
    apply_perm<T>(j0, A, lda, pr, s);
    const int64_t c2 = c1 + jb2, nr = n - c2;
    if (nr > 0) {
      apply_perm<T>(nr, A + c2, lda, pr, s);
      ops.apply_inv(jb, nr, invb, A + j0 * lda + c2, lda, tmp, s);
      if (rows > 0)
        ops.gemm(rows, nr, jb, A + c1 * lda + j0, lda, A + j0 * lda + c2, lda, A + c1 * lda + c2, lda, s);
    }
    if (jb2 > 0) {
      if (!begun) {
        ZH_CHECK(hipStreamWaitEvent(sp, ev_next, 0));
        attempted = panel_begin(c1, jb2);
      }
      panel_end(c1, jb2, blk + 1, attempted);
    }
  }
  if (s != caller) {
    hipEvent_t ev_out = aux_event(3);
    ZH_CHECK(hipEventRecord(ev_out, s));
    ZH_CHECK(hipStreamWaitEvent(caller, ev_out, 0));
  }
  report_leaf_error(info, caller);
}

template <typename T>
static void gemm_update(int64_t rows, int64_t cols, int64_t kb, const T* L, int64_t ldl, const T* U,
                        int64_t ldu, T* C, int64_t ldc, hipStream_t s) {
  void* tok = sizeof(T) == 8 ? profile_begin(kProfGemmF64, s) : nullptr;
  // row-major C -= L U is column-major C^T -= U^T L^T (gemm.zig:48-65)
  gemm_cm<T>(kNoTrans, kNoTrans, cols, rows, kb, T(-1), U, ldu, L, ldl, T(1), C, ldc, s, 0);
  // algorithmic traffic: L and U read once, C read and written once
  profile_end(tok, s, 2.0 * rows * cols * kb, (double)sizeof(T) * (rows * kb + kb * cols + 2.0 * rows * cols));
}
template <typename T>
static void trsm_update(int64_t kb, int64_t cols, const T* L11, int64_t ld, T* B, int64_t ldb, hipStream_t s) {
  trsm_cm<T>(kRight, kUpper, kNoTrans, kUnit, cols, kb, T(1), L11, ld, B, ldb, s);
}

template <typename T>
__global__ void set_identity_kernel(int64_t n, T* __restrict__ a) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n * n) a[e] = (e / n == e % n) ? T(1) : T(0);
}
template <typename T>
static void invert_unit_lower(int64_t kb, const T* L11, int64_t ld, T* inv, hipStream_t s) {
  hipLaunchKernelGGL(set_identity_kernel<T>, dim3((unsigned)((kb * kb + 255) / 256)), dim3(256), 0, s, kb, inv);
  ZH_LAUNCH_CHECK();
  trsm_update<T>(kb, kb, L11, ld, inv, kb, s);
}

// ---- X = L^-1 of a row-major unit-lower kb x kb block (FP64), in 64 x 64
// blocks: first every diagonal block X_ii = L_ii^-1 (one workgroup per block;
// lane c of one wave forms column c, x_rc = -sum_{k=c}^{r-1} l_rk x_kc with k
// ascending, in registers), then, for the block distances d = 1, 2, ... in
// turn, every block with i - j = d at once:
//   X_ij = -X_ii (sum_{k=j}^{i-1} L_ik X_kj)
// (one launch per distance, one workgroup per block, the 64 x 64 products on
// FP64 MFMA). The strictly upper part of X is written as zeros. kb / 64
// launches in all, against two per 32 columns for the blocked right-side
// triangular solve of the identity (invert_unit_lower); the values are those
// of another summation order.
namespace invblk {
constexpr int IB = 64, ILD = IB + 1;
typedef double d4 __attribute__((ext_vector_type(4)));

// 64 x 64 block of row-major rows [r0, r0 + 64) and columns [c0, c0 + 64) into
// LDS (row stride ILD); rows from nr on and columns from nc on are zeros (their
// loads are clamped into the block), all loads issued first
__device__ __forceinline__ void stage(double* __restrict__ S, const double* __restrict__ G, int64_t ld, int r0,
                                      int c0, int nr, int nc, int t) {
  const int b = t & 63, bc = b < nc ? b : nc - 1;
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int a = (t >> 6) + 4 * q;
    v[q] = G[(int64_t)(r0 + (a < nr ? a : nr - 1)) * ld + c0 + bc];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int a = (t >> 6) + 4 * q;
    S[a * ILD + b] = (a < nr && b < nc) ? v[q] : 0.0;
  }
}

// acc += P Q for the 32 x 32 quadrant (wm, wn) of a 64 x 64 product, P and Q
// 64 x 64 in LDS. Lane l supplies P[row l % 16][k l / 16] and Q[k l / 16][col
// l % 16] of each 16 x 16 x 4 step and holds C[row 4 r + l / 16][col l % 16].
__device__ __forceinline__ void mma64(const double* __restrict__ P, const double* __restrict__ Q, d4 (&acc)[2][2],
                                      int wm, int wn, int lane) {
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int ks = 0; ks < IB / 4; ++ks) {
    const int k = 4 * ks + lk;
    double a[2], b[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) a[x] = P[(wm * 32 + 16 * x + lr) * ILD + k];
#pragma unroll
    for (int y = 0; y < 2; ++y) b[y] = Q[k * ILD + wn * 32 + 16 * y + lr];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
  }
}

__global__ void __launch_bounds__(256) diag_kernel(int kb, const double* __restrict__ L, int64_t ld,
                                                   double* __restrict__ X, int64_t ldx) {
  __shared__ double Ls[IB * ILD], Xs[IB * ILD];
  const int i = blockIdx.x, r0 = i * IB, nr = min(IB, kb - r0);
  const int t = threadIdx.x;
  stage(Ls, L, ld, r0, r0, nr, nr, t);
  __syncthreads();
  if (t < IB) {
    // column c of L_ii^-1 in registers, right-looking: x_r -= l_rk x_k for
    // k ascending (terms with k < c are exact zeros); every lane reads the same
    // l_rk (an LDS broadcast), the loops fully unrolled
    const int c = t;
    double x[IB];
#pragma unroll
    for (int r = 0; r < IB; ++r) x[r] = r == c ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < IB; ++k)
#pragma unroll
      for (int r = k + 1; r < IB; ++r) x[r] = fma(-Ls[r * ILD + k], x[k], x[r]);
#pragma unroll
    for (int r = 0; r < IB; ++r) Xs[r * ILD + c] = x[r];
  }
  __syncthreads();
  // block row i of X: the diagonal block, zeros right of it
  for (int a = t >> 6; a < nr; a += 4)
    for (int b = r0 + (t & 63); b < kb; b += 64)
      X[(int64_t)(r0 + a) * ldx + b] = b < r0 + nr ? Xs[a * ILD + (b - r0)] : 0.0;
}

__global__ void __launch_bounds__(256) dist_kernel(int kb, int d, const double* __restrict__ L, int64_t ld,
                                                   double* __restrict__ X, int64_t ldx) {
  __shared__ double Ps[IB * ILD], Qs[IB * ILD];
  const int i = d + blockIdx.x, j = i - d, r0 = i * IB, c0 = j * IB, nr = min(IB, kb - r0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w & 1, wn = w >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
  // S = sum_{k=j}^{i-1} L_ik X_kj (blocks k < i are whole)
  for (int kk = j; kk < i; ++kk) {
    __syncthreads();
    stage(Ps, L, ld, r0, kk * IB, nr, IB, t);
    stage(Qs, X, ldx, kk * IB, c0, IB, IB, t);
    __syncthreads();
    mma64(Ps, Qs, acc, wm, wn, lane);
  }
  __syncthreads();
  // S to Qs, X_ii (rows and columns from nr on zero) to Ps
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Qs[(wm * 32 + 16 * x + 4 * r + lk) * ILD + wn * 32 + 16 * y + lr] = acc[x][y][r];
  stage(Ps, X, ldx, r0, r0, nr, nr, t);
  __syncthreads();
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
  mma64(Ps, Qs, acc, wm, wn, lane);
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = wm * 32 + 16 * x + 4 * r + lk;
        if (a < nr) X[(int64_t)(r0 + a) * ldx + c0 + wn * 32 + 16 * y + lr] = -acc[x][y][r];
      }
}
}  // namespace invblk

static void invert_unit_lower_blocked_f64(int64_t kb, const double* L11, int64_t ld, double* inv, hipStream_t s) {
  const int nbk = (int)((kb + invblk::IB - 1) / invblk::IB);
  hipLaunchKernelGGL(invblk::diag_kernel, dim3(nbk), dim3(256), 0, s, (int)kb, L11, ld, inv, kb);
  ZH_LAUNCH_CHECK();
  for (int d = 1; d < nbk; ++d) {
    hipLaunchKernelGGL(invblk::dist_kernel, dim3(nbk - d), dim3(256), 0, s, (int)kb, d, L11, ld, inv, kb);
    ZH_LAUNCH_CHECK();
  }
}
template <typename T>
static void apply_inv(int64_t kb, int64_t cols, const T* inv, T* B, int64_t ldb, T* tmp, hipStream_t s) {
  ZH_CHECK(hipMemcpy2DAsync(tmp, sizeof(T) * cols, B, sizeof(T) * ldb, sizeof(T) * cols, kb,
                            hipMemcpyDeviceToDevice, s));
  gemm_cm<T>(kNoTrans, kNoTrans, cols, kb, kb, T(1), tmp, cols, inv, kb, T(0), B, ldb, s, 0);
}

// ---- one-launch FP64 panels (panel_fused.hpp) -------------------------------
struct FusedState {
  void* ws = nullptr;
  uint32_t next_tag = 0;
  int max_resident = 0;
  int32_t* pairs[2] = {nullptr, nullptr};   // moves of the last two block columns (by parity)
};
static thread_local FusedState tl_fused[16];

static FusedState& fused_state(hipStream_t s) {
  int dev = 0;
  ZH_CHECK(hipGetDevice(&dev));
  FusedState& st = tl_fused[dev & 15];
  void* ws = workspace(panel::fused_ws_bytes(), 22);
  if (ws != st.ws || st.next_tag > 0x7fff0000u) {   // fresh workspace: every tag and barrier word zero
    ZH_CHECK(hipMemsetAsync(ws, 0, panel::fused_ws_bytes(), s));
    st.ws = ws;
    st.next_tag = 0;
  }
  if (st.max_resident == 0) {
    int per_cu = 0;
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, panel::panel_fused_kernel<double>, panel::FPT,
                                                          panel::FUSED_PAD_LDS));
    st.max_resident = std::max(0, per_cu) * device_cu_count();
  }
  st.pairs[0] = perm_pairs_buffer(23);
  st.pairs[1] = perm_pairs_buffer(24);
  return st;
}

// Before each one-launch panel: no moves yet, no barrier arrivals (a launch
// that gave up may have left some).
__global__ void clear_count_kernel(int32_t* __restrict__ pairs, unsigned* __restrict__ bar) {
  pairs[0] = 0;
  bar[0] = 0;
}

// Factors block column [j0, j0 + jb) of the row-major m x n matrix in one
// launch and puts its rows in LAPACK order (the moves, kept for the rest of
// the matrix, applied to the block column itself on this stream).
static void fused_panel_f64(int64_t m, double* A, int64_t lda, int64_t j0, int64_t jb, int32_t* ipiv, int* info,
                            hipStream_t s) {
  FusedState& st = fused_state(s);
  const int64_t M = m - j0;
  const int G = (int)((M + panel::FPT - 1) / panel::FPT);
  if (jb > panel::FW || G > panel::GMAX || G > st.max_resident)
    throw std::runtime_error("getrf: the block column is too tall for one resident panel grid");
  int32_t* pairs = st.pairs[(j0 / NB_F64) & 1];
  const panel::FusedWS fw = panel::fused_carve(st.ws);
  hipLaunchKernelGGL(clear_count_kernel, dim3(1), dim3(1), 0, s, pairs, fw.bar);
  const uint32_t base = st.next_tag;
  st.next_tag += (uint32_t)panel::FW + 1u;
  hipLaunchKernelGGL((panel::panel_fused_kernel<double>), dim3(G), dim3(panel::FPT), panel::FUSED_PAD_LDS, s, M,
                     (int)jb,
                     A + j0 * lda + j0, lda, ipiv + j0, j0, info, j0, pairs, fw, base,
                     panel_spin_limit());
  ZH_LAUNCH_CHECK();
  apply_perm<double>(jb, A + j0, lda, pairs, s);
}

static const int32_t* fused_block_perm(int64_t j0) {
  int dev = 0;
  ZH_CHECK(hipGetDevice(&dev));
  return tl_fused[dev & 15].pairs[(j0 / NB_F64) & 1];
}

// The share of a trailing update to apply beside the next panel (on the CUs
// the panel leaves free): a fixed 5 ms of update on 192 CUs at 44 TF/s, the
// rest then runs on all 256 CUs once the panel is done. (Measured on one
// MI355X at n = 32768: 718 ms per LU; sizing it to the one-launch panel's
// 7.4 ms plus 20% gave 729 ms.)

// Plain partial-pivoting LU of a row-major matrix in T (FP64 for dgetrf, FP32
// for sgetrf): the same driver, leaves and kernels, no low-precision update.
// FP64 factors each block column in one launch (panel_fused.hpp).
template <typename T>
void getrf_rm(int64_t m, int64_t n, T* A, int64_t lda, int32_t* ipiv, int* info, hipStream_t s) {
  // The pivot panels keep one workgroup per CU resident for the whole panel:
  // FP64 block columns taller than a height threshold are factored in one launch
  // (panel_fused.hpp, 512 rows per workgroup), shorter ones (and FP32) leaf by
  // leaf (256 / 512 rows per workgroup). When the first grid fits in a quarter
  // of the CUs, those CUs are kept free of trailing updates so the panel never
  // waits for GEMM tiles to drain (rounded up to a power of two from 8, so few
  // distinct masked streams exist).
  // The height threshold (panel_one_launch_rows, default 16384): measured per
  // panel height on one MI355X, the one-launch panel wins above 16384 rows
  // (where the leaf's 256-row grid no longer fits the reserved CUs) and loses
  // below (n = 16384: 242 ms against 172 ms), its in-kernel update of the
  // block column being slower than the MFMA GEMMs the leaf path launches
  // between leaves.
  const int64_t kFusedMinRows = panel_one_launch_rows();
  const int64_t first_rows = sizeof(T) == 8 && m > kFusedMinRows ? panel::FPT : panel::pt_of<T>();
  const int leaf_wgs = (int)((m + first_rows - 1) / first_rows);
  int spare = 0;
  if (leaf_wgs > 1) {
    spare = 8;
    while (spare < leaf_wgs) spare *= 2;
    if (spare > device_cu_count() / 4) spare = 0;
  }
  LuOps<T> ops{gemm_update<T>, trsm_update<T>, invert_unit_lower<T>, apply_inv<T>, nullptr, nullptr,
               nullptr, spare, nullptr};
  if constexpr (sizeof(T) == 8) {
    ops.invert = invert_unit_lower_blocked_f64;
    ops.panel = fused_panel_f64;
    ops.block_perm = fused_block_perm;
    ops.panel_min_rows = kFusedMinRows;
  }
  getrf_lookahead<T>(m, n, A, lda, ipiv, info, NB_F64, ops, s);
}
template void getrf_rm<float>(int64_t, int64_t, float*, int64_t, int32_t*, int*, hipStream_t);
template void getrf_rm<double>(int64_t, int64_t, double*, int64_t, int32_t*, int*, hipStream_t);

void getrf_f64_rm(int64_t m, int64_t n, double* A, int64_t lda, int32_t* ipiv, int* info,
                  hipStream_t s) {
  getrf_rm<double>(m, n, A, lda, ipiv, info, s);
}

template void getrf_lookahead<float>(int64_t, int64_t, float*, int64_t, int32_t*, int*, int64_t,
                                     const LuOps<float>&, hipStream_t);

__global__ void clear_flag_kernel(int* __restrict__ flag) { *flag = 0; }
__global__ void report_flag_kernel(const int* __restrict__ flag, int* __restrict__ info) {
  if (*flag) *info = kErrSync;
}

__global__ void clear_fused_kernel(int* __restrict__ flag, unsigned* __restrict__ bar) {
  *flag = 0;
  bar[0] = 0;   // arrivals (the generation word may keep any value)
}

void clear_leaf_error(hipStream_t s) {
  void* ws = nullptr;
  leaf_state(s, &ws);
  hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, s, panel::carve(ws).err);
  const panel::FusedWS fw = panel::fused_carve(fused_state(s).ws);
  hipLaunchKernelGGL(clear_fused_kernel, dim3(1), dim3(1), 0, s, fw.err, fw.bar);
  ZH_LAUNCH_CHECK();
}

void report_leaf_error(int* info, hipStream_t s) {
  void* ws = nullptr;
  leaf_state(s, &ws);
  hipLaunchKernelGGL(report_flag_kernel, dim3(1), dim3(1), 0, s, (const int*)panel::carve(ws).err, info);
  hipLaunchKernelGGL(report_flag_kernel, dim3(1), dim3(1), 0, s, (const int*)panel::fused_carve(fused_state(s).ws).err,
                     info);
  ZH_LAUNCH_CHECK();
}

}  // namespace zmlhip
