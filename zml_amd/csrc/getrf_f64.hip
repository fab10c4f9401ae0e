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

#include <algorithm>
#include <climits>
#include <vector>

namespace zmlhip {

constexpr int PW = 64;          // leaf width
constexpr int PT = 256;         // panel rows per workgroup
constexpr int GMAX = 1024;      // max workgroups of a leaf (m <= 262144)
constexpr int FLAG_STRIDE = 32; // uint32 per flag (128 bytes apart)
constexpr int NB_F64 = 256;     // block column width of the FP64 driver

struct LeafWS {
  uint32_t* flags;   // [GMAX * FLAG_STRIDE]
  double* val;       // [2][GMAX]
  int64_t* idx;      // [2][GMAX]
  double* rows;      // [2][GMAX][PW]
  double* rowk;      // [2][PW]
  int* nan;          // [2]
  int* err;          // [1]
};

size_t leaf_ws_bytes() {
  return sizeof(uint32_t) * GMAX * FLAG_STRIDE + sizeof(double) * 2 * GMAX +
         sizeof(int64_t) * 2 * GMAX + sizeof(double) * 2 * GMAX * PW + sizeof(double) * 2 * PW +
         sizeof(int) * 4 + 256;
}

static LeafWS carve_leaf_ws(void* base) {
  char* p = static_cast<char*>(base);
  LeafWS w;
  w.flags = reinterpret_cast<uint32_t*>(p); p += sizeof(uint32_t) * GMAX * FLAG_STRIDE;
  w.val = reinterpret_cast<double*>(p); p += sizeof(double) * 2 * GMAX;
  w.idx = reinterpret_cast<int64_t*>(p); p += sizeof(int64_t) * 2 * GMAX;
  w.rows = reinterpret_cast<double*>(p); p += sizeof(double) * 2 * GMAX * PW;
  w.rowk = reinterpret_cast<double*>(p); p += sizeof(double) * 2 * PW;
  w.nan = reinterpret_cast<int*>(p); p += sizeof(int) * 2;
  w.err = reinterpret_cast<int*>(p);
  return w;
}

namespace {

__device__ __forceinline__ uint32_t ld_flag(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void amax2(double& v, int64_t& i, double v2, int64_t i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

__device__ __forceinline__ void wave_amax(double& v, int64_t& i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double v2 = __shfl_xor(v, off);
    const int64_t i2 = __shfl_xor(i, off);
    amax2(v, i, v2, i2);
  }
}

__global__ void __launch_bounds__(PT)
getrf_leaf_kernel(int64_t M, int W, double* __restrict__ P, int64_t ld, int32_t* __restrict__ ipiv,
                  int64_t piv_base, int* __restrict__ info, int64_t info_base, LeafWS ws,
                  uint32_t flag_base, long spin_limit) {
#pragma clang fp contract(off)
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int64_t i = (int64_t)g * PT + t;
  const bool own = i < M;

  __shared__ double u[PW], rk[PW];
  __shared__ double s_v[PT / 64];
  __shared__ int64_t s_i[PT / 64];
  __shared__ int64_t s_p;
  __shared__ int s_pg, s_abort;

  double a[PW];
#pragma unroll
  for (int c = 0; c < PW; ++c) a[c] = (own && c < W) ? P[i * ld + c] : 0.0;
  if (t == 0) s_abort = 0;

  int first_zero = 0;
  const double sfmin = 0x1p-1022;   // lamch('S') for f64 (lamch.zig:34-42)

  for (int k = 0; k < W; ++k) {
    const int par = k & 1;
    double ak = 0.0;
#pragma unroll
    for (int c = 0; c < PW; ++c) if (c == k) ak = a[c];

    // ---- local candidate (NaN never wins; iamax.zig:32)
    double v = -1.0;
    int64_t vi = INT64_MAX;
    if (own && i >= k) {
      const double f = fabs(ak);
      if (f == f) { v = f; vi = i; }
    }
    wave_amax(v, vi);
    if (lane == 0) { s_v[wv] = v; s_i[wv] = vi; }
    __syncthreads();
    double bv = s_v[0];
    int64_t bi = s_i[0];
#pragma unroll
    for (int w = 1; w < PT / 64; ++w) amax2(bv, bi, s_v[w], s_i[w]);

    // ---- publish candidate row, row k and the workgroup result
    if (own && i == bi) {
      double* dst = ws.rows + ((size_t)par * GMAX + g) * PW;
#pragma unroll
      for (int c = 0; c < PW; ++c) dst[c] = a[c];
    }
    if (own && i == k) {
      double* dst = ws.rowk + (size_t)par * PW;
#pragma unroll
      for (int c = 0; c < PW; ++c) dst[c] = a[c];
      ws.nan[par] = (ak != ak) ? 1 : 0;
    }
    if (t == 0) {
      ws.val[par * GMAX + g] = bv;
      ws.idx[par * GMAX + g] = bi;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    const uint32_t target = flag_base + (uint32_t)k + 1u;
    if (t == 0) __hip_atomic_store(ws.flags + (size_t)g * FLAG_STRIDE, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);

    // ---- flag barrier: wave 0 waits until every workgroup has published step k
    if (wv == 0) {
      long spins = 0;
      while (true) {
        bool ok = true;
        for (int gg = lane; gg < G; gg += 64) ok = ok && (ld_flag(ws.flags + (size_t)gg * FLAG_STRIDE) >= target);
        if (__all(ok)) break;
        if (++spins > spin_limit) {
          if (lane == 0) { s_abort = 1; atomicExch(ws.err, 1); }
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    if (s_abort) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

    // ---- global winner
    if (wv == 0) {
      double gv = -1.0;
      int64_t gi = INT64_MAX;
      for (int gg = lane; gg < G; gg += 64) amax2(gv, gi, ws.val[par * GMAX + gg], ws.idx[par * GMAX + gg]);
      wave_amax(gv, gi);
      if (lane == 0) {
        const int nan_k = ws.nan[par];
        int64_t p = (nan_k || gi == INT64_MAX) ? k : gi;
        s_p = p;
        s_pg = nan_k ? -1 : (int)(p / PT);
      }
    }
    __syncthreads();
    const int64_t p = s_p;
    if (t < PW) {
      const double* src = s_pg < 0 ? ws.rowk + (size_t)par * PW
                                   : ws.rows + ((size_t)par * GMAX + s_pg) * PW;
      u[t] = src[t];
      rk[t] = ws.rowk[(size_t)par * PW + t];
    }
    __syncthreads();
    const double piv = u[k];
    const bool swap_ok = piv != 0.0;   // getrf2.zig:55: no interchange on an exact zero
    if (!swap_ok && first_zero == 0) first_zero = k + 1;
    if (g == 0 && t == 0) ipiv[k] = (int32_t)(piv_base + (swap_ok ? p : k) + 1);
    if (swap_ok && p != k) {
      if (own && i == k) {
#pragma unroll
        for (int c = 0; c < PW; ++c) a[c] = u[c];
      } else if (own && i == p) {
#pragma unroll
        for (int c = 0; c < PW; ++c) a[c] = rk[c];
      }
    }
    // ---- scale the column and update the rest of the leaf
    if (own && i > k) {
      double l = 0.0;
#pragma unroll
      for (int c = 0; c < PW; ++c) if (c == k) l = a[c];
      if (swap_ok) {
        if (fabs(piv) >= sfmin) {
          const double r = 1.0 / piv;
          l = l * r;
        } else {
          l = l / piv;
        }
      }
#pragma unroll
      for (int c = 0; c < PW; ++c) {
        if (c == k) a[c] = l;
        else if (c > k) a[c] = a[c] - l * u[c];
      }
    }
  }

#pragma unroll
  for (int c = 0; c < PW; ++c)
    if (own && c < W) P[i * ld + c] = a[c];
  if (g == 0 && t == 0 && first_zero != 0 && *info == 0) *info = (int)(info_base + first_zero);
}

}  // namespace

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
  uint32_t epoch = 0;
};
static thread_local LeafState tl_leaf[16];

// Factor one leaf (rows [r0, m), columns [c0, c0 + w)) of a row-major matrix.
static void leaf(int64_t m, double* A, int64_t lda, int64_t r0, int64_t c0, int w, int32_t* ipiv,
                 int* info, hipStream_t s) {
  int dev = 0;
  ZH_CHECK(hipGetDevice(&dev));
  LeafState& st = tl_leaf[dev & 15];
  void* ws = workspace(leaf_ws_bytes(), 3);
  if (ws != st.ws || st.epoch > 60000000u) {
    ZH_CHECK(hipMemsetAsync(ws, 0, leaf_ws_bytes(), s));
    st.ws = ws;
    st.epoch = 0;
  }
  const uint32_t base = st.epoch * (uint32_t)(PW + 1);
  st.epoch++;
  const int64_t M = m - r0;
  const int G = (int)((M + PT - 1) / PT);
  // every workgroup of the leaf must be resident at once (flag barrier)
  static thread_local int max_resident = 0;
  if (max_resident == 0) {
    int per_cu = 0;
    ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, getrf_leaf_kernel, PT, 0));
    max_resident = std::max(1, per_cu) * device_cu_count();
  }
  if (G > GMAX || G > max_resident)
    throw std::runtime_error("getrf: the matrix is too tall for one resident panel grid");
  LeafWS w_ = carve_leaf_ws(ws);
  hipLaunchKernelGGL(getrf_leaf_kernel, dim3(G), dim3(PT), 0, s, M, w, A + r0 * lda + c0, lda,
                     ipiv + r0, r0, info, c0, w_, base, 2000000L);
  ZH_LAUNCH_CHECK();
}

// Row-major FP64 getrf on device memory. ipiv (1-based, min(m,n) entries) and
// info live in device memory; info must be zero on entry.
void getrf_f64_rm(int64_t m, int64_t n, double* A, int64_t lda, int32_t* ipiv, int* info,
                  hipStream_t s) {
  const int64_t mn = std::min(m, n);
  for (int64_t j0 = 0; j0 < mn; j0 += NB_F64) {
    const int64_t jb = std::min<int64_t>(NB_F64, mn - j0);
    // ---- panel, in leaves
    for (int64_t k0 = j0; k0 < j0 + jb; k0 += PW) {
      const int kb = (int)std::min<int64_t>(PW, j0 + jb - k0);
      leaf(m, A, lda, k0, k0, kb, ipiv, info, s);
      // interchanges of this leaf on the other panel columns
      if (k0 > j0) laswp_rm(k0 - j0, A + j0, lda, (int)k0, (int)(k0 + kb - 1), ipiv, 1, s);
      const int64_t rc = j0 + jb - (k0 + kb);
      if (rc > 0) {
        laswp_rm(rc, A + k0 + kb, lda, (int)k0, (int)(k0 + kb - 1), ipiv, 1, s);
        trsm_rm_llnu(kb, rc, A + k0 * lda + k0, lda, A + k0 * lda + k0 + kb, lda, s);
        if (m - k0 - kb > 0)
          gemm_rm_nn(m - k0 - kb, rc, kb, -1.0, A + (k0 + kb) * lda + k0, lda, A + k0 * lda + k0 + kb,
                     lda, 1.0, A + (k0 + kb) * lda + k0 + kb, lda, s);
      }
    }
    // ---- interchanges outside the panel (getrf.zig:82-104)
    if (j0 > 0) laswp_rm(j0, A, lda, (int)j0, (int)(j0 + jb - 1), ipiv, 1, s);
    const int64_t nr = n - j0 - jb;
    if (nr > 0) {
      laswp_rm(nr, A + j0 + jb, lda, (int)j0, (int)(j0 + jb - 1), ipiv, 1, s);
      trsm_rm_llnu(jb, nr, A + j0 * lda + j0, lda, A + j0 * lda + j0 + jb, lda, s);
      if (m - j0 - jb > 0)
        gemm_rm_nn(m - j0 - jb, nr, jb, -1.0, A + (j0 + jb) * lda + j0, lda, A + j0 * lda + j0 + jb,
                   lda, 1.0, A + (j0 + jb) * lda + j0 + jb, lda, s);
    }
  }
}

int leaf_error_flag(hipStream_t s) {
  int dev = 0;
  ZH_CHECK(hipGetDevice(&dev));
  void* ws = workspace(leaf_ws_bytes(), 3);
  LeafWS w = carve_leaf_ws(ws);
  int h = 0;
  ZH_CHECK(hipMemcpyAsync(&h, w.err, sizeof(int), hipMemcpyDeviceToHost, s));
  ZH_CHECK(hipStreamSynchronize(s));
  return h;
}

}  // namespace zmlhip
