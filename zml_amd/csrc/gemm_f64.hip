// FP64 / FP32 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64 and
// v_mfma_f32_16x16x4_f32), column-major view.
//
//   C = alpha * op(A) * op(B) + beta * C        (reference blas/gemm.zig:68-1033)
//
// Row-major calls are mapped here by swapping the operands and m/n, exactly
// as zml does (gemm.zig:48-65). Only the arithmetic differs from zml: the
// matrix cores accumulate with fused multiply-adds, so results agree with
// zml's separately rounded loop to within the usual gamma_k |A||B| bound
// (bitwise on integer-valued data such as the reference's golden vectors).
//
// Tiling: a 256-thread workgroup computes a 128 x 128 tile of C; its four
// waves each own 64 x 64 (4 x 4 MFMA tiles of 16 x 16). K advances in slabs
// of 16 staged through double-buffered LDS (both operands stored k-major,
// so every MFMA operand fetch is a 16-lane contiguous read). The MFMA is
// issued as C^T = op(B)^T op(A)^T so that the 16 lanes of an output group
// hold 16 consecutive rows of one column of C: stores and beta-loads are
// 128-byte contiguous in column-major memory. The two element types differ
// only in the MFMA and in which output row a lane's accumulator r holds
// (FP64: lane / 16 + 4 r; FP32: 4 (lane / 16) + r).
//
// FP64 products whose shapes allow it (whole k-slabs, 16-byte aligned operand
// columns) take a second kernel with the same tile and arithmetic whose
// operand slabs go from global memory to LDS by the load unit
// (gemm_dma_f64_kernel below): 53.6 -> 64.4 TF/s for the K = 512 update of a
// 16384 x 16384 matrix, 59.5 -> 71.4 TF/s for 8192^3 on MI355X, same bits.
#include <cstdint>
#include <type_traits>

#include "zmlhip_internal.hpp"
#include "tiles.hpp"

namespace zmlhip {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// The MFMA of each element type and its accumulator layout.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 acc;
  static __device__ __forceinline__ acc step(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // column (within the 16 x 16 output tile) of accumulator r of a lane group lk
  static __device__ __forceinline__ int col(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma<float> {
  typedef f4 acc;
  static __device__ __forceinline__ acc step(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int lk, int r) { return 4 * lk + r; }
};

constexpr int BM = 128, BN = 128, BK = 16, PAD = 2;
constexpr int LDA_S = BM + PAD, LDB_S = BN + PAD;
constexpr int NT = 256;

// Maps the linear workgroup id to a (tile_m, tile_n) pair: consecutive ids
// land on different XCDs (dispatch is round-robin over 8 XCDs), so the ids
// are first regrouped so each XCD gets a contiguous range of tiles, and the
// range is walked in groups of 8 tile-rows to share operand panels in L2.
__device__ __forceinline__ void tile_coords(int64_t bid, int64_t ntm, int64_t ntn, int64_t& tm,
                                            int64_t& tn) {
  const int64_t nt = ntm * ntn;
  // XCD x receives the ids bid = x, x + 8, ...: give it the contiguous range
  // [x * q + min(x, r), ...) of tiles, q = nt / 8, r = nt % 8
  const int64_t x = bid & 7, q = nt >> 3, r8 = nt & 7;
  const int64_t pid = x * q + (x < r8 ? x : r8) + (bid >> 3);
  constexpr int64_t G = 8;
  const int64_t per_group = G * ntn;
  const int64_t group = pid / per_group;
  const int64_t first_m = group * G;
  const int64_t gm = (ntm - first_m) < G ? (ntm - first_m) : G;
  const int64_t r = pid - group * per_group;
  tm = first_m + r % gm;
  tn = r / gm;
}

template <typename T, bool TA, bool TB>
__global__ void __launch_bounds__(NT, 2)
gemm_kernel(int64_t m, int64_t n, int64_t k, T alpha, const T* __restrict__ A,
            int64_t lda, const T* __restrict__ B, int64_t ldb, T beta,
            T* __restrict__ C, int64_t ldc, int tri, int64_t tri_tiles) {
  typedef typename Mfma<T>::acc acc_t;
  __shared__ T As[2][BK * LDA_S];
  __shared__ T Bs[2][BK * LDB_S];

  const int64_t ntm = (m + BM - 1) / BM, ntn = (n + BN - 1) / BN;
  int64_t tm, tn;
  if (tri_tiles > 0) tri_tile_coords(blockIdx.x, tri_tiles, ntm, ntn, tri, tm, tn);
  else tile_coords(blockIdx.x, ntm, ntn, tm, tn);
  const int64_t row0 = tm * BM, col0 = tn * BN;
  if (tri == kLower && row0 + BM - 1 < col0) return;   // tile strictly above the diagonal
  if (tri == kUpper && col0 + BN - 1 < row0) return;   // tile strictly below the diagonal

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // ---- global -> register staging (8 elements of A and 8 of B per thread)
  T ra[8], rb[8];
  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int64_t i, l;
      if (!TA) { i = row0 + (t & 127); l = k0 + (t >> 7) + 2 * q; }   // m-contiguous
      else     { l = k0 + (t & 15);    i = row0 + (t >> 4) + 16 * q; } // k-contiguous
      const bool ok = i < m && l < k;
      ra[q] = ok ? (TA ? A[l + i * lda] : A[i + l * lda]) : T(0);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int64_t j, l;
      if (TB) { j = col0 + (t & 127); l = k0 + (t >> 7) + 2 * q; }    // n-contiguous
      else    { l = k0 + (t & 15);    j = col0 + (t >> 4) + 16 * q; }  // k-contiguous
      const bool ok = j < n && l < k;
      rb[q] = ok ? (TB ? B[j + l * ldb] : B[l + j * ldb]) : T(0);
    }
  };
  // Update form C <- C +- op(A) op(B) (beta 1, alpha +-1): the accumulators start
  // from C, loaded with the first operand slab, and B is staged negated for
  // alpha -1, so the epilogue only stores.
  const bool cinit = beta == T(1) && (alpha == T(1) || alpha == T(-1));
  const bool negb = cinit && alpha == T(-1);
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int i, l;
      if (!TA) { i = t & 127; l = (t >> 7) + 2 * q; }
      else     { l = t & 15;  i = (t >> 4) + 16 * q; }
      As[buf][l * LDA_S + i] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int j, l;
      if (TB) { j = t & 127; l = (t >> 7) + 2 * q; }
      else    { l = t & 15;  j = (t >> 4) + 16 * q; }
      Bs[buf][l * LDB_S + j] = negb ? -rb[q] : rb[q];
    }
  };

  const int lr = lane & 15, lk = lane >> 4;
  // lane holds C[i0 + lr][j0 + Mfma<T>::col(lk, r)] of each 16 x 16 tile
  auto c_ok = [&](int64_t i, int64_t j) {
    return i < m && j < n && !(tri == kLower && i < j) && !(tri == kUpper && i > j);
  };
  acc_t acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = acc_t{T(0), T(0), T(0), T(0)};
  if (cinit) {
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int64_t i = row0 + wm * 64 + 16 * x + lr;
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t j = col0 + wn * 64 + 16 * y + Mfma<T>::col(lk, r);
          if (c_ok(i, j)) acc[x][y][r] = C[i + j * ldc];
        }
    }
  }

  const int nk = (int)((k + BK - 1) / BK);
  load_tiles(0);
  store_tiles(0);
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tiles((int64_t)(kt + 1) * BK);
    const T* as = As[cur] + wm * 64 + lr;
    const T* bs = Bs[cur] + wn * 64 + lr;
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = 4 * ks + lk;
      T av[4], bv[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) av[x] = as[kr * LDA_S + 16 * x];
#pragma unroll
      for (int y = 0; y < 4; ++y) bv[y] = bs[kr * LDB_S + 16 * y];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
          acc[x][y] = Mfma<T>::step(bv[y], av[x], acc[x][y]);
    }
    if (kt + 1 < nk) store_tiles(cur ^ 1);
    __syncthreads();
  }

  // ---- epilogue
  if (cinit) {
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int64_t i = row0 + wm * 64 + 16 * x + lr;
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t j = col0 + wn * 64 + 16 * y + Mfma<T>::col(lk, r);
          if (c_ok(i, j)) C[i + j * ldc] = acc[x][y][r];
        }
    }
    return;
  }
  // general alpha, beta: per row tile the 16 C values are all requested before any
  // is used (one memory latency per row tile instead of one per element)
  const bool need_c = beta != T(0);
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const int64_t i = row0 + wm * 64 + 16 * x + lr;
    bool ok[4][4];
    T cv[4][4];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t j = col0 + wn * 64 + 16 * y + Mfma<T>::col(lk, r);
        ok[y][r] = i < m && j < n && !(tri == kLower && i < j) && !(tri == kUpper && i > j);
        cv[y][r] = (need_c && ok[y][r]) ? C[i + j * ldc] : T(0);
      }
    }
#pragma unroll
    for (int y = 0; y < 4; ++y) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (!ok[y][r]) continue;
        const int64_t j = col0 + wn * 64 + 16 * y + Mfma<T>::col(lk, r);
        const T v = alpha * acc[x][y][r];
        C[i + j * ldc] = beta == T(0) ? v : (beta == T(1) ? cv[y][r] + v : beta * cv[y][r] + v);
      }
    }
  }
}

// ---- FP64 with the operand slabs moved global -> LDS by the load unit
//
// Same tile, waves, MFMA issue order and epilogue as gemm_kernel (its results
// are identical bit for bit), but every thread issues 8 global_load_lds_dwordx4
// per 16-wide k-slab: each lane's 16 bytes land in LDS at 16 x its lane id
// after the instruction's (wave-uniform) base, without passing through VGPRs
// and without ds_write instructions. Two LDS buffers: slab kt + 1 is in flight
// while slab kt is multiplied; one barrier per slab, after each thread's own
// loads have landed (vmcnt 0).
// Slab layouts in LDS (doubles):
//  - operand contiguous along m (A) or n (op(B) = B^T): 16 k-rows of 128
//    elements, k-row l at l * LDA_S; a wave instruction moves one k-row (lane:
//    elements 2 lane, 2 lane + 1, so the m / n extent must be even);
//  - operand contiguous along k: 128 columns of 16, column c at 16 c, k pairs
//    stored at slot pair ^ sw(c), sw(c) = (c ^ c >> 3) & 7, so that the 16
//    lanes reading one k of 16 consecutive columns hit distinct LDS banks; a
//    wave instruction moves 8 columns (lane 8 jj + s: column 8 q + jj, the pair
//    s ^ sw).
// For the update form with alpha = -1 the A fragments are negated after their
// LDS reads (the register-staged kernel negates B before storing it): either
// way every product enters the accumulator as fl(c - a b).
constexpr int DBK = 16;
constexpr int DSLAB = DBK * LDA_S;   // doubles per operand buffer (>= 128 x 16)

typedef __attribute__((address_space(3))) void* lds_void_ptr;

__device__ __forceinline__ void lds_dma16(const double* g, double* l) {
  __builtin_amdgcn_global_load_lds((const void*)g, (lds_void_ptr)l, 16, 0, 0);
}

// One operand's slab mover and fragment reader. MNC: contiguous along m / n.
template <bool MNC>
struct DmaOperand {
  const double* p[4];   // per-thread source of each of the thread's 4 instructions (k0 = 0)
  int64_t ld;
  int o[4];             // K-contiguous: LDS offset of fragment x at ks = 0
  int base;             // MN-contiguous: LDS offset of fragment 0 at ks = 0

  __device__ __forceinline__ void init(const double* X, int64_t ldx, int64_t mn0, int64_t mn, int wave, int lane,
                                       int w, int lr, int lk) {
    ld = ldx;
    if (MNC) {
      const int64_t e = mn0 + 2 * lane < mn ? mn0 + 2 * lane : mn - 2;
#pragma unroll
      for (int c = 0; c < 4; ++c) p[c] = X + e + (int64_t)(wave + 4 * c) * ldx;
      base = w * 64 + lr + lk * LDA_S;
    } else {
      const int jj = lane >> 3, s = lane & 7;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int q = wave + 4 * c;
        const int64_t col = mn0 + 8 * q + jj < mn ? mn0 + 8 * q + jj : mn - 1;
        p[c] = X + 2 * (s ^ ((jj ^ q) & 7)) + col * ldx;
      }
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int c = w * 64 + 16 * x + lr;
        o[x] = c * DBK + 2 * ((lk >> 1) ^ ((c ^ (c >> 3)) & 7)) + (lk & 1);
      }
    }
  }
  __device__ __forceinline__ void issue(int64_t k0, double* buf, int wave) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (MNC) lds_dma16(p[c] + k0 * ld, buf + (wave + 4 * c) * LDA_S);
      else lds_dma16(p[c] + k0, buf + (wave + 4 * c) * 128);
    }
  }
  // fragment x (16 elements along m / n) at k = 4 ks + lk
  __device__ __forceinline__ double frag(const double* buf, int ks, int x) const {
    return MNC ? buf[base + 4 * ks * LDA_S + 16 * x] : buf[o[x] ^ (4 * ks)];
  }
};

template <bool TA, bool TB>
__global__ void __launch_bounds__(NT, 2)
gemm_dma_f64_kernel(int64_t m, int64_t n, int64_t k, double alpha, const double* __restrict__ A, int64_t lda,
                    const double* __restrict__ B, int64_t ldb, double beta, double* __restrict__ C, int64_t ldc,
                    int tri, int64_t tri_tiles) {
  __shared__ __attribute__((aligned(16))) double As0[DSLAB];
  __shared__ __attribute__((aligned(16))) double As1[DSLAB];
  __shared__ __attribute__((aligned(16))) double Bs0[DSLAB];
  __shared__ __attribute__((aligned(16))) double Bs1[DSLAB];

  const int64_t ntm = (m + BM - 1) / BM, ntn = (n + BN - 1) / BN;
  int64_t tm, tn;
  if (tri_tiles > 0) tri_tile_coords(blockIdx.x, tri_tiles, ntm, ntn, tri, tm, tn);
  else tile_coords(blockIdx.x, ntm, ntn, tm, tn);
  const int64_t row0 = tm * BM, col0 = tn * BN;
  if (tri == kLower && row0 + BM - 1 < col0) return;
  if (tri == kUpper && col0 + BN - 1 < row0) return;

  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;

  DmaOperand<!TA> opa;
  DmaOperand<TB> opb;
  opa.init(A, lda, row0, m, wave, lane, wm, lr, lk);
  opb.init(B, ldb, col0, n, wave, lane, wn, lr, lk);

  const bool cinit = beta == 1.0 && (alpha == 1.0 || alpha == -1.0);
  const unsigned long long sgn = (cinit && alpha == -1.0) ? 0x8000000000000000ull : 0ull;
  auto c_ok = [&](int64_t i, int64_t j) {
    return i < m && j < n && !(tri == kLower && i < j) && !(tri == kUpper && i > j);
  };
  d4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
  if (cinit) {
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int64_t i = row0 + wm * 64 + 16 * x + lr;
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t j = col0 + wn * 64 + 16 * y + Mfma<double>::col(lk, r);
          if (c_ok(i, j)) acc[x][y][r] = C[i + j * ldc];
        }
    }
  }

  auto compute = [&](const double* as, const double* bs) {
#pragma unroll
    for (int ks = 0; ks < DBK / 4; ++ks) {
      double av[4], bv[4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
        av[x] = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, opa.frag(as, ks, x)) ^ sgn);
#pragma unroll
      for (int y = 0; y < 4; ++y) bv[y] = opb.frag(bs, ks, y);
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = Mfma<double>::step(bv[y], av[x], acc[x][y]);
    }
  };
  // vmcnt(0), expcnt and lgkmcnt left at their maxima
  constexpr int kWaitVm0 = 0x0F70;
  const int nk = (int)(k / DBK);
  opa.issue(0, As0, wave);
  opb.issue(0, Bs0, wave);
  for (int kt = 0; kt < nk; kt += 2) {
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    if (kt + 1 < nk) {
      opa.issue((int64_t)(kt + 1) * DBK, As1, wave);
      opb.issue((int64_t)(kt + 1) * DBK, Bs1, wave);
    }
    compute(As0, Bs0);
    if (kt + 1 >= nk) break;
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    if (kt + 2 < nk) {
      opa.issue((int64_t)(kt + 2) * DBK, As0, wave);
      opb.issue((int64_t)(kt + 2) * DBK, Bs0, wave);
    }
    compute(As1, Bs1);
  }

  if (cinit) {
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int64_t i = row0 + wm * 64 + 16 * x + lr;
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t j = col0 + wn * 64 + 16 * y + Mfma<double>::col(lk, r);
          if (c_ok(i, j)) C[i + j * ldc] = acc[x][y][r];
        }
    }
    return;
  }
  const bool need_c = beta != 0.0;
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const int64_t i = row0 + wm * 64 + 16 * x + lr;
    bool ok[4][4];
    double cv[4][4];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t j = col0 + wn * 64 + 16 * y + Mfma<double>::col(lk, r);
        ok[y][r] = c_ok(i, j);
        cv[y][r] = (need_c && ok[y][r]) ? C[i + j * ldc] : 0.0;
      }
    }
#pragma unroll
    for (int y = 0; y < 4; ++y) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (!ok[y][r]) continue;
        const int64_t j = col0 + wn * 64 + 16 * y + Mfma<double>::col(lk, r);
        const double v = alpha * acc[x][y][r];
        C[i + j * ldc] = beta == 0.0 ? v : (beta == 1.0 ? cv[y][r] + v : beta * cv[y][r] + v);
      }
    }
  }
}

// Whether gemm_dma_f64_kernel takes the product: whole 16-wide k-slabs, 16-byte
// aligned operand columns, and an even extent along each operand's contiguous
// m / n direction.
inline bool dma_f64_ok(bool ta, bool tb, int64_t m, int64_t n, int64_t k, const void* a, int64_t lda,
                       const void* b, int64_t ldb) {
  if (k < DBK || k % DBK != 0) return false;
  if ((((uintptr_t)a) | ((uintptr_t)b)) & 15) return false;
  if ((lda | ldb) & 1) return false;
  if (!ta && (m & 1)) return false;
  if (tb && (n & 1)) return false;
  return true;
}

template <typename T>
__global__ void scale_kernel(int64_t m, int64_t n, T beta, T* __restrict__ c, int64_t ldc, int tri) {
  const int64_t total = m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % m, j = e / m;
    if (tri == kLower && i < j) continue;
    if (tri == kUpper && i > j) continue;
    T* p = c + i + j * ldc;
    *p = beta == T(0) ? T(0) : *p * beta;
  }
}

}  // namespace

template <typename T>
void scale_cm(int64_t m, int64_t n, T beta, T* c, int64_t ldc, hipStream_t s, int tri) {
  if (m <= 0 || n <= 0 || beta == T(1)) return;
  int64_t total = m * n;
  int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks), dim3(256), 0, s, m, n, beta, c, ldc, tri);
  ZH_LAUNCH_CHECK();
}
template void scale_cm<double>(int64_t, int64_t, double, double*, int64_t, hipStream_t, int);
template void scale_cm<float>(int64_t, int64_t, float, float*, int64_t, hipStream_t, int);

template <typename T>
void gemm_cm(int transa, int transb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda,
             const T* b, int64_t ldb, T beta, T* c, int64_t ldc, hipStream_t s, int tri) {
  if (m <= 0 || n <= 0) return;
  if (alpha == T(0) || k <= 0) {  // gemm.zig:104-140: only the beta scaling remains
    scale_cm<T>(m, n, beta, c, ldc, s, tri);
    return;
  }
  const bool ta = !is_notrans(transa), tb = !is_notrans(transb);
  const int64_t ntm = (m + BM - 1) / BM, ntn = (n + BN - 1) / BN;
  const int64_t tri_tiles = tri_tile_count(tri, ntm, ntn);
  const int64_t tiles = tri_tiles > 0 ? tri_tiles : ntm * ntn;
  dim3 grid((unsigned)tiles), block(NT);
  if constexpr (std::is_same<T, double>::value) {
    if (gemm_dma_enabled() && dma_f64_ok(ta, tb, m, n, k, a, lda, b, ldb)) {
      if (!ta && !tb) hipLaunchKernelGGL((gemm_dma_f64_kernel<false, false>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
      else if (ta && !tb) hipLaunchKernelGGL((gemm_dma_f64_kernel<true, false>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
      else if (!ta && tb) hipLaunchKernelGGL((gemm_dma_f64_kernel<false, true>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
      else hipLaunchKernelGGL((gemm_dma_f64_kernel<true, true>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
      ZH_LAUNCH_CHECK();
      return;
    }
  }
  if (!ta && !tb) hipLaunchKernelGGL((gemm_kernel<T, false, false>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
  else if (ta && !tb) hipLaunchKernelGGL((gemm_kernel<T, true, false>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
  else if (!ta && tb) hipLaunchKernelGGL((gemm_kernel<T, false, true>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
  else hipLaunchKernelGGL((gemm_kernel<T, true, true>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tri, tri_tiles);
  ZH_LAUNCH_CHECK();
}
template void gemm_cm<double>(int, int, int64_t, int64_t, int64_t, double, const double*, int64_t, const double*,
                              int64_t, double, double*, int64_t, hipStream_t, int);
template void gemm_cm<float>(int, int, int64_t, int64_t, int64_t, float, const float*, int64_t, const float*,
                             int64_t, float, float*, int64_t, hipStream_t, int);

void scale_f64_cm(int64_t m, int64_t n, double beta, double* c, int64_t ldc, hipStream_t s, int tri) {
  scale_cm<double>(m, n, beta, c, ldc, s, tri);
}

void gemm_f64_cm(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                 double* c, int64_t ldc, hipStream_t s, int tri) {
  gemm_cm<double>(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, tri);
}

}  // namespace zmlhip
