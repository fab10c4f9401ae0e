// Low-precision GEMMs of the mixed-precision factorizations, FP32 accumulate.
//
//   C = beta * C + alpha * rs_i * cs_j * sum_k A_ik B_kj        (row-major)
//
// C is FP32. rs / cs are optional per-row / per-column multipliers (the
// inverse power-of-two scales applied when the operands were rounded to
// FP16/BF16, so that A_ik * rs_i and B_kj * cs_j stayed in range).
//   * FP16 / BF16: v_mfma_f32_32x32x16_{f16,bf16}. Both operands k-contiguous:
//     A is m x K row-major, B is given transposed (BT: n x K row-major).
//   * FP32: v_mfma_f32_32x32x2_f32. A is m x K row-major, B is K x n
//     row-major (its natural layout: one element per lane per MFMA).
// With `lower`, tiles strictly above the diagonal of C are skipped (SYRK-style
// updates of a lower triangle; the upper part of diagonal tiles is written).
// A 256-thread workgroup computes 128 x 128 of C (2 x 2 waves of 64 x 64,
// each 2 x 2 MFMA blocks of 32 x 32); K advances through double-buffered
// LDS. Lane l of a 32 x 32 block holds column l % 32 and rows
// 8 (r / 4) + 4 ((l / 32) % 2) + r % 4 of register r, so every C access is
// 32 consecutive floats.
#include "zmlhip_internal.hpp"
#include "tiles.hpp"
#include "mixed.hpp"

namespace zmlhip {
namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

constexpr int MT = 128, NTL = 128, NTH = 256;

__device__ __forceinline__ void tile_of(int64_t bid, int64_t ntm, int64_t ntn, int64_t& tm, int64_t& tn) {
  const int64_t nt = ntm * ntn;
  const int64_t x = bid & 7, q = nt >> 3, r8 = nt & 7;
  const int64_t pid = x * q + (x < r8 ? x : r8) + (bid >> 3);
  constexpr int64_t G = 8;
  const int64_t per = G * ntn, grp = pid / per, fm = grp * G;
  const int64_t gm = (ntm - fm) < G ? (ntm - fm) : G;
  const int64_t r = pid - grp * per;
  tm = fm + r % gm;
  tn = r / gm;
}

// C is read before the K loop (its HBM latency then overlaps the MFMA work:
// for a rank-256 update the C traffic, not the arithmetic, bounds the kernel)
// and combined with the accumulators at the end. Rows and columns beyond m, n
// are clamped to the last valid ones, so every load is in bounds and none sits
// behind a branch (the values read there are never stored).
__device__ __forceinline__ void load_c(f16v (&cv)[2][2], int64_t m, int64_t n, float beta,
                                       const float* __restrict__ C, int64_t ldc, int64_t r0, int64_t c0,
                                       int lane) {
  if (beta == 0.0f) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 16; ++r) cv[bi][bj][r] = 0.0f;
    return;
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int64_t j = min(c0 + 32 * bj + (lane & 31), n - 1);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = min(r0 + 32 * bi + 8 * (r >> 2) + 4 * ((lane >> 5) & 1) + (r & 3), m - 1);
        cv[bi][bj][r] = C[i * ldc + j];
      }
    }
}

// C = beta C + alpha rs_i cs_j acc on a wave's 64 x 64 block. The scales are
// all requested before the first store (clamped like the C loads), and a block
// that lies inside C is stored without per-element tests.
__device__ __forceinline__ void epilogue(const f16v (&acc)[2][2], const f16v (&cv)[2][2], int64_t m,
                                         int64_t n, float alpha, float beta, const float* __restrict__ rs,
                                         const float* __restrict__ cs, float* __restrict__ C,
                                         int64_t ldc, int64_t r0, int64_t c0, int lane) {
  float sc[2], sr[2][16];
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) sc[bj] = alpha;
  if (cs) {
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) sc[bj] = alpha * cs[min(c0 + 32 * bj + (lane & 31), n - 1)];
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int r = 0; r < 16; ++r) sr[bi][r] = 1.0f;
  if (rs) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        sr[bi][r] = rs[min(r0 + 32 * bi + 8 * (r >> 2) + 4 * ((lane >> 5) & 1) + (r & 3), m - 1)];
  }
  auto value = [&](int bi, int bj, int r) {
    const float s = sc[bj] * sr[bi][r];
    return beta == 0.0f ? s * acc[bi][bj][r] : beta * cv[bi][bj][r] + s * acc[bi][bj][r];
  };
  if (r0 + 64 <= m && c0 + 64 <= n) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int64_t j = c0 + 32 * bj + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t i = r0 + 32 * bi + 8 * (r >> 2) + 4 * ((lane >> 5) & 1) + (r & 3);
          C[i * ldc + j] = value(bi, bj, r);
        }
      }
    return;
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int64_t j = c0 + 32 * bj + (lane & 31);
      if (j >= n) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = r0 + 32 * bi + 8 * (r >> 2) + 4 * ((lane >> 5) & 1) + (r & 3);
        if (i < m) C[i * ldc + j] = value(bi, bj, r);
      }
    }
}

// ---- FP16 / BF16 ---------------------------------------------------------------
template <typename H>
struct MfmaH;
template <> struct MfmaH<_Float16> {
  typedef h8 vec;
  static __device__ __forceinline__ f16v mma(vec a, vec b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};
template <> struct MfmaH<__bf16> {
  typedef b8 vec;
  static __device__ __forceinline__ f16v mma(vec a, vec b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};

constexpr int BKH = 32, PADH = 8, LDH = BKH + PADH;   // halfs per LDS row

template <typename H>
__global__ void __launch_bounds__(NTH, 2)
mp_gemm_h_kernel(int64_t m, int64_t n, int64_t K, float alpha, const H* __restrict__ A, int64_t lda,
                 const H* __restrict__ BT, int64_t ldb, float beta, const float* __restrict__ rs,
                 const float* __restrict__ cs, float* __restrict__ C, int64_t ldc, int lower, int64_t tri_tiles) {
  typedef typename MfmaH<H>::vec V;
  __shared__ __attribute__((aligned(16))) H As[2][MT * LDH];
  __shared__ __attribute__((aligned(16))) H Bs[2][NTL * LDH];
  const int64_t ntm = (m + MT - 1) / MT, ntn = (n + NTL - 1) / NTL;
  int64_t tm, tn;
  if (tri_tiles > 0) tri_tile_coords(blockIdx.x, tri_tiles, ntm, ntn, kLower, tm, tn);
  else tile_of(blockIdx.x, ntm, ntn, tm, tn);
  const int64_t row0 = tm * MT, col0 = tn * NTL;
  if (lower && col0 > row0 + MT - 1) return;   // tile strictly above the diagonal
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // each thread stages 2 x 16 bytes of A and of B per k-slab: row t / 4 (+64), k chunk t % 4
  uint4 ra[2], rb[2];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int r = (t >> 2) + 64 * p, kc = (t & 3) * 8;
      const int64_t kk = k0 + kc;
      const int64_t ia = row0 + r, jb = col0 + r;
      uint4 va = {0, 0, 0, 0}, vb = {0, 0, 0, 0};
      if (ia < m) {
        if (kk + 8 <= K) va = *reinterpret_cast<const uint4*>(A + ia * lda + kk);
        else {
          H tmp[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) tmp[e] = (kk + e < K) ? A[ia * lda + kk + e] : H(0.0f);
          va = *reinterpret_cast<uint4*>(tmp);
        }
      }
      if (jb < n) {
        if (kk + 8 <= K) vb = *reinterpret_cast<const uint4*>(BT + jb * ldb + kk);
        else {
          H tmp[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) tmp[e] = (kk + e < K) ? BT[jb * ldb + kk + e] : H(0.0f);
          vb = *reinterpret_cast<uint4*>(tmp);
        }
      }
      ra[p] = va;
      rb[p] = vb;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int r = (t >> 2) + 64 * p, kc = (t & 3) * 8;
      *reinterpret_cast<uint4*>(&As[buf][r * LDH + kc]) = ra[p];
      *reinterpret_cast<uint4*>(&Bs[buf][r * LDH + kc]) = rb[p];
    }
  };

  f16v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

  f16v cv[2][2];
  load_c(cv, m, n, beta, C, ldc, row0 + wm * 64, col0 + wn * 64, lane);
  const int nk = (int)((K + BKH - 1) / BKH);
  load(0);
  store(0);
  __syncthreads();
  const int li = lane & 31, lk = (lane >> 5) * 8;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((int64_t)(kt + 1) * BKH);
#pragma unroll
    for (int ks = 0; ks < BKH / 16; ++ks) {
      V av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x)
        av[x] = *reinterpret_cast<const V*>(&As[cur][(wm * 64 + 32 * x + li) * LDH + 16 * ks + lk]);
#pragma unroll
      for (int y = 0; y < 2; ++y)
        bv[y] = *reinterpret_cast<const V*>(&Bs[cur][(wn * 64 + 32 * y + li) * LDH + 16 * ks + lk]);
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = MfmaH<H>::mma(av[x], bv[y], acc[x][y]);
    }
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
  epilogue(acc, cv, m, n, alpha, beta, rs, cs, C, ldc, row0 + wm * 64, col0 + wn * 64, lane);
}

// Same product with the operand slabs streamed straight into LDS
// (global_load_lds_dwordx4: a wave writes 1 KB, 16 rows of 64 bytes, lane l
// the 16 bytes at 16 l) through a ring of S stages: slab kt + S - 1 is
// requested while slab kt is multiplied, one barrier per slab and no
// registers for staging. Within a row the four 16-byte k-chunks are stored
// XOR-swizzled by (row >> 1) & 3 (each lane picks the global chunk that lands
// at its position), so the MFMA operand reads of 8 consecutive rows hit
// distinct banks. Needs K % 32 == 0 and 16-byte aligned rows; rows beyond m
// (columns beyond n) read the last valid row, whose products land only in C
// entries that are not stored. Operands and products are those of
// mp_gemm_h_kernel, in the same order.
// Defaults of the streamed kernel: ring stages, workgroups per CU the
// register budget is fitted to, and whether C is requested before the K loop
// or after it. Alone (m = n = 32256 and 16384, K = 512, FP16), three stages,
// three workgroups per CU and C read after the products move 3.03 and 3.11
// TB/s against 2.84 and 2.92 TB/s for four stages, two workgroups and C first;
// but inside the mixed LU the third workgroup per CU takes the slots the panel
// stream's kernels need, and the solve slows from 138 to 153 ms at n = 32768.
// Four workgroups per CU (two stages) fall to 1.6 TB/s.
constexpr int DL_STAGES = 4, DL_OCC = 2;
constexpr bool DL_PREFETCH_C = true;
__device__ __forceinline__ int dl_swz(int r) { return (r >> 1) & 3; }

template <int N>
__device__ __forceinline__ void wait_vm() {
  // s_waitcnt vmcnt(N), the other counters unconstrained (gfx9 encoding)
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

template <typename H, int S = DL_STAGES, int OCC = DL_OCC, bool PFC = DL_PREFETCH_C>
__global__ void __launch_bounds__(NTH, OCC)
mp_gemm_h_dl_kernel(int64_t m, int64_t n, int64_t K, float alpha, const H* __restrict__ A, int64_t lda,
                    const H* __restrict__ BT, int64_t ldb, float beta, const float* __restrict__ rs,
                    const float* __restrict__ cs, float* __restrict__ C, int64_t ldc, int lower,
                    int64_t tri_tiles) {
  typedef typename MfmaH<H>::vec V;
  constexpr int ROWB = 64;                 // bytes per LDS row: 32 halfs
  __shared__ __attribute__((aligned(1024))) unsigned char lds[S][2][MT * ROWB];
  const int64_t ntm = (m + MT - 1) / MT, ntn = (n + NTL - 1) / NTL;
  int64_t tm, tn;
  if (tri_tiles > 0) tri_tile_coords(blockIdx.x, tri_tiles, ntm, ntn, kLower, tm, tn);
  else tile_of(blockIdx.x, ntm, ntn, tm, tn);
  const int64_t row0 = tm * MT, col0 = tn * NTL;
  if (lower && col0 > row0 + MT - 1) return;   // tile strictly above the diagonal
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // wave w requests rows 32 w .. 32 w + 31 of both operands, two 16-row instructions each
  const H* ga[2];
  const H* gb[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int r = 32 * wave + 16 * j + (lane >> 2);
    const int c = (lane & 3) ^ dl_swz(r);
    const int64_t ia = row0 + r < m ? row0 + r : m - 1;
    const int64_t jb = col0 + r < n ? col0 + r : n - 1;
    ga[j] = A + ia * lda + 8 * c;
    gb[j] = BT + jb * ldb + 8 * c;
  }
  auto issue = [&](int kt, int st) {
    const int64_t ko = (int64_t)kt * BKH;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      __builtin_amdgcn_global_load_lds((const void*)(ga[j] + ko),
                                       (__attribute__((address_space(3))) void*)&lds[st][0][(32 * wave + 16 * j) * ROWB],
                                       16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(gb[j] + ko),
                                       (__attribute__((address_space(3))) void*)&lds[st][1][(32 * wave + 16 * j) * ROWB],
                                       16, 0, 0);
    }
  };

  f16v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
  f16v cv[2][2];
  if (PFC) load_c(cv, m, n, beta, C, ldc, row0 + wm * 64, col0 + wn * 64, lane);

  const int nk = (int)(K / BKH);
#pragma unroll
  for (int s0 = 0; s0 < S - 1; ++s0)
    if (s0 < nk) issue(s0, s0);
  const int li = lane & 31, hk = lane >> 5;
  for (int kt = 0; kt < nk; ++kt) {
    // this wave's part of slab kt has landed once at most the later slabs' requests are pending
    if (kt + S - 2 < nk) wait_vm<4 * (S - 2)>();
    else wait_vm<0>();
    // every wave's part of slab kt is in LDS and stage (kt - 1) % S is no longer read:
    // a bare barrier (the fence of __syncthreads would wait for every request in flight)
    asm volatile("s_barrier" ::: "memory");
    if (kt + S - 1 < nk) issue(kt + S - 1, (kt + S - 1) % S);
    const int st = kt % S;
#pragma unroll
    for (int ks = 0; ks < BKH / 16; ++ks) {
      const int c = 2 * ks + hk;   // logical 16-byte chunk of the MFMA operand
      V av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int r = wm * 64 + 32 * x + li;
        av[x] = *reinterpret_cast<const V*>(&lds[st][0][r * ROWB + 16 * (c ^ dl_swz(r))]);
      }
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const int r = wn * 64 + 32 * y + li;
        bv[y] = *reinterpret_cast<const V*>(&lds[st][1][r * ROWB + 16 * (c ^ dl_swz(r))]);
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = MfmaH<H>::mma(av[x], bv[y], acc[x][y]);
    }
  }
  if (!PFC) load_c(cv, m, n, beta, C, ldc, row0 + wm * 64, col0 + wn * 64, lane);
  epilogue(acc, cv, m, n, alpha, beta, rs, cs, C, ldc, row0 + wm * 64, col0 + wn * 64, lane);
}

// ---- FP32 ------------------------------------------------------------------------
constexpr int BKF = 16;
constexpr int LDAF = BKF + 1;        // A rows in LDS: [i][k], padded
constexpr int LDBF = NTL;            // B rows in LDS: [k][j]

// VEC: A, B 16-byte aligned with lda, ldb (and the K slice) multiples of 4, so
// whole 8-float runs inside the matrices load as two float4 each.
template <bool VEC>
__global__ void __launch_bounds__(NTH, 2)
mp_gemm_f_kernel(int64_t m, int64_t n, int64_t K, float alpha, const float* __restrict__ A,
                 int64_t lda, const float* __restrict__ B, int64_t ldb, float beta,
                 const float* __restrict__ rs, const float* __restrict__ cs, float* __restrict__ C,
                 int64_t ldc, int lower, int64_t tri_tiles, int64_t kslice, float* __restrict__ part) {
  if (part) {   // split K: slice blockIdx.y of K, plain partial product into part[slice] (m x n)
    const int64_t k0 = blockIdx.y * kslice;
    A += k0;
    B += k0 * ldb;
    K = K - k0 < kslice ? K - k0 : kslice;
    C = part + blockIdx.y * m * n;
    ldc = n;
    alpha = 1.0f;
    beta = 0.0f;
    rs = nullptr;
    cs = nullptr;
  }
  __shared__ float As[2][MT * LDAF];
  __shared__ float Bs[2][BKF * LDBF];
  const int64_t ntm = (m + MT - 1) / MT, ntn = (n + NTL - 1) / NTL;
  int64_t tm, tn;
  if (tri_tiles > 0) tri_tile_coords(blockIdx.x, tri_tiles, ntm, ntn, kLower, tm, tn);
  else tile_of(blockIdx.x, ntm, ntn, tm, tn);
  const int64_t row0 = tm * MT, col0 = tn * NTL;
  if (lower && col0 > row0 + MT - 1) return;   // tile strictly above the diagonal
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // A tile 128 x 16: thread -> row t / 2 (8 floats); B tile 16 x 128: row t / 16 (8 floats).
  // Every load stays in bounds: rows of A beyond m and columns of B beyond n
  // read the last valid one (their products land only in C entries that are not
  // stored); k beyond K reads the last valid k and is zeroed after the load.
  float ra[8], rb[8];
  auto load = [&](int64_t k0) {
    if (K <= 0) {   // nothing to load (the product is zero)
#pragma unroll
      for (int e = 0; e < 8; ++e) ra[e] = rb[e] = 0.0f;
      return;
    }
    const int r = t >> 1, kc = (t & 1) * 8;
    const int64_t ia = row0 + r < m ? row0 + r : m - 1;
    const float* pa = A + ia * lda;
    if (VEC && k0 + kc + 8 <= K) {
      const float4 v0 = *reinterpret_cast<const float4*>(pa + k0 + kc);
      const float4 v1 = *reinterpret_cast<const float4*>(pa + k0 + kc + 4);
      ra[0] = v0.x; ra[1] = v0.y; ra[2] = v0.z; ra[3] = v0.w;
      ra[4] = v1.x; ra[5] = v1.y; ra[6] = v1.z; ra[7] = v1.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t kk = k0 + kc + e;
        ra[e] = pa[kk < K ? kk : K - 1];
      }
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (k0 + kc + e >= K) ra[e] = 0.0f;
    }
    const int kr = t >> 4, jc = (t & 15) * 8;
    const int64_t kk = k0 + kr;
    const float* pb = B + (kk < K ? kk : K - 1) * ldb;
    if (VEC && col0 + jc + 8 <= n) {
      const float4 v0 = *reinterpret_cast<const float4*>(pb + col0 + jc);
      const float4 v1 = *reinterpret_cast<const float4*>(pb + col0 + jc + 4);
      rb[0] = v0.x; rb[1] = v0.y; rb[2] = v0.z; rb[3] = v0.w;
      rb[4] = v1.x; rb[5] = v1.y; rb[6] = v1.z; rb[7] = v1.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t j = col0 + jc + e;
        rb[e] = pb[j < n ? j : n - 1];
      }
    }
    if (kk >= K) {
#pragma unroll
      for (int e = 0; e < 8; ++e) rb[e] = 0.0f;
    }
  };
  auto store = [&](int buf) {
    const int r = t >> 1, kc = (t & 1) * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) As[buf][r * LDAF + kc + e] = ra[e];
    const int kr = t >> 4, jc = (t & 15) * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) Bs[buf][kr * LDBF + jc + e] = rb[e];
  };

  f16v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

  f16v cv[2][2];
  load_c(cv, m, n, beta, C, ldc, row0 + wm * 64, col0 + wn * 64, lane);
  const int nk = (int)((K + BKF - 1) / BKF);
  load(0);
  store(0);
  __syncthreads();
  const int li = lane & 31, lk = lane >> 5;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((int64_t)(kt + 1) * BKF);
#pragma unroll
    for (int ks = 0; ks < BKF / 2; ++ks) {
      float av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = As[cur][(wm * 64 + 32 * x + li) * LDAF + 2 * ks + lk];
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = Bs[cur][(2 * ks + lk) * LDBF + wn * 64 + 32 * y + li];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
  epilogue(acc, cv, m, n, alpha, beta, rs, cs, C, ldc, row0 + wm * 64, col0 + wn * 64, lane);
}

// mp_gemm_f_kernel's float4 loads apply: 16-byte aligned operands, leading
// dimensions (and the K offset of a split-K slice) multiples of 4 floats
bool f32_vec_ok(const float* A, int64_t lda, const float* B, int64_t ldb, int64_t kslice) {
  return lda % 4 == 0 && ldb % 4 == 0 && kslice % 4 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
         (reinterpret_cast<uintptr_t>(B) & 15) == 0;
}

// the streamed variant applies: whole 32-wide k-slabs, 16-byte aligned rows
bool dl_ok(int64_t K, const void* A, int64_t lda, const void* BT, int64_t ldb) {
  return K >= BKH && K % BKH == 0 && lda % 8 == 0 && ldb % 8 == 0 &&
         (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(BT) & 15) == 0;
}

}  // namespace

void mp_gemm_f16(int64_t m, int64_t n, int64_t K, float alpha, const _Float16* A, int64_t lda,
                 const _Float16* BT, int64_t ldb, float beta, const float* rs, const float* cs,
                 float* C, int64_t ldc, hipStream_t s, int lower) {
  if (m <= 0 || n <= 0) return;
  const int64_t ntm = (m + MT - 1) / MT, ntn = (n + NTL - 1) / NTL;
  const int64_t tri_tiles = lower ? tri_tile_count(kLower, ntm, ntn) : 0;
  const int64_t tiles = tri_tiles > 0 ? tri_tiles : ntm * ntn;
  if (dl_ok(K, A, lda, BT, ldb))
    hipLaunchKernelGGL((mp_gemm_h_dl_kernel<_Float16>), dim3((unsigned)tiles), dim3(NTH), 0, s, m, n, K, alpha,
                       A, lda, BT, ldb, beta, rs, cs, C, ldc, lower, tri_tiles);
  else
    hipLaunchKernelGGL((mp_gemm_h_kernel<_Float16>), dim3((unsigned)tiles), dim3(NTH), 0, s, m, n, K, alpha,
                       A, lda, BT, ldb, beta, rs, cs, C, ldc, lower, tri_tiles);
  ZH_LAUNCH_CHECK();
}

void mp_gemm_bf16(int64_t m, int64_t n, int64_t K, float alpha, const __bf16* A, int64_t lda,
                  const __bf16* BT, int64_t ldb, float beta, const float* rs, const float* cs,
                  float* C, int64_t ldc, hipStream_t s, int lower) {
  if (m <= 0 || n <= 0) return;
  const int64_t ntm = (m + MT - 1) / MT, ntn = (n + NTL - 1) / NTL;
  const int64_t tri_tiles = lower ? tri_tile_count(kLower, ntm, ntn) : 0;
  const int64_t tiles = tri_tiles > 0 ? tri_tiles : ntm * ntn;
  if (dl_ok(K, A, lda, BT, ldb))
    hipLaunchKernelGGL((mp_gemm_h_dl_kernel<__bf16>), dim3((unsigned)tiles), dim3(NTH), 0, s, m, n, K, alpha,
                       A, lda, BT, ldb, beta, rs, cs, C, ldc, lower, tri_tiles);
  else
    hipLaunchKernelGGL((mp_gemm_h_kernel<__bf16>), dim3((unsigned)tiles), dim3(NTH), 0, s, m, n, K, alpha,
                       A, lda, BT, ldb, beta, rs, cs, C, ldc, lower, tri_tiles);
  ZH_LAUNCH_CHECK();
}

void mp_gemm_f32(int64_t m, int64_t n, int64_t K, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, const float* rs, const float* cs, float* C,
                 int64_t ldc, hipStream_t s, int lower) {
  if (m <= 0 || n <= 0) return;
  const int64_t ntm = (m + MT - 1) / MT, ntn = (n + NTL - 1) / NTL;
  const int64_t tri_tiles = lower ? tri_tile_count(kLower, ntm, ntn) : 0;
  const int64_t tiles = tri_tiles > 0 ? tri_tiles : ntm * ntn;
  if (f32_vec_ok(A, lda, B, ldb, K))
    hipLaunchKernelGGL(mp_gemm_f_kernel<true>, dim3((unsigned)tiles), dim3(NTH), 0, s, m, n, K, alpha, A, lda,
                       B, ldb, beta, rs, cs, C, ldc, lower, tri_tiles, (int64_t)0, (float*)nullptr);
  else
    hipLaunchKernelGGL(mp_gemm_f_kernel<false>, dim3((unsigned)tiles), dim3(NTH), 0, s, m, n, K, alpha, A, lda,
                       B, ldb, beta, rs, cs, C, ldc, lower, tri_tiles, (int64_t)0, (float*)nullptr);
  ZH_LAUNCH_CHECK();
}

namespace {
// C = alpha sum_s P[s] + beta C (slices summed in order)
__global__ void splitk_reduce_kernel(int64_t m, int64_t n, int slices, const float* __restrict__ P, float alpha,
                                     float beta, float* __restrict__ C, int64_t ldc) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= m * n) return;
  const int64_t i = e / n, j = e - i * n;
  float acc = 0.0f;
  for (int sl = 0; sl < slices; ++sl) acc += P[sl * m * n + e];
  float* c = C + i * ldc + j;
  *c = beta == 0.0f ? alpha * acc : beta * *c + alpha * acc;
}
}  // namespace

int64_t mp_gemm_f32_splitk_floats(int64_t m, int64_t n, int64_t K) {
  const int64_t slices = std::max<int64_t>(1, std::min<int64_t>(8, K / 64));
  return slices * m * n;
}

void mp_gemm_f32_splitk(int64_t m, int64_t n, int64_t K, float alpha, const float* A, int64_t lda, const float* B,
                        int64_t ldb, float beta, float* C, int64_t ldc, float* part, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  const int64_t slices = std::max<int64_t>(1, std::min<int64_t>(8, K / 64));
  const int64_t kslice = (K + slices - 1) / slices;
  const int64_t ntm = (m + MT - 1) / MT, ntn = (n + NTL - 1) / NTL;
  if (f32_vec_ok(A, lda, B, ldb, kslice))
    hipLaunchKernelGGL(mp_gemm_f_kernel<true>, dim3((unsigned)(ntm * ntn), (unsigned)slices), dim3(NTH), 0, s, m, n,
                       K, 1.0f, A, lda, B, ldb, 0.0f, (const float*)nullptr, (const float*)nullptr, C, ldc, 0,
                       (int64_t)0, kslice, part);
  else
    hipLaunchKernelGGL(mp_gemm_f_kernel<false>, dim3((unsigned)(ntm * ntn), (unsigned)slices), dim3(NTH), 0, s, m,
                       n, K, 1.0f, A, lda, B, ldb, 0.0f, (const float*)nullptr, (const float*)nullptr, C, ldc, 0,
                       (int64_t)0, kslice, part);
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((m * n + 255) / 256)), dim3(256), 0, s, m, n,
                     (int)slices, (const float*)part, alpha, beta, C, ldc);
  ZH_LAUNCH_CHECK();
}

}  // namespace zmlhip
