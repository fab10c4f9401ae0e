// One-launch panel factorization of the FP64 LU (zml lapack/getrf2.zig:14-210
// for a whole NB-wide block column).
//
// The block column (rows [0, M) x columns [0, W), row-major, W <= 512) is
// factored by one persistent grid whose threads each own one panel row for
// the whole launch, instead of one launch per 64-column leaf with the
// interchanges, the triangular solve and the update of the rest of the block
// column as separate launches in between.
//
// Rows never move during the launch ("virtual" partial pivoting): every row
// carries its current LAPACK position. At column k the candidates are the
// rows at positions >= k; the pivot is the largest |a_ik|, ties going to the
// lowest position (iamax.zig:32's strict '>', applied to the current row
// order exactly as zml's interchanged rows present it); the pivot row takes
// position k and the row that held position k takes the pivot's old
// position, which is the interchange LAPACK records (ipiv[k] = old position
// + 1). A NaN at position k, or no finite candidate, keeps position k; so does
// an exactly zero pivot (getrf2.zig:55, info records it). Multipliers are
// fl(a * fl(1/p)) when |p| >= sfmin and a / p otherwise (getrf2.zig:65-87),
// updates separately rounded: the per-element operation sequence of zml's
// unblocked LU inside every 64-column leaf.
//
// Per leaf of 64 columns:
//   1. the column steps, each one cross-workgroup exchange of tagged words
//      (the protocol of panel.hpp's leaf kernel, with positions in the keys);
//   2. every row writes its leaf columns back; one grid barrier;
//   3. every workgroup forms U12 = L11^-1 A12 of the leaf's pivot rows for
//      the block column's remaining columns (64 at a time, in LDS; each
//      workgroup computes the same values), the pivot rows' owners keep it
//      aside (ubuf), and every row still unpivoted applies A22 -= L21 U12 to
//      itself.
// Other workgroups may still be reading a pivot row's A12 while its owner has
// its U12, so U12 reaches the pivot rows only at the end: by then every
// workgroup has taken part in a later column step, which it enters only after
// its last read.
// At the end the grid emits the moves that put the rows in LAPACK order
// (destination position <- physical row, absolute row numbers, as
// apply_perm_kernel reads them), the interchanges and info. The caller
// applies the moves to the block column and to the rest of the matrix.
#pragma once

#include "panel.hpp"

namespace zmlhip {
namespace panel {

constexpr int FW = 512;                 // widest block column
constexpr int FREC = 4 + 2 * W;         // candidate / position-k record words
// Rows (= threads) per workgroup: 512, one row each in 248 registers, so a
// CU holds one workgroup (two waves per SIMD) and n = 32768 needs 64 of them.
constexpr int FPT = 512;
// Dynamic LDS the launch adds to the kernel's own (about 77 KB): with it a
// panel workgroup takes more than 160 KB - 66.5 KB of the CU's LDS, so it never
// shares its CU with another panel workgroup or with a trailing-update GEMM
// workgroup (66.5 KB), whose waves would otherwise slow every column step
// (its registers, 256 per lane at two waves per SIMD, already fill the CU's).
constexpr int FUSED_PAD_LDS = 24 * 1024;

struct FusedWS {
  uint64_t* val;    // [2][GMAX][3]  |a| (two words) and position of each workgroup's candidate
  uint64_t* rec;    // [2][GMAX][FREC] candidate: (word 0) position, (1) physical row, leaf values
  uint64_t* rowk;   // [2][FREC]       row at position k: (word 1) NaN flag, (2) physical row, leaf values
  uint64_t* ubuf;   // [FW][FW]        U12 of the pivot row at each position (one value per word)
  unsigned* bar;    // [2] grid barrier: arrivals, generation
  int* err;
};

inline size_t fused_ws_bytes() {
  return sizeof(uint64_t) * (2 * GMAX * 3 + 2 * GMAX * FREC + 2 * FREC + (size_t)FW * FW) + 512;
}
inline FusedWS fused_carve(void* base) {
  FusedWS w;
  uint64_t* p = static_cast<uint64_t*>(base);
  w.val = p; p += 2 * GMAX * 3;
  w.rec = p; p += 2 * GMAX * FREC;
  w.rowk = p; p += 2 * FREC;
  w.ubuf = p; p += (size_t)FW * FW;
  w.bar = reinterpret_cast<unsigned*>(p);
  w.err = reinterpret_cast<int*>(p) + 16;
  return w;
}

// Grid-wide barrier of a fully resident grid (generation counting). Every
// workgroup's global stores before the barrier are visible to every
// workgroup's loads after it (release / acquire at agent scope around the
// arrival). A wait that exceeds spin_limit polls sets *err and *abort.
__device__ __forceinline__ void grid_barrier(unsigned* bar, int G, long spin_limit, int* err, int* abort) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned gen = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == (unsigned)G - 1) {
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar + 1, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (spin_limit < 0 || ++spins > spin_limit) {
          *abort = 1;
          atomicExch(err, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
  }
  __syncthreads();
}

// (|a|, position) maximum: larger |a| first, then the lower position.
__device__ __forceinline__ bool better(double v2, int p2, double v, int p) {
  return v2 > v || (v2 == v && p2 < p);
}

template <typename T>
__global__ void __launch_bounds__(FPT, 1)
panel_fused_kernel(int64_t M, int Wd, T* __restrict__ P, int64_t ld, int32_t* __restrict__ ipiv,
                   int64_t row_base, int* __restrict__ info, int64_t info_base, int32_t* __restrict__ pairs,
                   FusedWS ws, uint32_t tag_base, long spin_limit) {
#pragma clang fp contract(off)
  constexpr int PT = FPT, NW = PT / 64, CH = 64, TPC = PT / CH;   // TPC: threads per U12 column
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int64_t i = (int64_t)g * PT + t;
  const bool own = i < M;

  __shared__ T u[W], rk[W];
  __shared__ T s_v[NW];
  __shared__ int s_pos[NW];
  __shared__ int64_t s_phys[NW];
  __shared__ T stage[NW][W], stage_k[W];
  __shared__ int s_knan, s_abort;
  __shared__ int s_ppos, s_fz;
  __shared__ int64_t s_pphys;
  __shared__ int32_t s_ipiv[FW];
  __shared__ int64_t s_prow[FW];          // physical row of the pivot at each position
  __shared__ T L11[W][W + 1];
  __shared__ T U12[W][CH + 1];

  int pos = (int)i;                       // current LAPACK position of this row
  const T sfmin = sizeof(T) == 8 ? T(0x1p-1022) : T(0x1p-126);
  if (t == 0) { s_abort = 0; s_fz = 0; }
  __syncthreads();

  T* const ub = reinterpret_cast<T*>(ws.ubuf);
  T a[W];
  int kdone = 0;                          // columns finished (for the give-up path)
  for (int c0 = 0; c0 < Wd; c0 += W) {
    const int wn = Wd - c0 < W ? Wd - c0 : W;
    const bool active = own && pos >= c0;   // rows not pivoted by earlier leaves
#pragma unroll
    for (int c = 0; c < W; ++c) a[c] = (active && c < wn) ? P[i * ld + c0 + c] : T(0);
    T ak = a[0];

    for (int kk = 0; kk < wn; ++kk) {
      const int k = c0 + kk;
      // ---- workgroup candidate: (largest |a_ik|, lowest position) over rows at positions >= k
      T v = T(-1);
      if (own && pos >= k) {
        const T f = fabs(ak);
        if (f == f) v = f;
      }
      // the wave's (largest |a|, lowest position)
      const WaveKey wk = wave_best(WaveKey{(double)v, (own && pos >= k) ? pos : INT32_MAX, 0});
      const T wmax = (T)wk.v;
      const int wpos = wk.v >= 0.0 ? wk.idx : INT32_MAX;
      if (wpos != INT32_MAX && pos == wpos && own) {
#pragma unroll
        for (int c = 0; c < W; ++c) stage[wv][c] = a[c];
        s_phys[wv] = i;
      }
      if (lane == 0) { s_v[wv] = wmax; s_pos[wv] = wpos; }
      if (own && pos == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) stage_k[c] = a[c];
        s_knan = (ak != ak);
      }
      __syncthreads();
      T bv = s_v[0];
      int bp = s_pos[0], bw = 0;
#pragma unroll
      for (int w = 1; w < NW; ++w)
        if (better(s_v[w], s_pos[w], bv, bp)) { bv = s_v[w]; bp = s_pos[w]; bw = w; }
      const bool have = bp != INT32_MAX;

      const int slot = k & 1;
      const uint32_t tag = tag_base + (uint32_t)k + 1u;
      uint64_t* my_rec = ws.rec + ((size_t)slot * GMAX + g) * FREC;
      uint64_t* rkrec = ws.rowk + (size_t)slot * FREC;
      // publish this workgroup's candidate row and, from its owner, the row at position k
      if (have && wv == bw) {
        st_val<T>(my_rec + 4 + 2 * lane, tag, stage[bw][lane]);
        if (lane == 0) {
          st64(my_rec, tagged(tag, (uint32_t)bp));
          st64(my_rec + 1, tagged(tag, (uint32_t)s_phys[bw]));
        }
      }
      // (the wave holding the row at position k: one lane of one wave in the grid)
      const bool kowner = own && pos == k;
      if (__ballot(kowner)) {
        st_val<T>(rkrec + 4 + 2 * lane, tag, stage_k[lane]);
        if (lane == 0) st64(rkrec + 1, tagged(tag, s_knan ? 1u : 0u));
        if (kowner) st64(rkrec + 2, tagged(tag, (uint32_t)i));
      }
      if (t == 0) {
        uint64_t* q = ws.val + ((size_t)slot * GMAX + g) * 3;
        st_val<T>(q, tag, have ? bv : T(-1));
        st64(q + 2, tagged(tag, (uint32_t)(have ? bp : INT32_MAX)));
      }
      if (wv == 0 && spin_limit < 0) {
        if (lane == 0) { s_abort = 1; atomicExch(ws.err, 1); }
      } else if (wv == 0) {
        long spins = 0;
        T gv = T(-1);
        int gp = INT32_MAX, gg = 0;
        while (true) {
          bool ok = true;
          gv = T(-1); gp = INT32_MAX; gg = 0;
          for (int w = lane; w < G; w += 64) {
            const uint64_t* q = ws.val + ((size_t)slot * GMAX + w) * 3;
            const uint64_t hi = ld64(q), lo = ld64(q + 1), pw = ld64(q + 2);
            if (tag_of(hi) != tag || tag_of(lo) != tag || tag_of(pw) != tag) ok = false;
            else {
              const T v2 = val_of<T>(hi, lo);
              const int p2 = (int)payload(pw);
              if (better(v2, p2, gv, gp)) { gv = v2; gp = p2; gg = w; }
            }
          }
          if (__all(ok)) break;
          if (++spins > spin_limit) { if (lane == 0) { s_abort = 1; atomicExch(ws.err, 1); } break; }
          __builtin_amdgcn_s_sleep(1);
        }
        {
          const WaveKey bk = wave_best(WaveKey{(double)gv, gp, gg});
          gv = (T)bk.v;
          gp = bk.idx;
          gg = bk.pay;
        }
        if (!s_abort) {
          const uint64_t* r = ws.rec + ((size_t)slot * GMAX + gg) * FREC;
          const bool want_r = gv >= T(0);
          while (true) {
            const uint64_t rh = ld64(r + 4 + 2 * lane), rl = ld64(r + 5 + 2 * lane);
            const uint64_t kh = ld64(rkrec + 4 + 2 * lane), kl = ld64(rkrec + 5 + 2 * lane);
            const uint64_t pw = ld64(r), phw = ld64(r + 1), nw = ld64(rkrec + 1), kw = ld64(rkrec + 2);
            const bool ok = (!want_r || (tag_of(rh) == tag && tag_of(rl) == tag && tag_of(pw) == tag &&
                                         tag_of(phw) == tag)) &&
                            tag_of(kh) == tag && tag_of(kl) == tag && tag_of(nw) == tag && tag_of(kw) == tag;
            if (__all(ok)) {
              // position k keeps its row for a NaN there, no finite candidate, or
              // an exactly zero maximum (then every candidate is zero and the
              // row at position k is the first of them)
              const T row_w = want_r ? val_of<T>(rh, rl) : T(0);
              const bool keep_k = payload(nw) != 0 || !want_r || !(__shfl(row_w, kk) != T(0));
              const T piv_w = row_w;
              rk[lane] = val_of<T>(kh, kl);
              u[lane] = keep_k ? val_of<T>(kh, kl) : piv_w;
              if (lane == 0) {
                s_ppos = keep_k ? k : (int)payload(pw);
                s_pphys = keep_k ? (int64_t)payload(kw) : (int64_t)payload(phw);
              }
              break;
            }
            if (++spins > spin_limit) { if (lane == 0) { s_abort = 1; atomicExch(ws.err, 1); } break; }
            __builtin_amdgcn_s_sleep(1);
          }
        }
      }
      __syncthreads();
      if (s_abort) goto give_up;

      {
        const int ppos = s_ppos;
        const int64_t pphys = s_pphys;
        const T piv = u[kk];
        const bool nz = piv != T(0);
        if (t == 0) {
          s_ipiv[k] = (int32_t)(row_base + ppos + 1);
          s_prow[k] = pphys;
          if (!nz && s_fz == 0) s_fz = k + 1;
        }
        const int old = pos;
        const bool was_cand = own && old >= k;
        const bool is_piv = own && i == pphys;
        if (is_piv) pos = k;
        else if (own && old == k) pos = ppos;
        if (was_cand && !is_piv) {
          T l = ak;
          if (nz) {
            if (fabs(piv) >= sfmin) {
              const T r = T(1) / piv;
              l = l * r;
            } else {
              l = l / piv;
            }
          }
          update_row<T>(a, l, u, kk);
        }
        ak = pick<T>(a, kk + 1);
        kdone = k + 1;
      }
    }

    // ---- the leaf's columns back to memory; the rest of the block column
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (active && c < wn) P[i * ld + c0 + c] = a[c];
    const int r0 = c0 + wn;
    if (r0 >= Wd) break;
    grid_barrier(ws.bar, G, spin_limit, ws.err, &s_abort);
    if (s_abort) goto give_up;
    {
      // L11: the pivot rows' leaf columns (unit lower part), in pivot order
      {
        constexpr int PER_L = W * W / PT;
        T v[PER_L];
#pragma unroll
        for (int r = 0; r < PER_L; ++r) {
          const int e = t + r * PT, q = e / W, j = e % W;
          v[r] = j < q ? P[s_prow[c0 + q] * ld + c0 + j] : T(0);
        }
#pragma unroll
        for (int r = 0; r < PER_L; ++r) {
          const int e = t + r * PT;
          L11[e / W][e % W] = v[r];
        }
      }
      const bool upd = own && pos >= r0;   // rows still unpivoted
      // this row's next SUB columns of A22, requested one step ahead of their use
      constexpr int SUB = 8;
      T nxt[SUB];
      auto load_next = [&](int col, int lim) {
#pragma unroll
        for (int c = 0; c < SUB; ++c) nxt[c] = (upd && col + c < lim) ? P[i * ld + col + c] : T(0);
      };
      load_next(r0, r0 + CH < Wd ? r0 + CH : Wd);
      for (int cc = r0; cc < Wd; cc += CH) {
        const int cw = Wd - cc < CH ? Wd - cc : CH;
        // A12 chunk of the pivot rows
        {
          constexpr int PER_U = W * CH / PT;
          T v[PER_U];
#pragma unroll
          for (int r = 0; r < PER_U; ++r) {
            const int e = t + r * PT, q = e / CH, c = e % CH;
            v[r] = c < cw ? P[s_prow[c0 + q] * ld + cc + c] : T(0);
          }
#pragma unroll
          for (int r = 0; r < PER_U; ++r) {
            const int e = t + r * PT;
            U12[e / CH][e % CH] = v[r];
          }
        }
        __syncthreads();
        // forward substitution, TPC threads per column (one wave holds a
        // column's threads): u_q = a_q - sum_{j < q} L11[q][j] u_j
        {
          const int c = t / TPC, r = t % TPC;
          for (int q = 1; q < wn; ++q) {
            T s = T(0);
            for (int j = r; j < q; j += TPC) s += L11[q][j] * U12[j][c];
#pragma unroll
            for (int off = 1; off < TPC; off <<= 1) s += __shfl_xor(s, off);
            if (r == 0) U12[q][c] = U12[q][c] - s;
          }
        }
        __syncthreads();
        // the pivot rows' owners keep their U12 values aside (stored at the end)
        for (int e = t; e < wn * cw; e += PT) {
          const int q = e / cw, c = e % cw;
          if (s_prow[c0 + q] / PT == g) ub[(size_t)(c0 + q) * FW + cc + c] = U12[q][c];
        }
        // every unpivoted row: a_i -= l_i U12 (chunk), 16 columns at a time
        if (upd) {
#pragma unroll
          for (int c4 = 0; c4 < CH; c4 += SUB) {
            if (c4 < cw) {
              T acc[SUB];
#pragma unroll
              for (int c = 0; c < SUB; ++c) acc[c] = nxt[c];
              // the next SUB columns: the rest of this chunk, else the next chunk's first
              if (c4 + SUB < cw) load_next(cc + c4 + SUB, cc + cw);
              else load_next(cc + CH, cc + 2 * CH < Wd ? cc + 2 * CH : Wd);
#pragma unroll
              for (int q = 0; q < W; ++q) {
                if (q < wn) {
                  const T lq = a[q];
#pragma unroll
                  for (int c = 0; c < SUB; ++c) acc[c] = fma(-lq, U12[q][c4 + c], acc[c]);
                }
              }
#pragma unroll
              for (int c = 0; c < SUB; ++c)
                if (c4 + c < cw) P[i * ld + cc + c4 + c] = acc[c];
            }
          }
        }
        __syncthreads();
      }
    }
  }

  __syncthreads();   // s_ipiv, s_fz of the last column step
  // ---- U12 into the pivot rows: each thread stores the elements it set aside
  // (same loops, same element-to-thread mapping, so it reads its own stores)
  for (int c0 = 0; c0 + W < Wd; c0 += W)
    for (int cc = c0 + W; cc < Wd; cc += CH) {
      const int cw = Wd - cc < CH ? Wd - cc : CH;
      for (int e = t; e < W * cw; e += PT) {
        const int q = e / cw, c = e % cw;
        const int64_t row = s_prow[c0 + q];
        if (row / PT == g) P[row * ld + cc + c] = ub[(size_t)(c0 + q) * FW + cc + c];
      }
    }
  // ---- row moves into LAPACK order, interchanges, info
  if (own && pos != (int)i) {
    const int q = atomicAdd(pairs, 1);
    pairs[1 + 2 * q] = (int32_t)(row_base + pos);
    pairs[2 + 2 * q] = (int32_t)(row_base + i);
  }
  if (g == 0) {
    for (int k = t; k < Wd; k += PT) ipiv[k] = s_ipiv[k];
    if (t == 0 && s_fz != 0 && *info == 0) *info = (int)(info_base + s_fz);
  }
  return;

give_up:
  // the columns not reached keep their rows: identity interchanges, no moves
  if (g == 0)
    for (int k = t; k < Wd; k += PT) ipiv[k] = k < kdone ? s_ipiv[k] : (int32_t)(row_base + k + 1);
}

}  // namespace panel
}  // namespace zmlhip
