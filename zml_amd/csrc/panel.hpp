// LU panel machinery shared by the FP64 and the mixed-precision drivers:
//   * the leaf kernel: partial-pivoting LU of a tall M x W row-major panel,
//     one row per thread held in registers, all workgroups resident;
//   * row-interchange application through the composed permutation.
//
// Cross-workgroup exchange (leaf kernel, more than one workgroup). Per
// column k every workgroup publishes, into slot k mod 2,
//   val[g]  = |a| of its candidate (largest |a_ik| among its rows)
//   rec[g]  = {row index, candidate row}
//   rowk    = {NaN flag, row k}              (owner of row k only)
// Every 64-bit word carries a 32-bit tag (launch base + k + 1) above 32
// payload bits (a double is split over two words), stored and loaded as
// device-scope relaxed atomics, which the memory model keeps coherent
// across XCDs without cache-wide fences. A reader knows a word is current
// from its tag alone, so writers never wait for their stores: a reader polls
// the G val words, picks the winner (largest |a|, lowest workgroup on ties =
// lowest row, since workgroups own increasing row ranges: iamax.zig:32's
// strict '>'), then fetches the winner's record and row k, re-reading any
// word whose tag is not yet current. A slot is rewritten two columns later,
// when every workgroup has provably finished reading it (it has published
// the column in between).
#pragma once

#include "zmlhip_internal.hpp"

namespace zmlhip {
namespace panel {

constexpr int W = 64;            // leaf width (columns held in registers)
// threads (= rows) per workgroup: FP64 rows take 128 VGPRs, so one wave per
// SIMD keeps the whole register file (incl. AGPRs) available to a thread.
template <typename T> constexpr int pt_of() { return sizeof(T) == 8 ? 256 : 512; }
constexpr int GMAX = 256;        // workgroups per leaf
constexpr int REC = 2 + 2 * W;   // words per record: index/flag + 2 per value (+ pad)

struct WS {
  uint64_t* val;    // [2][GMAX][2]
  uint64_t* rec;    // [2][GMAX][REC]
  uint64_t* rowk;   // [2][REC]
  int* err;
};

inline size_t ws_bytes() {
  return sizeof(uint64_t) * (2 * GMAX * 2 + 2 * GMAX * REC + 2 * REC) + 256;
}
inline WS carve(void* base) {
  WS w;
  uint64_t* p = static_cast<uint64_t*>(base);
  w.val = p; p += 2 * GMAX * 2;
  w.rec = p; p += 2 * GMAX * REC;
  w.rowk = p; p += 2 * REC;
  w.err = reinterpret_cast<int*>(p);
  return w;
}

__device__ __forceinline__ void st64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld64(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t tagged(uint32_t tag, uint32_t payload) {
  return ((uint64_t)tag << 32) | payload;
}
__device__ __forceinline__ uint32_t tag_of(uint64_t w) { return (uint32_t)(w >> 32); }
__device__ __forceinline__ uint32_t payload(uint64_t w) { return (uint32_t)w; }
// a value travels as a double split over two tagged words
template <typename T>
__device__ __forceinline__ void st_val(uint64_t* p, uint32_t tag, T v) {
  const uint64_t b = (uint64_t)__double_as_longlong((double)v);
  st64(p, tagged(tag, (uint32_t)(b >> 32)));
  st64(p + 1, tagged(tag, (uint32_t)b));
}
template <typename T>
__device__ __forceinline__ T val_of(uint64_t hi, uint64_t lo) {
  return (T)__longlong_as_double((long long)(((uint64_t)payload(hi) << 32) | payload(lo)));
}

// (value, index) maximum with the lowest index winning ties.
template <typename V, typename I>
__device__ __forceinline__ void amax(V& v, I& i, V v2, I i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

// ---- wave-wide (value, index) maximum through DPP ---------------------------
// The key: a value (|a| >= 0, or -1 for "no candidate"), an index that breaks
// ties (lowest wins: a row, a position or a workgroup) and a payload carried
// along with the winner. The reduction runs in registers: quad permutations,
// row rotations by 4 and 8, then the row broadcasts 15 and 31 (CDNA DPP), so
// lane 63 ends with the maximum over the wave, which every lane then reads.
// Keys are totally ordered, so the order in which lanes meet does not matter.
struct WaveKey {
  double v;
  int idx;
  int pay;
};
__device__ __forceinline__ bool key_better(const WaveKey& a, const WaveKey& b) {
  return a.v > b.v || (a.v == b.v && a.idx < b.idx);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ WaveKey key_dpp(const WaveKey& k) {
  // lanes whose source is outside the pattern keep their own key
  const uint64_t b = (uint64_t)__double_as_longlong(k.v);
  const int hi = (int)(b >> 32), lo = (int)b;
  WaveKey r;
  const uint32_t rhi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
  const uint32_t rlo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
  r.v = __longlong_as_double((long long)(((uint64_t)rhi << 32) | rlo));
  r.idx = __builtin_amdgcn_update_dpp(k.idx, k.idx, CTRL, ROW_MASK, 0xf, false);
  r.pay = __builtin_amdgcn_update_dpp(k.pay, k.pay, CTRL, ROW_MASK, 0xf, false);
  return r;
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ void key_step(WaveKey& k) {
  const WaveKey o = key_dpp<CTRL, ROW_MASK>(k);
  if (key_better(o, k)) k = o;
}
__device__ __forceinline__ WaveKey wave_best(WaveKey k) {
  key_step<0xb1>(k);         // quad_perm [1,0,3,2]
  key_step<0x4e>(k);         // quad_perm [2,3,0,1]
  key_step<0x124>(k);        // row_ror:4
  key_step<0x128>(k);        // row_ror:8  (every lane: its row's best)
  key_step<0x142, 0xa>(k);   // row_bcast:15 into rows 1 and 3
  key_step<0x143, 0xc>(k);   // row_bcast:31 into rows 2 and 3 (lane 63: the wave's best)
  const uint64_t b = (uint64_t)__double_as_longlong(k.v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(b >> 32), 63);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)b, 63);
  WaveKey r;
  r.v = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
  r.idx = __builtin_amdgcn_readlane(k.idx, 63);
  r.pay = __builtin_amdgcn_readlane(k.pay, 63);
  return r;
}

// Rank-1 update of one row for column k: a[k] <- l, a[c] <- a[c] - l u[c]
// for c > k, separately rounded. The register row is handled in groups of 8
// columns so that every index is a compile-time constant; k is uniform, so
// the group tests are scalar branches and only the group holding k needs
// per-column predication.
template <int G8, typename T>
__device__ __forceinline__ void update_group(T (&a)[W], T l, const T* __restrict__ u, int k) {
#pragma clang fp contract(off)
  constexpr int c0 = 8 * G8;
  if (c0 > k) {
#pragma unroll
    for (int c = c0; c < c0 + 8; ++c) a[c] = a[c] - l * u[c];
  } else if (c0 + 7 >= k) {
#pragma unroll
    for (int c = c0; c < c0 + 8; ++c) {
      if (c == k) a[c] = l;
      else if (c > k) a[c] = a[c] - l * u[c];
    }
  }
}

template <int G8, typename T>
__device__ __forceinline__ void pick_group(const T (&a)[W], int c, T& out) {
  constexpr int c0 = 8 * G8;
  if (c >= c0 && c < c0 + 8) {
#pragma unroll
    for (int x = c0; x < c0 + 8; ++x) if (x == c) out = a[x];
  }
}

// a[c] for a uniform, run-time c (0 if c >= W).
template <typename T>
__device__ __forceinline__ T pick(const T (&a)[W], int c) {
  T out = T(0);
  pick_group<0>(a, c, out); pick_group<1>(a, c, out); pick_group<2>(a, c, out); pick_group<3>(a, c, out);
  pick_group<4>(a, c, out); pick_group<5>(a, c, out); pick_group<6>(a, c, out); pick_group<7>(a, c, out);
  return out;
}

// The empty asm statements keep the compiler from hoisting every group's
// pivot-row reads (LDS) ahead of the first group: only 8 values of u are live
// at a time, which keeps the FP64 leaf free of register spills.
#define ZH_GROUP_FENCE() asm volatile("" ::: "memory")
template <typename T>
__device__ __forceinline__ void update_row(T (&a)[W], T l, const T* __restrict__ u, int k) {
  update_group<0>(a, l, u, k); ZH_GROUP_FENCE(); update_group<1>(a, l, u, k); ZH_GROUP_FENCE();
  update_group<2>(a, l, u, k); ZH_GROUP_FENCE(); update_group<3>(a, l, u, k); ZH_GROUP_FENCE();
  update_group<4>(a, l, u, k); ZH_GROUP_FENCE(); update_group<5>(a, l, u, k); ZH_GROUP_FENCE();
  update_group<6>(a, l, u, k); ZH_GROUP_FENCE(); update_group<7>(a, l, u, k);
}
#undef ZH_GROUP_FENCE

// Leaf LU: rows [0, M) x columns [0, Wn) of the row-major panel P (Wn <= 64).
// ipiv[k] = piv_base + pivot row + 1 (1-based, global); *info gets
// info_base + first zero pivot (1-based) when it is still 0.
// MULTI: more than one workgroup (exchange through ws), else LDS only.
// A wait that exceeds spin_limit polls (or any wait, when spin_limit < 0)
// sets *ws.err and ends the kernel in every workgroup; the factors are then
// invalid and the host reports the give-up.
template <typename T, bool MULTI>
__global__ void __launch_bounds__(pt_of<T>())
leaf_kernel(int64_t M, int Wn, T* __restrict__ P, int64_t ld, int32_t* __restrict__ ipiv,
            int64_t piv_base, int* __restrict__ info, int64_t info_base, WS ws, uint32_t tag_base,
            long spin_limit) {
#pragma clang fp contract(off)
  constexpr int PT = pt_of<T>();
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  constexpr int NW = PT / 64;
  const int64_t i = (int64_t)g * PT + t;
  const bool own = i < M;

  __shared__ T u[W], rk[W];
  __shared__ T s_v[NW];
  __shared__ int64_t s_i[NW];
  __shared__ int64_t s_p;
  __shared__ int s_nan, s_abort;
  __shared__ int32_t s_piv[W];
  __shared__ T stage[MULTI ? NW : 1][W], stage_k[W];
  __shared__ int s_knan;

  T a[W];
#pragma unroll
  for (int c = 0; c < W; ++c) a[c] = (own && c < Wn) ? P[i * ld + c] : T(0);
  if (t == 0) { s_abort = 0; s_nan = 0; }

  int first_zero = 0;
  const T sfmin = sizeof(T) == 8 ? T(0x1p-1022) : T(0x1p-126);   // lamch('S')
  T ak = a[0];

  for (int k = 0; k < Wn; ++k) {
    // ---- workgroup candidate: largest |a_ik| over own rows i >= k (NaN never
    // wins). The wave maximum is found on the value alone, the lowest lane
    // holding it by a ballot (lanes hold increasing rows).
    T v = T(-1);
    if (own && i >= k) {
      const T f = fabs(ak);
      if (f == f) v = f;
    }
    const WaveKey wk = wave_best(WaveKey{(double)v, lane, 0});
    const T wmax = (T)wk.v;
    const int wlane = wk.v >= 0.0 ? wk.idx : -1;   // lowest lane (= row) holding the maximum
    if (lane == 0) {
      s_v[wv] = wmax;
      s_i[wv] = wlane >= 0 ? (int64_t)g * PT + wv * 64 + wlane : INT64_MAX;
    }
    if (MULTI) {
      // stage the wave's candidate row and row k for publishing (the barrier
      // below orders these writes before the publishing wave reads them)
      if (lane == wlane) {
#pragma unroll
        for (int c = 0; c < W; ++c) stage[wv][c] = a[c];
      }
      if (own && i == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) stage_k[c] = a[c];
        s_knan = (ak != ak);
      }
    }
    __syncthreads();
    T bv = s_v[0];
    int64_t bi = s_i[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) amax(bv, bi, s_v[w], s_i[w]);

    if (!MULTI) {
      if (own && i == bi) {
#pragma unroll
        for (int c = 0; c < W; ++c) u[c] = a[c];
        s_p = i;
      }
      if (own && i == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) rk[c] = a[c];
        s_nan = (ak != ak);
      }
      __syncthreads();
      if (t == 0 && (s_nan || bi == INT64_MAX)) s_p = k;
      if (s_nan && t < W) u[t] = rk[t];
      __syncthreads();
    } else {
      const int slot = k & 1;
      const uint32_t tag = tag_base + (uint32_t)k + 1u;
      uint64_t* my_rec = ws.rec + ((size_t)slot * GMAX + g) * REC;
      uint64_t* rkrec = ws.rowk + (size_t)slot * REC;
      // publish the candidate row and row k from the staging rows, one value
      // per lane (lane c stores column c)
      const int64_t base_row = (int64_t)g * PT;
      if (bi != INT64_MAX && wv == (int)((bi - base_row) >> 6)) {
        st_val<T>(my_rec + 2 + 2 * lane, tag, stage[wv][lane]);
        if (lane == 0) st64(my_rec, tagged(tag, (uint32_t)bi));
      }
      if (k >= base_row && k < base_row + PT && wv == (int)((k - base_row) >> 6)) {
        st_val<T>(rkrec + 2 + 2 * lane, tag, stage_k[lane]);
        if (lane == 0) st64(rkrec, tagged(tag, s_knan ? 1u : 0u));
      }
      if (t == 0) st_val<T>(ws.val + ((size_t)slot * GMAX + g) * 2, tag, bi == INT64_MAX ? T(-1) : bv);
      if (wv == 0 && spin_limit < 0) {   // give-up forced by the host (error-path tests)
        if (lane == 0) { s_abort = 1; atomicExch(ws.err, 1); }
      } else if (wv == 0) {
        long spins = 0;
        // (1) every workgroup's |a|
        T gv = T(-1);
        int gg = INT32_MAX;
        while (true) {
          bool ok = true;
          gv = T(-1);
          gg = INT32_MAX;
          for (int w = lane; w < G; w += 64) {
            const uint64_t* q = ws.val + ((size_t)slot * GMAX + w) * 2;
            const uint64_t hi = ld64(q), lo = ld64(q + 1);
            if (tag_of(hi) != tag || tag_of(lo) != tag) ok = false;
            else amax(gv, gg, val_of<T>(hi, lo), w);
          }
          if (__all(ok)) break;
          if (++spins > spin_limit) { if (lane == 0) { s_abort = 1; atomicExch(ws.err, 1); } break; }
          __builtin_amdgcn_s_sleep(1);
        }
        {
          const WaveKey bk = wave_best(WaveKey{(double)gv, gg, 0});
          gv = (T)bk.v;
          gg = bk.idx;
        }
        // (2) the winner's record and row k
        if (!s_abort) {
          const uint64_t* r = ws.rec + ((size_t)slot * GMAX + (gg == INT32_MAX ? 0 : gg)) * REC;
          const bool want_r = gv >= T(0);
          while (true) {
            const uint64_t rh = ld64(r + 2 + 2 * lane), rl = ld64(r + 3 + 2 * lane);
            const uint64_t kh = ld64(rkrec + 2 + 2 * lane), kl = ld64(rkrec + 3 + 2 * lane);
            const uint64_t pw = ld64(r), nw = ld64(rkrec);
            const bool ok = (!want_r || (tag_of(rh) == tag && tag_of(rl) == tag && tag_of(pw) == tag)) &&
                            tag_of(kh) == tag && tag_of(kl) == tag && tag_of(nw) == tag;
            if (__all(ok)) {
              const bool nanf = payload(nw) != 0;
              rk[lane] = val_of<T>(kh, kl);
              u[lane] = (nanf || !want_r) ? val_of<T>(kh, kl) : val_of<T>(rh, rl);
              if (lane == 0) s_p = (nanf || !want_r) ? (int64_t)k : (int64_t)payload(pw);
              break;
            }
            if (++spins > spin_limit) { if (lane == 0) { s_abort = 1; atomicExch(ws.err, 1); } break; }
            __builtin_amdgcn_s_sleep(1);
          }
        }
      }
      __syncthreads();
      if (s_abort) {
        // gave up: the columns not reached keep their rows (identity pivots), so
        // the interchanges the caller applies next stay inside the matrix
        if (g == 0 && t < Wn) ipiv[t] = t < k ? s_piv[t] : (int32_t)(piv_base + t + 1);
        return;
      }
    }

    const int64_t p = s_p;
    const T piv = u[k];
    const bool nz = piv != T(0);   // getrf2.zig:55: no interchange on an exact zero
    if (!nz && first_zero == 0) first_zero = k + 1;
    // pivots are written out at the end: a global store here would make the
    // next barrier wait for its completion
    if (t == 0) s_piv[k] = (int32_t)(piv_base + (nz ? p : k) + 1);
    T l = ak;
    if (nz && p != k) {
      if (own && i == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = u[c];
      } else if (own && i == p) {
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = rk[c];
        l = rk[k];
      }
    }
    if (own && i > k) {
      if (nz) {
        if (fabs(piv) >= sfmin) {
          const T r = T(1) / piv;
          l = l * r;
        } else {
          l = l / piv;
        }
      }
      update_row<T>(a, l, u, k);
    }
    ak = pick<T>(a, k + 1);   // candidate value of the next column
  }

  // nothing is in flight any more: the row's stores go out back to back
  __builtin_amdgcn_s_waitcnt(0xF70);
  if (own) {
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < Wn) P[i * ld + c] = a[c];
  }
  if (g == 0 && t < Wn) ipiv[t] = s_piv[t];
  if (g == 0 && t == 0 && first_zero != 0 && *info == 0) *info = (int)(info_base + first_zero);
}

// ---- look-ahead leaf (more than one workgroup) -------------------------------
// The leaf LU above with the exchange of column k + 1 started before the
// update of column k is finished: once column k's pivot row is known, every
// row applies column k's update to column k + 1 only, the candidates of
// column k + 1 are chosen and published, and only then the update of columns
// k + 2 .. W - 1 runs, while the other workgroups' candidates travel. A
// published row therefore still lacks column k's update in columns > k + 1;
// its multiplier l (stored at column k) comes with it, and the reader applies
// a_c - l u_c itself, separately rounded as the owner does, so every value
// and pivot is the leaf kernel's bit for bit. u and row k are double-buffered
// by column parity: column k's pivot row is read by the update of its tail
// after column k + 1's exchange has begun.

// Column k's update of a row, split in two: the head (a[k] <- l, a[k+1] <-
// a[k+1] - l u[k+1]) and the tail (a[c] <- a[c] - l u[c], c >= k + 2), in the
// 8-column groups update_row uses, separately rounded.
template <int G8, typename T>
__device__ __forceinline__ void head_group(T (&a)[W], T l, const T* __restrict__ u, int k) {
#pragma clang fp contract(off)
  constexpr int c0 = 8 * G8;
  if (c0 + 7 >= k && c0 <= k + 1) {
#pragma unroll
    for (int c = c0; c < c0 + 8; ++c) {
      if (c == k) a[c] = l;
      else if (c == k + 1) a[c] = a[c] - l * u[c];
    }
  }
}
template <typename T>
__device__ __forceinline__ void update_head(T (&a)[W], T l, const T* __restrict__ u, int k) {
  head_group<0>(a, l, u, k); head_group<1>(a, l, u, k); head_group<2>(a, l, u, k); head_group<3>(a, l, u, k);
  head_group<4>(a, l, u, k); head_group<5>(a, l, u, k); head_group<6>(a, l, u, k); head_group<7>(a, l, u, k);
}
// a[c] <- a[c] - l u[c] for c >= cf (cf uniform), 8 columns at a time
template <int G8, typename T>
__device__ __forceinline__ void tail_group(T (&a)[W], T l, const T* __restrict__ u, int cf) {
#pragma clang fp contract(off)
  constexpr int c0 = 8 * G8;
  if (c0 >= cf) {
#pragma unroll
    for (int c = c0; c < c0 + 8; ++c) a[c] = a[c] - l * u[c];
  } else if (c0 + 7 >= cf) {
#pragma unroll
    for (int c = c0; c < c0 + 8; ++c)
      if (c >= cf) a[c] = a[c] - l * u[c];
  }
}
#define ZH_GROUP_FENCE() asm volatile("" ::: "memory")
template <typename T>
__device__ __forceinline__ void update_tail(T (&a)[W], T l, const T* __restrict__ u, int cf) {
  tail_group<0>(a, l, u, cf); ZH_GROUP_FENCE(); tail_group<1>(a, l, u, cf); ZH_GROUP_FENCE();
  tail_group<2>(a, l, u, cf); ZH_GROUP_FENCE(); tail_group<3>(a, l, u, cf); ZH_GROUP_FENCE();
  tail_group<4>(a, l, u, cf); ZH_GROUP_FENCE(); tail_group<5>(a, l, u, cf); ZH_GROUP_FENCE();
  tail_group<6>(a, l, u, cf); ZH_GROUP_FENCE(); tail_group<7>(a, l, u, cf);
}
#undef ZH_GROUP_FENCE

// Candidates of column kc (rows >= kc): the wave and workgroup maxima, the
// candidate row and row kc staged in LDS, then published (records and val).
template <typename T, int NW>
__device__ __forceinline__ void la_candidate_publish(int kc, const T (&a)[W], T ak, bool own, int64_t i,
                                                     int64_t base_row, int g, int t, int lane, int wv, const WS& ws,
                                                     uint32_t tag_base, T (*stage)[W], T* stage_k, T* s_v,
                                                     int64_t* s_i, int* s_knan) {
  constexpr int PT = NW * 64;
  T v = T(-1);
  if (own && i >= kc) {
    const T f = fabs(ak);
    if (f == f) v = f;
  }
  const WaveKey wk = wave_best(WaveKey{(double)v, lane, 0});
  const int wlane = wk.v >= 0.0 ? wk.idx : -1;
  if (lane == 0 && wv < NW) {
    s_v[wv] = (T)wk.v;
    s_i[wv] = wlane >= 0 ? base_row + wv * 64 + wlane : INT64_MAX;
  }
  if (lane == wlane && wv < NW) {
#pragma unroll
    for (int c = 0; c < W; ++c) stage[wv][c] = a[c];
  }
  if (own && i == kc) {
#pragma unroll
    for (int c = 0; c < W; ++c) stage_k[c] = a[c];
    *s_knan = (ak != ak);
  }
  __syncthreads();
  T bv = s_v[0];
  int64_t bi = s_i[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) amax(bv, bi, s_v[w], s_i[w]);
  const int slot = kc & 1;
  const uint32_t tag = tag_base + (uint32_t)kc + 1u;
  uint64_t* my_rec = ws.rec + ((size_t)slot * GMAX + g) * REC;
  uint64_t* rkrec = ws.rowk + (size_t)slot * REC;
  if (bi != INT64_MAX && wv == (int)((bi - base_row) >> 6)) {
    st_val<T>(my_rec + 2 + 2 * lane, tag, stage[wv][lane]);
    if (lane == 0) st64(my_rec, tagged(tag, (uint32_t)bi));
  }
  if (kc >= base_row && kc < base_row + PT && wv == (int)((kc - base_row) >> 6)) {
    st_val<T>(rkrec + 2 + 2 * lane, tag, stage_k[lane]);
    if (lane == 0) st64(rkrec, tagged(tag, *s_knan ? 1u : 0u));
  }
  if (t == 0) st_val<T>(ws.val + ((size_t)slot * GMAX + g) * 2, tag, bi == INT64_MAX ? T(-1) : bv);
}

// Wave 0: the winner of column kc and row kc, completed with column kc - 1's
// update, into ub / rkb[kc & 1] and *s_p; *s_abort on a give-up.
template <typename T>
__device__ __forceinline__ void la_poll(int kc, int G, int lane, const WS& ws, uint32_t tag_base, long spin_limit,
                                        T (*ub)[W], T (*rkb)[W], int64_t* s_p, int* s_abort) {
#pragma clang fp contract(off)
  const int slot = kc & 1;
  const uint32_t tag = tag_base + (uint32_t)kc + 1u;
  const uint64_t* rkrec = ws.rowk + (size_t)slot * REC;
  if (spin_limit < 0) {   // give-up forced by the host (error-path tests)
    if (lane == 0) { *s_abort = 1; atomicExch(ws.err, 1); }
    return;
  }
  long spins = 0;
  T gv = T(-1);
  int gg = INT32_MAX;
  while (true) {
    bool ok = true;
    gv = T(-1);
    gg = INT32_MAX;
    for (int w = lane; w < G; w += 64) {
      const uint64_t* q = ws.val + ((size_t)slot * GMAX + w) * 2;
      const uint64_t hi = ld64(q), lo = ld64(q + 1);
      if (tag_of(hi) != tag || tag_of(lo) != tag) ok = false;
      else amax(gv, gg, val_of<T>(hi, lo), w);
    }
    if (__all(ok)) break;
    if (++spins > spin_limit) { if (lane == 0) { *s_abort = 1; atomicExch(ws.err, 1); } return; }
    __builtin_amdgcn_s_sleep(1);
  }
  {
    const WaveKey bk = wave_best(WaveKey{(double)gv, gg, 0});
    gv = (T)bk.v;
    gg = bk.idx;
  }
  const uint64_t* r = ws.rec + ((size_t)slot * GMAX + (gg == INT32_MAX ? 0 : gg)) * REC;
  const bool want_r = gv >= T(0);
  while (true) {
    const uint64_t rh = ld64(r + 2 + 2 * lane), rl = ld64(r + 3 + 2 * lane);
    const uint64_t kh = ld64(rkrec + 2 + 2 * lane), kl = ld64(rkrec + 3 + 2 * lane);
    const uint64_t pw = ld64(r), nw = ld64(rkrec);
    const bool ok = (!want_r || (tag_of(rh) == tag && tag_of(rl) == tag && tag_of(pw) == tag)) &&
                    tag_of(kh) == tag && tag_of(kl) == tag && tag_of(nw) == tag;
    if (__all(ok)) {
      T rv = val_of<T>(rh, rl), kv = val_of<T>(kh, kl);
      if (kc > 0) {
        // both rows still lack column kc - 1's update beyond column kc; their
        // multipliers sit at column kc - 1
        const int lc = kc - 1;
        const T lr = __shfl(rv, lc), lk = __shfl(kv, lc);
        if (lane > kc) {
          const T up = ub[lc & 1][lane];
          rv = rv - lr * up;
          kv = kv - lk * up;
        }
      }
      const bool nanf = payload(nw) != 0;
      rkb[slot][lane] = kv;
      ub[slot][lane] = (nanf || !want_r) ? kv : rv;
      if (lane == 0) *s_p = (nanf || !want_r) ? (int64_t)kc : (int64_t)payload(pw);
      return;
    }
    if (++spins > spin_limit) { if (lane == 0) { *s_abort = 1; atomicExch(ws.err, 1); } return; }
    __builtin_amdgcn_s_sleep(1);
  }
}

// threads of a look-ahead leaf workgroup: the row waves and one exchange wave
template <typename T> constexpr int la_threads() { return pt_of<T>() + 64; }

template <typename T>
__global__ void __launch_bounds__(la_threads<T>())
leaf_la_kernel(int64_t M, int Wn, T* __restrict__ P, int64_t ld, int32_t* __restrict__ ipiv,
               int64_t piv_base, int* __restrict__ info, int64_t info_base, WS ws, uint32_t tag_base,
               long spin_limit) {
#pragma clang fp contract(off)
  constexpr int PT = pt_of<T>();
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  constexpr int NW = PT / 64;
  // waves 0 .. NW-1 hold the rows; wave NW only exchanges (it polls while
  // the row waves update the tail of the previous column)
  const int64_t i = (int64_t)g * PT + t;
  const bool own = t < PT && i < M;
  const int64_t base_row = (int64_t)g * PT;

  __shared__ T ub[2][W], rkb[2][W];
  __shared__ T s_v[NW];
  __shared__ int64_t s_i[NW];
  __shared__ int64_t s_p;
  __shared__ int s_abort, s_knan;
  __shared__ int32_t s_piv[W];
  __shared__ T stage[NW][W], stage_k[W];

  T a[W];
#pragma unroll
  for (int c = 0; c < W; ++c) a[c] = (own && c < Wn) ? P[i * ld + c] : T(0);
  if (t == 0) { s_abort = 0; s_knan = 0; }

  int first_zero = 0;
  const T sfmin = sizeof(T) == 8 ? T(0x1p-1022) : T(0x1p-126);   // lamch('S')
  T ak = a[0];   // this row's current value in the column being searched

  la_candidate_publish<T, NW>(0, a, ak, own, i, base_row, g, t, lane, wv, ws, tag_base, stage, stage_k, s_v,
                             s_i, &s_knan);
  if (wv == NW) la_poll<T>(0, G, lane, ws, tag_base, spin_limit, ub, rkb, &s_p, &s_abort);
  for (int k = 0; k < Wn; ++k) {
    __syncthreads();   // column k's pivot row is in, column k - 1's tail updated
    if (s_abort) {
      // gave up: the columns not reached keep their rows (identity pivots), so
      // the interchanges the caller applies next stay inside the matrix
      if (g == 0 && t < Wn) ipiv[t] = t < k ? s_piv[t] : (int32_t)(piv_base + t + 1);
      return;
    }
    const T* u = ub[k & 1];
    const T* rk = rkb[k & 1];
    const int64_t p = s_p;
    const T piv = u[k];
    const bool nz = piv != T(0);   // getrf2.zig:55: no interchange on an exact zero
    if (!nz && first_zero == 0) first_zero = k + 1;
    if (t == 0) s_piv[k] = (int32_t)(piv_base + (nz ? p : k) + 1);
    T l = ak;
    if (nz && p != k) {
      if (own && i == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = u[c];
      } else if (own && i == p) {
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = rk[c];
        l = rk[k];
      }
    }
    const bool upd = own && i > k;
    if (upd) {
      if (nz) {
        if (fabs(piv) >= sfmin) {
          const T r = T(1) / piv;
          l = l * r;
        } else {
          l = l / piv;
        }
      }
      update_head<T>(a, l, u, k);
    }
    ak = pick<T>(a, k + 1);   // candidate value of the next column
    if (k + 1 < Wn) {
      la_candidate_publish<T, NW>(k + 1, a, ak, own, i, base_row, g, t, lane, wv, ws, tag_base, stage, stage_k,
                                  s_v, s_i, &s_knan);
      if (wv == NW) la_poll<T>(k + 1, G, lane, ws, tag_base, spin_limit, ub, rkb, &s_p, &s_abort);
    }
    if (upd) update_tail<T>(a, l, u, k + 2);
  }

  // nothing is in flight any more: the row's stores go out back to back
  __builtin_amdgcn_s_waitcnt(0xF70);
  if (own) {
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < Wn) P[i * ld + c] = a[c];
  }
  if (g == 0 && t < Wn) ipiv[t] = s_piv[t];
  if (g == 0 && t == 0 && first_zero != 0 && *info == 0) *info = (int)(info_base + first_zero);
}

// ---- row interchanges through the composed permutation ----------------------
constexpr int PERM_MAX = 1024;  // pairs (2 x block of 512 interchanges)

// Composes the interchanges i <-> ipiv[i]-1 for i = k1..k2 (ascending, every
// ipiv[i]-1 >= i) into a list of moves dst <- src over the rows touched.
// One wave: the rows beyond the block enter an open-addressing table
// (row -> slot) in parallel, each interchange's partner slot is looked up in
// parallel, one lane then walks the interchanges in order (a swap of two LDS
// words each), and the slots whose content changed become moves, compacted by
// a ballot. pairs[0] = count, then (dst, src) pairs in slot order.
constexpr int PERM_HT = 2 * PERM_MAX;   // power of two, load factor <= 1/4
static_assert((PERM_HT & (PERM_HT - 1)) == 0, "table size must be a power of two");
__device__ __forceinline__ uint32_t perm_hash(int p) { return ((uint32_t)p * 2654435761u) & (PERM_HT - 1); }
__global__ void __launch_bounds__(64)
compose_perm_kernel(int k1, int k2, const int32_t* __restrict__ ipiv, int32_t* __restrict__ pairs) {
  __shared__ int32_t pos[PERM_MAX], cont[PERM_MAX], partner[PERM_MAX / 2];
  __shared__ int32_t hkey[PERM_HT], hslot[PERM_HT];
  __shared__ int s_n;
  const int lane = threadIdx.x;
  const int kb = k2 - k1 + 1;
  for (int q = lane; q < kb; q += 64) { pos[q] = k1 + q; cont[q] = k1 + q; }
  for (int q = lane; q < PERM_HT; q += 64) hkey[q] = -1;
  if (lane == 0) s_n = kb;
  __syncthreads();
  // the distinct rows beyond the block enter the table
  for (int q = lane; q < kb; q += 64) {
    const int p = ipiv[k1 + q] - 1;
    if (p > k2) {
      uint32_t h = perm_hash(p);
      for (;;) {
        const int prev = atomicCAS(&hkey[h], -1, p);
        if (prev == -1 || prev == p) break;
        h = (h + 1) & (PERM_HT - 1);
      }
    }
  }
  __syncthreads();
  // each of them gets a slot after the block's
  for (int h = lane; h < PERM_HT; h += 64) {
    const int key = hkey[h];
    if (key >= 0) {
      const int sl = atomicAdd(&s_n, 1);
      hslot[h] = sl;
      pos[sl] = key;
      cont[sl] = key;
    }
  }
  __syncthreads();
  // the slot each interchange swaps with
  for (int q = lane; q < kb; q += 64) {
    const int p = ipiv[k1 + q] - 1;
    int qp;
    if (p <= k2) {
      qp = p - k1;
    } else {
      uint32_t h = perm_hash(p);
      while (hkey[h] != p) h = (h + 1) & (PERM_HT - 1);
      qp = hslot[h];
    }
    partner[q] = qp;
  }
  __syncthreads();
  if (lane == 0) {
    for (int q = 0; q < kb; ++q) {
      const int qp = partner[q];
      if (qp != q) {
        const int tmp = cont[q];
        cont[q] = cont[qp];
        cont[qp] = tmp;
      }
    }
  }
  __syncthreads();
  const int n = s_n;
  int c = 0;
  for (int q0 = 0; q0 < n; q0 += 64) {
    const int q = q0 + lane;
    const bool moved = q < n && cont[q] != pos[q];
    const uint64_t m = __ballot(moved);
    if (moved) {
      const int at = c + (int)__popcll(m & ((1ull << lane) - 1ull));
      pairs[1 + 2 * at] = pos[q];
      pairs[2 + 2 * at] = cont[q];
    }
    c += (int)__popcll(m);
  }
  if (lane == 0) pairs[0] = c;
}

// Applies the moves to columns [0, ncols) of a row-major matrix: every
// workgroup owns 16 columns, copies the move list into LDS, reads all source
// rows (no load waits on another: the row numbers come from LDS), then writes
// them.
template <typename T>
__global__ void __launch_bounds__(256)
apply_perm_kernel(int64_t ncols, T* __restrict__ a, int64_t lda, const int32_t* __restrict__ pairs) {
  constexpr int CW = 16, RG = 256 / CW, RPT = PERM_MAX / RG;   // 64 moves per thread
  __shared__ int32_t mv[2 * PERM_MAX];                          // (destination, source) per move
  const int n = min(pairs[0], PERM_MAX);
  for (int q = threadIdx.x; q < 2 * n; q += 256) mv[q] = pairs[1 + q];
  __syncthreads();
  const int col = threadIdx.x % CW, q0 = threadIdx.x / CW;
  const int64_t c = blockIdx.x * (int64_t)CW + col;
  const int64_t cl = c < ncols ? c : ncols - 1;   // loads stay in bounds; only c < ncols is written
  T v[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    if (r * RG >= n) break;
    const int q = q0 + r * RG;
    v[r] = a[(int64_t)mv[2 * (q < n ? q : 0) + 1] * lda + cl];
  }
  // a column chunk belongs to one workgroup: all of its reads (every load
  // returned: s_waitcnt vmcnt(0)) precede all of its writes
  __builtin_amdgcn_s_waitcnt(0xF70);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    if (r * RG >= n) break;
    const int q = q0 + r * RG;
    if (q < n && c < ncols) a[(int64_t)mv[2 * q] * lda + c] = v[r];
  }
}

}  // namespace panel
}  // namespace zmlhip
