// Batched mixed-precision solves of many small independent systems (n <= 256).
//
// No counterpart in zml (its getrf/potrf work on one matrix per call); per
// system the arithmetic is the one of the large mixed solves: an FP32
// factorization and FP64 iterative refinement against the FP64 A.
//
// A plan sorts the systems into size buckets (n <= 32, 64, 128, 256). Every
// bucket runs one persistent kernel whose workgroups take systems from an
// atomic work queue; a workgroup keeps its system's FP32 factor in registers
// (a TY x TX thread grid, RI x RJ values per thread, NMAX = TY RI = TX RJ) from the
// factorization through every refinement step, and reads the FP64 matrix
// from HBM only for the residuals. The thread-grid shape of every bucket is
// chosen by timing the candidate shapes on the bucket's own systems
// (autotune), and the queue resets plus the bucket kernels are captured once
// into a hipGraph, so that one graph launch solves the whole batch.
//
// The FP32 factorization is a Gauss-Jordan inversion in place: general
// systems pivot on the largest |a| of the column among the rows not yet
// chosen (lowest row on ties), with virtual interchanges (a pivot row is
// marked done and stays where it is); SPD systems pivot on the diagonal. The
// result M holds inv(A) up to that row order, inv(A)[k][l] =
// M[prow[k]][invp[l]], so every later solve is one parallel matrix-vector
// product instead of two sequences of n dependent steps.
// Refinement per right-hand side: x = 0, r = b; repeat x += inv32(A) r,
// r = b - A x in FP64, until ||r||_inf <= tol ||A||_inf ||x||_inf n u; a
// system whose backward error stops halving, or that runs out of steps,
// reports -1.
#include "zmlhip_internal.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/zmlhip.h"

namespace zmlhip {
namespace {

struct SysDesc {
  const double* a;   // column-major, ld n
  const double* b;   // column-major n x nrhs, ld n
  double* x;
  int n, nrhs, spd, pad;
};

struct BatchArgs {
  const SysDesc* sys;
  const int* list;   // systems of this bucket
  int count;
  int* counter;      // work-queue head (reset before every run)
  int* status;       // per system: refinement steps of its worst right-hand side, or -1
  int max_iter;
  double tol;
};

__device__ __forceinline__ uint64_t bkey(float v, int r) {
  if (v != v) return (uint64_t)(0x7fffffffu - (uint32_t)r);
  return ((uint64_t)__float_as_uint(v) << 32) | (0xffffffffu - (uint32_t)r);
}
__device__ __forceinline__ int bkey_row(uint64_t k) {
  const uint32_t lo = (uint32_t)k;
  return lo >= 0x80000000u ? (int)(0xffffffffu - lo) : (int)(0x7fffffffu - lo);
}

template <int NW>
__device__ __forceinline__ double block_max(double v, double* red, int t) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double m = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) m = fmax(m, red[w]);
  __syncthreads();
  return m;
}

template <int TY, int TX, int RI, int RJ>
__global__ void __launch_bounds__(TY * TX) batched_solve_kernel(BatchArgs g) {
  constexpr int NT = TY * TX, NMAX = TY * RI, NW = (NT + 63) / 64, PARTS = NT / NMAX;
  static_assert(TX * RJ == NMAX && TX <= 64 && RI <= 32, "thread grid shape");
  static_assert(PARTS >= 1 && PARTS * NMAX == NT, "residual split");
  __shared__ uint64_t wbest[NW];
  __shared__ float lcol[NMAX], urow[NMAX], cv[NMAX], yv[NMAX];
  __shared__ int prow[NMAX];
  __shared__ double xv[NMAX], rv[NMAX], red[NW];
  __shared__ double psum[NT], pabs[NT];   // partial residual sums and |a| row sums
  __shared__ int s_idx, s_bad;
  const int t = threadIdx.x, tx = t % TX, ty = t / TX, lane = t & 63, wv = t >> 6;
  for (;;) {
    if (t == 0) s_idx = atomicAdd(g.counter, 1);
    __syncthreads();
    const int idx = s_idx;
    __syncthreads();
    if (idx >= g.count) break;
    const int sid = g.list[idx];
    const SysDesc d = g.sys[sid];
    const int n = d.n;

    // ---- FP32 Gauss-Jordan inversion, in registers
    // a thread row that lies inside n x n is loaded without masks (its RJ
    // loads in flight together); rows at the edge test every element
    float a[RI][RJ];
#pragma unroll
    for (int i = 0; i < RI; ++i) {
      const int row = ty + TY * i;
      if (row < n && tx + TX * (RJ - 1) < n) {
#pragma unroll
        for (int j = 0; j < RJ; ++j) a[i][j] = (float)d.a[row + (int64_t)(tx + TX * j) * n];
      } else {
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          const int col = tx + TX * j;
          a[i][j] = (row < n && col < n) ? (float)d.a[row + (int64_t)col * n] : 0.0f;
        }
      }
    }
    unsigned done = 0;
#pragma unroll
    for (int i = 0; i < RI; ++i)
      if (ty + TY * i >= n) done |= 1u << i;
    if (t == 0) s_bad = 0;
#pragma unroll
    for (int jj = 0; jj < RJ; ++jj) {
      for (int kq = 0; kq < TX; ++kq) {
        const int k = jj * TX + kq;
        if (k >= n) break;
        int p = k;
        if (!d.spd) {
          uint64_t best = 0;
          if (tx == kq) {
#pragma unroll
            for (int i = 0; i < RI; ++i)
              if (!((done >> i) & 1u)) {
                const uint64_t key = bkey(fabsf(a[i][jj]), ty + TY * i);
                best = key > best ? key : best;
              }
          }
#pragma unroll
          for (int off = TX; off < 64; off <<= 1) {
            const uint64_t o = __shfl_xor(best, off);
            best = o > best ? o : best;
          }
          if (lane == kq) wbest[wv] = best;
        }
        // also orders this step's urow / lcol writes after the previous step's reads
        __syncthreads();
        if (!d.spd) {
          uint64_t gk = wbest[0];
#pragma unroll
          for (int w = 1; w < NW; ++w) gk = wbest[w] > gk ? wbest[w] : gk;
          p = bkey_row(gk);
        }
        const int pt = p % TY, pi = p / TY;
        if (ty == pt) {   // the pivot row
#pragma unroll
          for (int i = 0; i < RI; ++i)
            if (i == pi) {
#pragma unroll
              for (int j = 0; j < RJ; ++j) urow[tx + TX * j] = a[i][j];
            }
          done |= 1u << pi;
        }
        if (tx == kq) {   // column k: the multipliers of every other row
#pragma unroll
          for (int i = 0; i < RI; ++i) lcol[ty + TY * i] = (ty + TY * i == p) ? 0.0f : a[i][jj];
        }
        if (t == 0) prow[k] = p;
        __syncthreads();
        const float pv = urow[k];
        if (t == 0 && !(pv != 0.0f && pv == pv)) s_bad = 1;
        const float rp = 1.0f / pv;
        float fi[RI], uj[RJ];
#pragma unroll
        for (int i = 0; i < RI; ++i) fi[i] = lcol[ty + TY * i];
#pragma unroll
        for (int j = 0; j < RJ; ++j) uj[j] = urow[tx + TX * j] * rp;
#pragma unroll
        for (int i = 0; i < RI; ++i)
#pragma unroll
          for (int j = 0; j < RJ; ++j) a[i][j] -= fi[i] * uj[j];
        if (ty == pt) {
#pragma unroll
          for (int i = 0; i < RI; ++i)
            if (i == pi) {
#pragma unroll
              for (int j = 0; j < RJ; ++j) a[i][j] = uj[j];
            }
        }
        if (tx == kq) {
#pragma unroll
          for (int i = 0; i < RI; ++i) a[i][jj] = (ty + TY * i == p) ? rp : -fi[i] * rp;
        }
      }
    }
    __syncthreads();
    // a now holds M with inv(A)[k][l] = M[prow[k]][invp[l]]: x = inv(A) r is
    // c_j = r[prow[j]], y = M c, x_k = y[prow[k]]

    // ---- FP64 refinement of every right-hand side
    int worst = s_bad ? -1 : 0;
    double anorm = 0.0;
    for (int q = 0; q < d.nrhs && worst >= 0; ++q) {
      const double* b = d.b + (int64_t)q * n;
      if (t < NMAX) {
        xv[t] = 0.0;
        rv[t] = t < n ? b[t] : 0.0;
      }
      __syncthreads();
      double prev = INFINITY;
      int it = 0;
      for (;;) {
        if (t < NMAX) cv[t] = t < n ? (float)rv[prow[t]] : 0.0f;
        __syncthreads();
        float acc[RI];
#pragma unroll
        for (int i = 0; i < RI; ++i) {
          float s = 0.0f;
#pragma unroll
          for (int j = 0; j < RJ; ++j) s = fmaf(a[i][j], cv[tx + TX * j], s);
          acc[i] = s;
        }
#pragma unroll
        for (int off = TX / 2; off > 0; off >>= 1)
#pragma unroll
          for (int i = 0; i < RI; ++i) acc[i] += __shfl_xor(acc[i], off);
        if (tx == 0) {
#pragma unroll
          for (int i = 0; i < RI; ++i) yv[ty + TY * i] = acc[i];
        }
        __syncthreads();
        if (t < n) xv[t] += (double)yv[prow[t]];
        __syncthreads();
        // r = b - A x, ||A||_inf on the first pass. Thread t takes row t % NMAX
        // and the columns t / NMAX + PARTS l, eight loads in flight at a time (a
        // tail of fewer than eight one by one); the PARTS partial sums of a row
        // are then added in a fixed order.
        {
          const int row = t % NMAX, part = t / NMAX;
          double accd = 0.0, asum = 0.0;
          if (row < n) {
            int j0 = part;
            for (; j0 + 7 * PARTS < n; j0 += 8 * PARTS) {   // eight whole columns: no masks
              double v[8];
#pragma unroll
              for (int u = 0; u < 8; ++u) v[u] = d.a[row + (int64_t)(j0 + u * PARTS) * n];
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                accd -= v[u] * xv[j0 + u * PARTS];
                asum += fabs(v[u]);
              }
            }
            for (; j0 < n; j0 += PARTS) {
              const double v = d.a[row + (int64_t)j0 * n];
              accd -= v * xv[j0];
              asum += fabs(v);
            }
          }
          psum[t] = accd;
          pabs[t] = asum;
        }
        __syncthreads();
        double ri = 0.0, ai = 0.0;
        if (t < n) {
          double accd = b[t], asum = 0.0;
#pragma unroll
          for (int q = 0; q < PARTS; ++q) {
            accd += psum[q * NMAX + t];
            asum += pabs[q * NMAX + t];
          }
          rv[t] = accd;
          ri = fabs(accd);
          ai = asum;
        }
        if (anorm == 0.0) anorm = block_max<NW>(ai, red, t);
        const double rmax = block_max<NW>(ri, red, t);
        const double xmax = block_max<NW>(t < n ? fabs(xv[t]) : 0.0, red, t);
        ++it;
        const double be = rmax / (anorm * xmax * (double)n * 0x1p-53);
        if (be <= g.tol) break;
        if (!(be == be) || it >= g.max_iter || (it >= 2 && be > 0.5 * prev)) {
          it = -1;
          break;
        }
        prev = be;
      }
      if (t < n) d.x[(int64_t)q * n + t] = xv[t];
      worst = it < 0 ? -1 : std::max(worst, it);
      __syncthreads();
    }
    if (t == 0) g.status[sid] = worst;
    __syncthreads();
  }
}

__global__ void reset_counters_kernel(int* c, int k) {
  if ((int)threadIdx.x < k) c[threadIdx.x] = 0;
}

// ---- variants and buckets -------------------------------------------------------------
struct Variant {
  int ty, tx, ri, rj;   // thread grid TY x TX, RI x RJ values per thread
  void (*kernel)(BatchArgs);
};
const Variant kVariants[] = {
    {8, 8, 4, 4, batched_solve_kernel<8, 8, 4, 4>},        // n <= 32, one wave
    {8, 8, 8, 8, batched_solve_kernel<8, 8, 8, 8>},        // n <= 64, one wave
    {16, 16, 4, 4, batched_solve_kernel<16, 16, 4, 4>},    // n <= 64, four waves
    {16, 16, 8, 8, batched_solve_kernel<16, 16, 8, 8>},    // n <= 128
    {32, 32, 4, 4, batched_solve_kernel<32, 32, 4, 4>},    // n <= 128, sixteen waves
    {32, 32, 8, 8, batched_solve_kernel<32, 32, 8, 8>},    // n <= 256, sixteen waves (register spills)
    {16, 32, 16, 8, batched_solve_kernel<16, 32, 16, 8>},  // n <= 256, eight waves
};
constexpr int kBuckets = 4;
const int kBucketMax[kBuckets] = {32, 64, 128, 256};
const std::vector<int> kCandidates[kBuckets] = {{0, 1, 2}, {1, 2}, {3, 4}, {6, 5}};

}  // namespace

struct BatchPlan {
  int count = 0;
  SysDesc* d_sys = nullptr;
  int* d_lists = nullptr;
  int* d_counters = nullptr;
  int* d_status = nullptr;
  int bucket_count[kBuckets] = {};
  int bucket_offset[kBuckets] = {};
  int variant[kBuckets] = {};
  unsigned grid[kBuckets] = {};
  BatchArgs args[kBuckets];
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : {(void*)d_sys, (void*)d_lists, (void*)d_counters, (void*)d_status})
      if (p) (void)hipFree(p);
  }
};

namespace {

unsigned resident_grid(int v) {
  int per_cu = 0;
  const int nt = kVariants[v].ty * kVariants[v].tx;
  ZH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kVariants[v].kernel),
                                                        nt, 0));
  return (unsigned)(std::max(per_cu, 1) * device_cu_count());
}

void launch_bucket(const BatchPlan& P, int b, int v, hipStream_t s) {
  const int nt = kVariants[v].ty * kVariants[v].tx;
  const unsigned grid = std::min<unsigned>(P.grid[b], (unsigned)std::max(1, P.bucket_count[b]));
  BatchArgs args = P.args[b];
  void* kargs[] = {&args};
  ZH_CHECK(hipLaunchKernel(reinterpret_cast<const void*>(kVariants[v].kernel), dim3(grid), dim3(nt), kargs, 0, s));
}

}  // namespace

BatchPlan* batch_plan_create(int count, const int* n, const int* nrhs, const int* spd,
                             const double* const* a, const double* const* b, double* const* x,
                             const zmlhip_ir_opts* opts, int autotune, hipStream_t s) {
  std::unique_ptr<BatchPlan> P(new BatchPlan);
  P->count = count;
  std::vector<SysDesc> h(count);
  std::vector<std::vector<int>> lists(kBuckets);
  for (int i = 0; i < count; ++i) {
    h[i] = SysDesc{a[i], b[i], x[i], n[i], nrhs[i], spd ? (spd[i] ? 1 : 0) : 0, 0};
    int bk = 0;
    while (bk < kBuckets - 1 && n[i] > kBucketMax[bk]) ++bk;
    lists[bk].push_back(i);
  }
  std::vector<int> flat;
  for (int bk = 0; bk < kBuckets; ++bk) {
    // larger systems first inside a bucket: the queue ends on the short ones
    std::stable_sort(lists[bk].begin(), lists[bk].end(), [&](int u, int v) { return n[u] > n[v]; });
    P->bucket_offset[bk] = (int)flat.size();
    P->bucket_count[bk] = (int)lists[bk].size();
    flat.insert(flat.end(), lists[bk].begin(), lists[bk].end());
  }
  ZH_CHECK(hipMalloc(&P->d_sys, sizeof(SysDesc) * std::max(count, 1)));
  ZH_CHECK(hipMalloc(&P->d_lists, sizeof(int) * std::max<size_t>(flat.size(), 1)));
  ZH_CHECK(hipMalloc(&P->d_counters, sizeof(int) * 64));
  ZH_CHECK(hipMalloc(&P->d_status, sizeof(int) * std::max(count, 1)));
  ZH_CHECK(hipMemcpyAsync(P->d_sys, h.data(), sizeof(SysDesc) * count, hipMemcpyHostToDevice, s));
  if (!flat.empty())
    ZH_CHECK(hipMemcpyAsync(P->d_lists, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice, s));
  const int max_iter = opts && opts->max_iter > 0 ? opts->max_iter : 10;
  const double tol = opts && opts->tol > 0.0 ? opts->tol : 1.0;
  for (int bk = 0; bk < kBuckets; ++bk) {
    P->args[bk] = BatchArgs{P->d_sys, P->d_lists + P->bucket_offset[bk], P->bucket_count[bk],
                            P->d_counters + bk, P->d_status, max_iter, tol};
    P->variant[bk] = kCandidates[bk].back();
  }
  // autotune: time every candidate shape on the bucket's own systems
  for (int bk = 0; bk < kBuckets; ++bk) {
    const auto& cand = kCandidates[bk];
    if (P->bucket_count[bk] == 0) continue;
    float best = INFINITY;
    if (!autotune || cand.size() == 1) {
      P->grid[bk] = resident_grid(P->variant[bk]);
      continue;
    }
    hipEvent_t e0, e1;
    ZH_CHECK(hipEventCreate(&e0));
    ZH_CHECK(hipEventCreate(&e1));
    for (int v : cand) {
      P->grid[bk] = resident_grid(v);
      float ms = 0.0f;
      for (int rep = 0; rep < 2; ++rep) {   // the first run warms up
        hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, P->d_counters, kBuckets);
        ZH_CHECK(hipEventRecord(e0, s));
        launch_bucket(*P, bk, v, s);
        ZH_CHECK(hipEventRecord(e1, s));
        ZH_CHECK(hipEventSynchronize(e1));
        ZH_CHECK(hipEventElapsedTime(&ms, e0, e1));
      }
      if (ms < best) {
        best = ms;
        P->variant[bk] = v;
      }
    }
    P->grid[bk] = resident_grid(P->variant[bk]);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  // one graph: reset the queues, then the buckets as independent branches (their
  // workgroups only take work from their own queues, so the small buckets can use
  // the CUs the large one leaves idle as its queue drains)
  hipStream_t cs[kBuckets];
  hipEvent_t ev[kBuckets + 1];
  for (int i = 0; i < kBuckets; ++i) ZH_CHECK(hipStreamCreateWithFlags(&cs[i], hipStreamNonBlocking));
  for (int i = 0; i <= kBuckets; ++i) ZH_CHECK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
  ZH_CHECK(hipStreamBeginCapture(cs[0], hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, cs[0], P->d_counters, kBuckets);
  ZH_CHECK(hipEventRecord(ev[kBuckets], cs[0]));
  for (int i = 1; i < kBuckets; ++i) ZH_CHECK(hipStreamWaitEvent(cs[i], ev[kBuckets], 0));
  for (int bk = kBuckets - 1, i = 0; bk >= 0; --bk, ++i)   // the largest bucket on the capturing stream
    if (P->bucket_count[bk] > 0) launch_bucket(*P, bk, P->variant[bk], cs[i]);
  for (int i = 1; i < kBuckets; ++i) {
    ZH_CHECK(hipEventRecord(ev[i], cs[i]));
    ZH_CHECK(hipStreamWaitEvent(cs[0], ev[i], 0));
  }
  ZH_CHECK(hipStreamEndCapture(cs[0], &P->graph));
  for (int i = 0; i < kBuckets; ++i) ZH_CHECK(hipStreamDestroy(cs[i]));
  for (int i = 0; i <= kBuckets; ++i) ZH_CHECK(hipEventDestroy(ev[i]));
  ZH_CHECK(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
  ZH_CHECK(hipStreamSynchronize(s));
  return P.release();
}

void batch_plan_execute(BatchPlan* P, hipStream_t s) { ZH_CHECK(hipGraphLaunch(P->exec, s)); }

void batch_plan_status(BatchPlan* P, int* status, hipStream_t s) {
  ZH_CHECK(hipMemcpyAsync(status, P->d_status, sizeof(int) * P->count, hipMemcpyDeviceToHost, s));
  ZH_CHECK(hipStreamSynchronize(s));
}

void batch_plan_variants(const BatchPlan* P, int* tg, int* r, int* counts) {
  for (int bk = 0; bk < kBuckets; ++bk) {
    tg[bk] = kVariants[P->variant[bk]].ty;
    r[bk] = kVariants[P->variant[bk]].ri;
    counts[bk] = P->bucket_count[bk];
  }
}

void batch_plan_destroy(BatchPlan* P) { delete P; }

}  // namespace zmlhip
