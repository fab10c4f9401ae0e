// Per-thread runtime state of libzmlhip: last-error text, the stream of the
// synchronous host-pointer API and growable device workspaces.
#include "zmlhip_internal.hpp"

#include <algorithm>
#include <atomic>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace zmlhip {

namespace {
thread_local std::string tl_last_error;

struct Buf {
  void* ptr = nullptr;
  size_t bytes = 0;
};
constexpr int kSlots = 32;
constexpr int kAux = 4, kEvents = 12;
constexpr int kSparing = 8;   // CU-sparing streams kept per device (distinct reservations)
struct DeviceState {
  Buf slot[kSlots];
  hipStream_t stream = nullptr;
  hipStream_t aux[kAux] = {};
  int reserved_cus[kSparing] = {};
  hipStream_t reserved[kSparing] = {};   // every CU but the first reserved_cus[i]
  hipEvent_t ev[kEvents] = {};
  hipEvent_t fence = nullptr;   // end of the previous call of this thread on this device
  bool fence_set = false;
};
struct ThreadState {
  std::unordered_map<int, DeviceState> dev;
  ~ThreadState() {
    for (auto& kv : dev) {
      int cur = 0;
      if (hipGetDevice(&cur) != hipSuccess) continue;
      if (hipSetDevice(kv.first) != hipSuccess) continue;
      for (auto& b : kv.second.slot)
        if (b.ptr) (void)hipFree(b.ptr);
      if (kv.second.stream) (void)hipStreamDestroy(kv.second.stream);
      for (auto& st : kv.second.aux)
        if (st) (void)hipStreamDestroy(st);
      for (auto& st : kv.second.reserved)
        if (st) (void)hipStreamDestroy(st);
      for (auto& e : kv.second.ev)
        if (e) (void)hipEventDestroy(e);
      if (kv.second.fence) (void)hipEventDestroy(kv.second.fence);
      (void)hipSetDevice(cur);
    }
  }
};
thread_local ThreadState tl_state;

DeviceState& current() {
  int d = 0;
  ZH_CHECK(hipGetDevice(&d));
  return tl_state.dev[d];
}
}  // namespace

void set_last_error(const std::string& msg) { tl_last_error = msg; }

// ---- optional timing of the dominant kernels (HIP events on their streams) ----
namespace {
struct ProfRec {
  int kind;
  hipEvent_t a, b;
  double flops, bytes;
};
thread_local bool tl_prof_on = false;
thread_local std::vector<ProfRec> tl_prof;
}  // namespace

bool profiling() { return tl_prof_on; }
void profile_enable(bool on) { tl_prof_on = on; }
void* profile_begin(int kind, hipStream_t s) {
  if (!tl_prof_on) return nullptr;
  ProfRec r{kind, nullptr, nullptr, 0.0, 0.0};
  ZH_CHECK(hipEventCreate(&r.a));
  ZH_CHECK(hipEventCreate(&r.b));
  ZH_CHECK(hipEventRecord(r.a, s));
  tl_prof.push_back(r);
  return reinterpret_cast<void*>(tl_prof.size());
}
void profile_end(void* token, hipStream_t s, double flops, double bytes) {
  if (!token) return;
  ProfRec& r = tl_prof[reinterpret_cast<size_t>(token) - 1];
  ZH_CHECK(hipEventRecord(r.b, s));
  r.flops = flops;
  r.bytes = bytes;
}
void profile_collect(int kind, double* ms, long long* launches, double* flops, double* bytes) {
  double t = 0, f = 0, b = 0;
  long long c = 0;
  std::vector<ProfRec> keep;
  for (auto& r : tl_prof) {
    if (r.kind != kind) { keep.push_back(r); continue; }
    ZH_CHECK(hipEventSynchronize(r.b));
    float e = 0.f;
    ZH_CHECK(hipEventElapsedTime(&e, r.a, r.b));
    t += e; f += r.flops; b += r.bytes; ++c;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  tl_prof.swap(keep);
  *ms = t; *launches = c; *flops = f; *bytes = b;
}
const char* last_error() { return tl_last_error.c_str(); }

void* workspace(size_t bytes, int slot) {
  if (slot < 0 || slot >= kSlots) throw std::runtime_error("workspace: slot out of range");
  DeviceState& ds = current();
  Buf& b = ds.slot[slot];
  if (b.bytes < bytes) {
    if (b.ptr) {
      // previous users of this slot are stream-ordered on this thread's streams;
      // the device must be idle before the memory is released
      ZH_CHECK(hipDeviceSynchronize());
      ZH_CHECK(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t sz = align_up(bytes, 1 << 20);
    hipError_t e = hipMalloc(&b.ptr, sz);
    if (e != hipSuccess) {
      b.ptr = nullptr;
      throw HipError(e, "workspace allocation failed");
    }
    b.bytes = sz;
  }
  return b.ptr;
}

hipStream_t host_api_stream() {
  DeviceState& ds = current();
  if (!ds.stream) ZH_CHECK(hipStreamCreateWithFlags(&ds.stream, hipStreamNonBlocking));
  return ds.stream;
}

hipStream_t aux_stream(int idx) {
  DeviceState& ds = current();
  hipStream_t& st = ds.aux[idx % kAux];
  if (!st) ZH_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  return st;
}

hipStream_t stream_sparing_cus(int reserved) {
  DeviceState& ds = current();
  const int cus = device_cu_count();
  reserved = std::max(1, std::min(reserved, cus - 1));
  for (int i = 0; i < kSparing; ++i)
    if (ds.reserved[i] && ds.reserved_cus[i] == reserved) return ds.reserved[i];
  for (int i = 0; i < kSparing; ++i) {
    if (ds.reserved[i]) continue;
    std::vector<uint32_t> mask((cus + 31) / 32, 0u);
    for (int c = reserved; c < cus; ++c) mask[c / 32] |= 1u << (c % 32);
    ZH_CHECK(hipExtStreamCreateWithCUMask(&ds.reserved[i], (uint32_t)mask.size(), mask.data()));
    ds.reserved_cus[i] = reserved;
    return ds.reserved[i];
  }
  throw std::runtime_error("stream_sparing_cus: too many distinct reservations");
}

hipEvent_t aux_event(int idx) {
  DeviceState& ds = current();
  hipEvent_t& e = ds.ev[idx % kEvents];
  if (!e) ZH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return e;
}

void call_begin(hipStream_t s) {
  DeviceState& ds = current();
  if (ds.fence_set) ZH_CHECK(hipStreamWaitEvent(s, ds.fence, 0));
}

void call_end(hipStream_t s) {
  DeviceState& ds = current();
  if (!ds.fence) ZH_CHECK(hipEventCreateWithFlags(&ds.fence, hipEventDisableTiming));
  ZH_CHECK(hipEventRecord(ds.fence, s));
  ds.fence_set = true;
}

namespace {
std::atomic<long> g_panel_spin_limit{2000000L};
}
long panel_spin_limit() { return g_panel_spin_limit.load(std::memory_order_relaxed); }
void set_panel_spin_limit(long polls) { g_panel_spin_limit.store(polls, std::memory_order_relaxed); }

namespace {
constexpr long kOneLaunchRowsDefault = 16384;
std::atomic<long> g_one_launch_rows{kOneLaunchRowsDefault};
}
long panel_one_launch_rows() { return g_one_launch_rows.load(std::memory_order_relaxed); }
void set_panel_one_launch_rows(long rows) {
  g_one_launch_rows.store(rows < 0 ? kOneLaunchRowsDefault : rows, std::memory_order_relaxed);
}

namespace {
std::atomic<int> g_gemm_dma{1};
std::atomic<int> g_leaf_la{1};
}
bool leaf_lookahead() { return g_leaf_la.load(std::memory_order_relaxed) != 0; }
void set_leaf_lookahead(int on) { g_leaf_la.store(on ? 1 : 0, std::memory_order_relaxed); }
bool gemm_dma_enabled() { return g_gemm_dma.load(std::memory_order_relaxed) != 0; }
void set_gemm_dma(int on) { g_gemm_dma.store(on ? 1 : 0, std::memory_order_relaxed); }

int device_cu_count() {
  int d = 0;
  ZH_CHECK(hipGetDevice(&d));
  static std::mutex mu;
  static std::unordered_map<int, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(d);
  if (it != cache.end()) return it->second;
  int cus = 0;
  ZH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d));
  cache[d] = cus;
  return cus;
}

}  // namespace zmlhip
