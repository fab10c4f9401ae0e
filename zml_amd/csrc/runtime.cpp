// Per-thread runtime state of libzmlhip: last-error text, the stream of the
// synchronous host-pointer API and growable device workspaces.
#include "zmlhip_internal.hpp"

#include <mutex>
#include <string>
#include <unordered_map>

namespace zmlhip {

namespace {
thread_local std::string tl_last_error;

struct Buf {
  void* ptr = nullptr;
  size_t bytes = 0;
};
constexpr int kSlots = 8;
constexpr int kAux = 4, kEvents = 8;
struct DeviceState {
  Buf slot[kSlots];
  hipStream_t stream = nullptr;
  hipStream_t aux[kAux] = {};
  hipEvent_t ev[kEvents] = {};
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
      for (auto& e : kv.second.ev)
        if (e) (void)hipEventDestroy(e);
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
const char* last_error() { return tl_last_error.c_str(); }

void* workspace(size_t bytes, int slot) {
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

hipEvent_t aux_event(int idx) {
  DeviceState& ds = current();
  hipEvent_t& e = ds.ev[idx % kEvents];
  if (!e) ZH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return e;
}

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
