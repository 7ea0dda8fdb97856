// runtime.hip -- implementation of runtime.hpp.
#include "runtime.hpp"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>

namespace rb {

namespace {
thread_local int g_err = kOk;
thread_local char g_msg[512] = "";
std::mutex g_mu;
thread_local hipStream_t g_stream = nullptr;

struct Buf {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::map<std::tuple<int, hipStream_t, int>, Buf>& buffers() {
  static auto* m = new std::map<std::tuple<int, hipStream_t, int>, Buf>();
  return *m;
}
std::map<std::pair<int, hipStream_t>, Buf>& ticket_bufs() {
  static auto* m = new std::map<std::pair<int, hipStream_t>, Buf>();
  return *m;
}
DeviceInfo g_info[64];

// One recursive mutex per (device, stream), created on first use and kept
// (a stream handle that is destroyed and reused maps to the same mutex).
std::recursive_mutex* stream_mutex(hipStream_t s) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    dev = -1;
  }
  // the last (device, stream) this thread used: no global lock on repeated calls
  thread_local int last_dev = -2;
  thread_local hipStream_t last_s = nullptr;
  thread_local std::recursive_mutex* last_m = nullptr;
  if (last_m && last_dev == dev && last_s == s) return last_m;
  std::lock_guard<std::mutex> lk(g_mu);
  static auto* m = new std::map<std::pair<int, hipStream_t>, std::recursive_mutex*>();
  std::recursive_mutex*& p = (*m)[std::make_pair(dev, s)];
  if (!p) p = new std::recursive_mutex();
  last_dev = dev;
  last_s = s;
  last_m = p;
  return p;
}
}  // namespace

StreamGuard::StreamGuard(hipStream_t s) : mu_(stream_mutex(s)) { static_cast<std::recursive_mutex*>(mu_)->lock(); }
StreamGuard::~StreamGuard() { static_cast<std::recursive_mutex*>(mu_)->unlock(); }

void set_error(int code, const char* fmt, ...) {
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  fprintf(stderr, "reproblas_hip: %s\n", g_msg);
}

void clear_error() {
  g_err = kOk;
  g_msg[0] = 0;
  (void)hipGetLastError();  // drop a stale HIP error left by an earlier failed call
}
int last_error() { return g_err; }
const char* last_error_message() { return g_msg; }

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  set_error(kHipError, "%s failed: %s", what, hipGetErrorString(e));
  return false;
}

hipStream_t current_stream() { return g_stream; }
void set_current_stream(hipStream_t s) { g_stream = s; }

const DeviceInfo* device_info() {
  int dev = -1;
  if (!RB_HIP(hipGetDevice(&dev))) return nullptr;
  if (dev < 0 || dev >= 64) {
    set_error(kNoDevice, "device ordinal %d out of range", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceInfo& d = g_info[dev];
  if (d.device < 0) {
    hipDeviceProp_t prop;
    if (!RB_HIP(hipGetDeviceProperties(&prop, dev))) return nullptr;
    d.cus = prop.multiProcessorCount;
    d.warp = prop.warpSize;
    d.lds_per_block = prop.sharedMemPerBlock;
    d.device = dev;
  }
  return &d;
}

void* workspace(hipStream_t s, int slot, size_t bytes) {
  int dev = -1;
  if (!RB_HIP(hipGetDevice(&dev))) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  Buf& b = buffers()[std::make_tuple(dev, s, slot)];
  if (b.bytes < bytes) {
    if (b.ptr) {
      // the stream may still read the old buffer
      if (!RB_HIP(hipStreamSynchronize(s))) return nullptr;
      if (!RB_HIP(hipFree(b.ptr))) return nullptr;
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    if (!RB_HIP(hipMalloc(&b.ptr, want))) return nullptr;
    b.bytes = want;
  }
  return b.ptr;
}

unsigned* tickets(hipStream_t s, int count) {
  int dev = -1;
  if (!RB_HIP(hipGetDevice(&dev))) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  Buf& b = ticket_bufs()[std::make_pair(dev, s)];
  size_t bytes = sizeof(unsigned) * (size_t)count;
  if (b.bytes < bytes) {
    void* p = nullptr;
    size_t want = bytes < 1024 ? 1024 : bytes;
    if (!RB_HIP(hipMalloc(&p, want))) return nullptr;
    // synchronous clear at creation; kernels keep the counters at zero after
    if (!RB_HIP(hipMemset(p, 0, want))) return nullptr;
    if (b.ptr) {
      if (!RB_HIP(hipStreamSynchronize(s))) return nullptr;
      if (!RB_HIP(hipFree(b.ptr))) return nullptr;
    }
    b.ptr = p;
    b.bytes = want;
  }
  return (unsigned*)b.ptr;
}

bool is_device_pointer(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory reports an error; clear it
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

bool is_device_accessible(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  memset(&attr, 0, sizeof(attr));
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) return true;
  if (attr.type != hipMemoryTypeHost || attr.devicePointer == nullptr) return false;
  // pinned host memory: the runtime must also hand out a device address for it
  void* dp = nullptr;
  if (hipHostGetDevicePointer(&dp, const_cast<void*>(p), 0) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return dp != nullptr;
}

bool require_device_accessible(const void* p, const char* what, bool may_be_null) {
  if (!p && may_be_null) return true;
  if (is_device_accessible(p)) return true;
  set_error(kBadArgument, "%s is not device-accessible memory (device, managed or pinned host)", what);
  return false;
}

}  // namespace rb
