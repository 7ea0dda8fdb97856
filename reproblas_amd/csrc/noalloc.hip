// noalloc.hip -- reproblas_hip_rddot_device_noalloc: reproblas_hip_rddot_device
// with its scratch (the completion ticket and the grid accumulators through
// which the workgroups merge) in a fixed device array of this file instead
// of the library's grown-on-demand workspace, for callers that must not
// allocate (code under hipGraph capture, tight serving loops).
//
// The array lives in the code object: every GPU the library runs on holds
// one copy, at its own address, until the process ends. Host code never
// passes the array's name to a kernel (in host code it names a host-side
// stand-in); the entry point looks up each device's address once with
// hipGetSymbolAddress and passes that.
#include <mutex>

#include "../../include/reproblas_hip.h"
#include "api_helpers.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rb {
namespace {

// 64 KiB: the 32 KiB sync block (ticket and grid accumulators, which the kernels
// leave zeroed) in the last kBlas1SyncBytes; the rest would hold workgroup
// records, which the dot does not use (it merges through the accumulators
// whatever the grid), so the dot needs no more than this at any size.
constexpr size_t kScratchBytes = 64 * 1024;
constexpr size_t kRecordBytes = kScratchBytes - kBlas1SyncBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void clear_sync_kernel(unsigned* sync) {
  for (size_t i = threadIdx.x; i < kBlas1SyncBytes / sizeof(unsigned); i += blockDim.x) sync[i] = 0u;
}

struct DeviceScratch {
  unsigned char* ptr = nullptr;  // device address of `scratch` on that GPU
};
std::mutex g_scratch_mu;
DeviceScratch g_scratch[64];

// Address of this device's array; looked up (and the ticket cleared) on the
// first call on the device, with the device current.
unsigned char* device_scratch(hipStream_t s) {
  int dev = -1;
  if (!RB_HIP(hipGetDevice(&dev))) return nullptr;
  if (dev < 0 || dev >= 64) {
    set_error(kNoDevice, "device ordinal %d out of range", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  DeviceScratch& d = g_scratch[dev];
  if (!d.ptr) {
    void* p = nullptr;
    if (!RB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return nullptr;
    if (!p) {
      set_error(kHipError, "hipGetSymbolAddress returned no address for the scratch array");
      return nullptr;
    }
    // the kernels leave the sync block zeroed; start it there (a kernel, not
    // a memset, so the call may also be captured into a graph)
    clear_sync_kernel<<<1, 256, 0, s>>>((unsigned*)((unsigned char*)p + kRecordBytes));
    if (!RB_HIP(hipGetLastError())) return nullptr;
    d.ptr = (unsigned char*)p;
  }
  return d.ptr;
}

}  // namespace
}  // namespace rb

using namespace rb;

extern "C" {

int reproblas_hip_rddot_device_noalloc(int fold, int64_t N, const double* X, int64_t incX, const double* Y,
                                       int64_t incY, double* result, void* stream) {
  clear_error();
  const hipStream_t s = (hipStream_t)stream;
  if (!api::check_fold<double>(fold)) return last_error();
  if (N < 0) N = 0;
  if (N > 0 && (!require_device_accessible(X, "X") || !require_device_accessible(Y, "Y"))) return last_error();
  if (!require_device_accessible(result, "result")) return last_error();
  unsigned char* p = device_scratch(s);
  if (!p) return last_error();
  if (!blas1_scratch(fold, kDot, N, X, incX, Y, incY, result, p, kRecordBytes, (unsigned*)(p + kRecordBytes), s))
    return last_error();
  return kOk;
}

}  // extern "C"
