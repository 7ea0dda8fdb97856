// runtime.hpp -- host-side runtime of the HIP library: error state, the
// library stream, device properties, cached device workspaces and staging of
// host operands.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rb {

// Error codes reported through reproblas_hip_last_error().
enum Status : int {
  kOk = 0,
  kBadFold = 1,      // fold outside the range the library supports
  kBadArgument = 2,  // negative size, bad leading dimension, null pointer
  kHipError = 3,     // a HIP runtime call failed
  kNoDevice = 4,     // no usable GPU
  kUnsupported = 5,  // a case the library does not handle
};

void set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
int last_error();
const char* last_error_message();

// Checks a HIP call; on failure records kHipError with the call site and
// returns false.
bool hip_ok(hipError_t e, const char* what);
#define RB_HIP(call) ::rb::hip_ok((call), #call)

// Stream used by the synchronous entry points of the calling host thread
// (default: the null stream). Each thread has its own setting.
hipStream_t current_stream();
void set_current_stream(hipStream_t s);

// Serialises the host side of library calls that share a (device, stream):
// every entry point holds it from its first workspace() / tickets() fetch
// through its last enqueue (synchronous calls: through their final
// synchronize), so concurrent host threads never stage into, or regrow and
// free, a buffer another thread's call is still using. Recursive: an entry
// point may call another under the same guard. Calls on different streams
// do not contend.
class StreamGuard {
 public:
  explicit StreamGuard(hipStream_t s);
  ~StreamGuard();
  StreamGuard(const StreamGuard&) = delete;
  StreamGuard& operator=(const StreamGuard&) = delete;

 private:
  void* mu_;
};
#define RB_STREAM_GUARD(s) ::rb::StreamGuard rb_stream_guard_((hipStream_t)(s))

struct DeviceInfo {
  int device = -1;
  int cus = 0;
  int warp = 64;
  size_t lds_per_block = 0;
};
// Properties of the current device (cached).
const DeviceInfo* device_info();

// Device memory of at least `bytes` owned by the library, one slot per
// (device, stream, slot id), grown on demand and kept until process exit.
// Calls that share a stream are ordered by that stream on the device, and
// their host sides are serialised by StreamGuard, so sharing a slot between
// them is safe. Callers must hold StreamGuard(s).
void* workspace(hipStream_t s, int slot, size_t bytes);

// Zero-initialised device counters (one per slot) used by the single-pass
// grid reductions; every kernel that takes one leaves it at zero. Callers
// must hold StreamGuard(s).
unsigned* tickets(hipStream_t s, int count);

// Pointer kind of `p`: true when a kernel may dereference it directly.
bool is_device_pointer(const void* p);
// True when a kernel may dereference p: device or managed memory, or pinned
// host memory mapped for the GPU. Pageable host memory is not.
bool is_device_accessible(const void* p);
// Sets kBadArgument naming `what` and returns false unless p is device
// accessible (null passes when `may_be_null`).
bool require_device_accessible(const void* p, const char* what, bool may_be_null = false);

enum WorkspaceSlot : int {
  kSlotPartials = 0,
  kSlotResult = 1,
  kSlotStageX = 2,
  kSlotStageY = 3,
  kSlotStageZ = 4,
  kSlotBinnedOut = 5,
  kSlotBatch = 6,
  kSlotGemmMax = 7,  // per-chunk row/column maxima of the gemm operands
  kSlotGemvX = 8,    // fl(alpha * x) gathered contiguously for the row-dot gemv
  kSlotZA = 9,       // complex gemm/gemv: real image of op(A)
  kSlotZB = 10,      // complex gemm: real image of op(B); gemv: the two right-hand vectors
  kSlotZC = 11,      // complex gemm/gemv: the outputs' real and imaginary binned values
  kSlotCount = 12,
};

}  // namespace rb
