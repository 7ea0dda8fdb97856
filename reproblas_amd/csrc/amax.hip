// amax.hip -- binnedBLAS_damax / damaxm / samax / samaxm on the GPU: max |x_i|
// and max |fl(x_i*y_i)| (reference src/binnedBLAS/damax.ccog:45,
// damaxm.ccog:47, samax.ccog, samaxm.ccog). The max ignores NaN terms (IEEE maxNum); the reference's
// answer on NaN input depends on which of its SIMD sections ran, so no bit
// pattern is promised there.
#include "../../include/binnedBLAS.h"
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace rb;

namespace {

// |t| as an unsigned key: for non-negative values the bit patterns order like
// the values, so an atomic max of keys is an exact, order-free max.
template <class F>
__global__ __launch_bounds__(256) void amax_kernel(int64_t n, const F* x, int64_t incx, const F* y, int64_t incy,
                                                   unsigned long long* key) {
  F m = F(0);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const F t = y ? x[i * incx] * y[i * incy] : x[i * incx];
    m = fmax(m, fabs(t));
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) m = fmax(m, __shfl_xor(m, s, 64));
  if ((threadIdx.x & 63) == 0 && m > F(0)) atomicMax(key, (unsigned long long)bits(m));
}

__global__ void clear_key(unsigned long long* key) { *key = 0ull; }

template <class F>
F device_amax(int N, const F* X, int incX, const F* Y, int incY) {
  clear_error();
  if (N <= 0) return F(0);
  hipStream_t s = current_stream();
  RB_STREAM_GUARD(s);
  const bool dev = is_device_pointer(X) && (!Y || is_device_pointer(Y));
  const int64_t ax = incX < 0 ? -(int64_t)incX : incX, ay = incY < 0 ? -(int64_t)incY : incY;
  const F* dx = X;
  const F* dy = Y;
  if (!dev) {
    // stage the spans (host operands)
    const int64_t sx = (int64_t)(N - 1) * ax + 1;
    F* bx = (F*)workspace(s, kSlotStageX, sizeof(F) * sx);
    if (!bx) return F(0);
    if (!RB_HIP(hipMemcpyAsync(bx, incX < 0 ? X + (int64_t)(N - 1) * incX : X, sizeof(F) * sx,
                               hipMemcpyHostToDevice, s)))
      return F(0);
    dx = incX < 0 ? bx + (N - 1) * ax : bx;
    if (Y) {
      const int64_t sy = (int64_t)(N - 1) * ay + 1;
      F* by = (F*)workspace(s, kSlotStageY, sizeof(F) * sy);
      if (!by) return F(0);
      if (!RB_HIP(hipMemcpyAsync(by, incY < 0 ? Y + (int64_t)(N - 1) * incY : Y, sizeof(F) * sy,
                                 hipMemcpyHostToDevice, s)))
        return F(0);
      dy = incY < 0 ? by + (N - 1) * ay : by;
    }
  }
  unsigned long long* key = (unsigned long long*)workspace(s, kSlotResult, 64);
  if (!key) return F(0);
  clear_key<<<1, 1, 0, s>>>(key);
  const DeviceInfo* d = device_info();
  if (!d) return F(0);
  int64_t blocks = ((int64_t)N + 255) / 256;
  if (blocks > d->cus * 4) blocks = d->cus * 4;
  amax_kernel<F><<<(unsigned)blocks, 256, 0, s>>>(N, dx, incX, dy, incY, key);
  if (!RB_HIP(hipGetLastError())) return F(0);
  unsigned long long h = 0;
  if (!RB_HIP(hipMemcpyAsync(&h, key, sizeof(h), hipMemcpyDeviceToHost, s))) return F(0);
  if (!RB_HIP(hipStreamSynchronize(s))) return F(0);
  return fp_from_bits<F>((typename Fmt<F>::U)h);
}

}  // namespace

extern "C" {
double binnedBLAS_damax(const int N, const double* X, const int incX) {
  return device_amax<double>(N, X, incX, nullptr, 0);
}
double binnedBLAS_damaxm(const int N, const double* X, const int incX, const double* Y, const int incY) {
  return device_amax(N, X, incX, Y, incY);
}
float binnedBLAS_samax(const int N, const float* X, const int incX) { return device_amax<float>(N, X, incX, nullptr, 0); }
float binnedBLAS_samaxm(const int N, const float* X, const int incX, const float* Y, const int incY) {
  return device_amax(N, X, incX, Y, incY);
}
}
