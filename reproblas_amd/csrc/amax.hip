// amax.hip -- binnedBLAS_damax / damaxm / samax / samaxm on the GPU: max |x_i|
// and max |fl(x_i*y_i)| (reference src/binnedBLAS/damax.ccog:45,
// damaxm.ccog:47, samax.ccog, samaxm.ccog).
//
// NaN terms follow the reference's scalar sections (damax.ccog:282-300,
// damaxm.ccog:355-366): the running `m = (m > t ? m : t)` from 0 takes a NaN
// term and drops it at the next term, so the result is the max of the terms
// after the last NaN term, or that NaN when it is the last term. (The
// reference's SIMD sections keep one such running value per vector lane, so
// their answer on NaN input depends on the lane layout; the scalar one is
// the order-defined answer.) Without NaN it is the max over all terms.
#include "../../include/binnedBLAS.h"
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace rb;

namespace {

// |t| as an unsigned key: for non-negative values the bit patterns order like
// the values, so an atomic max of keys is an exact, order-free max.
// key[0]: max |t| over the non-NaN terms from i0 on; key[1]: 1 + the index
// of the last NaN term (0: none).
template <class F>
__global__ __launch_bounds__(256) void amax_kernel(int64_t i0, int64_t n, const F* x, int64_t incx, const F* y,
                                                   int64_t incy, unsigned long long* key) {
  F m = F(0);
  unsigned long long last = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const F t = fabs(y ? x[i * incx] * y[i * incy] : x[i * incx]);
    if (t != t) {
      last = (unsigned long long)i + 1;
    } else {
      m = fmax(m, t);
    }
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) {
    m = fmax(m, __shfl_xor(m, s, 64));
    const unsigned long long o = __shfl_xor(last, s, 64);
    last = o > last ? o : last;
  }
  if ((threadIdx.x & 63) == 0) {
    if (m > F(0)) atomicMax(key, (unsigned long long)bits(m));
    if (last) atomicMax(key + 1, last);
  }
}

// |t_j| into *out (the result when the last NaN term is the last term).
template <class F>
__global__ void term_kernel(int64_t j, const F* x, int64_t incx, const F* y, int64_t incy, F* out) {
  *out = fabs(y ? x[j * incx] * y[j * incy] : x[j * incx]);
}

__global__ void clear_key(unsigned long long* key) {
  key[0] = 0ull;
  key[1] = 0ull;
}

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
  const DeviceInfo* d = device_info();
  if (!d) return F(0);
  // pass over [i0, N): the max of the non-NaN terms and the last NaN term
  auto pass = [&](int64_t i0, unsigned long long* h) {
    clear_key<<<1, 1, 0, s>>>(key);
    int64_t blocks = (N - i0 + 255) / 256;
    if (blocks > d->cus * 4) blocks = d->cus * 4;
    amax_kernel<F><<<(unsigned)blocks, 256, 0, s>>>(i0, N, dx, incX, dy, incY, key);
    if (!RB_HIP(hipGetLastError())) return false;
    if (!RB_HIP(hipMemcpyAsync(h, key, 2 * sizeof(*h), hipMemcpyDeviceToHost, s))) return false;
    return RB_HIP(hipStreamSynchronize(s));
  };
  unsigned long long h[2] = {0, 0};
  if (!pass(0, h)) return F(0);
  if (h[1] == 0) return fp_from_bits<F>((typename Fmt<F>::U)h[0]);
  // a NaN term: the result is the max after the last one, or that NaN
  const int64_t j = (int64_t)h[1] - 1;
  if (j == (int64_t)N - 1) {
    F* t = (F*)(key + 2);
    term_kernel<F><<<1, 1, 0, s>>>(j, dx, incX, dy, incY, t);
    F v = F(0);
    if (!RB_HIP(hipGetLastError())) return F(0);
    if (!RB_HIP(hipMemcpyAsync(&v, t, sizeof(F), hipMemcpyDeviceToHost, s))) return F(0);
    if (!RB_HIP(hipStreamSynchronize(s))) return F(0);
    return v;
  }
  if (!pass(j + 1, h)) return F(0);
  return fp_from_bits<F>((typename Fmt<F>::U)h[0]);
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
