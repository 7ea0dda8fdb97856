// api_helpers.hpp -- host-side helpers shared by the C ABI files (api.hip,
// api_complex.hip): fold checks, staging of host operands into the
// per-stream workspace, the synchronous BLAS1 driver and gathering of
// manually strided binned values.
#pragma once
#include <math.h>
#include <string.h>

#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rb {
namespace api {

constexpr double kNaN = __builtin_nan("");

template <class F = double>
inline bool check_fold(int fold) {
  if (fold >= 2 && fold <= Fmt<F>::max_fold) return true;
  set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
  return false;
}

// Device view of a strided vector of n elements. Host data is copied into the
// workspace slot; the returned pointer p satisfies p[i*inc] = X[i*inc].
template <class F>
const F* stage_vector(const F* X, int64_t n, int64_t inc, int slot, hipStream_t s, bool* staged) {
  *staged = false;
  if (n <= 0) return X;
  if (is_device_pointer(X)) return X;
  const int64_t ainc = inc < 0 ? -inc : (inc == 0 ? 0 : inc);
  const int64_t span = (n - 1) * ainc + 1;
  F* d = (F*)workspace(s, slot, sizeof(F) * (size_t)span);
  if (!d) return nullptr;
  const F* lo = (inc < 0) ? X + (n - 1) * inc : X;
  if (!RB_HIP(hipMemcpyAsync(d, lo, sizeof(F) * (size_t)span, hipMemcpyHostToDevice, s))) return nullptr;
  *staged = true;
  return (inc < 0) ? d + (n - 1) * ainc : d;
}

// Copies a staged strided vector back to the host span it came from.
template <class F>
bool unstage_vector(F* X, const F* d, int64_t n, int64_t inc, hipStream_t s) {
  const int64_t ainc = inc < 0 ? -inc : (inc == 0 ? 0 : inc);
  const int64_t span = (n - 1) * ainc + 1;
  F* lo = (inc < 0) ? X + (n - 1) * inc : X;
  const F* dlo = (inc < 0) ? d - (n - 1) * ainc : d;
  return RB_HIP(hipMemcpyAsync(lo, dlo, sizeof(F) * (size_t)span, hipMemcpyDeviceToHost, s));
}

// The BLAS1 launcher of each element type.
inline bool blas1_of(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
              const double* sin, double scale_in, const double* scale_in_dev, double* sout, double* scale_out,
              double* result, hipStream_t s) {
  return blas1(fold, term, n, x, incx, y, incy, sin, scale_in, scale_in_dev, sout, scale_out, result, s);
}
inline bool blas1_of(int fold, Term term, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy,
              const float* sin, float scale_in, const float* scale_in_dev, float* sout, float* scale_out,
              float* result, hipStream_t s) {
  return blas1_s(fold, term, n, x, incx, y, incy, sin, scale_in, scale_in_dev, sout, scale_out, result, s);
}

// Pinned host landing buffer for scalar results of the synchronous calls.
inline double* pinned_scalars() {
  static thread_local double* p = nullptr;
  if (!p) {
    if (!RB_HIP(hipHostMalloc((void**)&p, (16 + 4 * kMaxFold) * sizeof(double), hipHostMallocDefault))) p = nullptr;
  }
  return p;
}

// Runs one BLAS1 reduction synchronously. Y (2*fold, optional) and scale
// (optional) are host or device; returns the converted result.
template <class F>
F run_blas1(int fold, Term term, int64_t n, const F* X, int64_t incX, const F* Yv, int64_t incY, F* state,
            F* scale) {
  clear_error();
  if (!check_fold<F>(fold)) return (F)kNaN;
  if (n < 0) n = 0;
  hipStream_t s = current_stream();
  RB_STREAM_GUARD(s);
  bool sx = false, sy = false;
  const F* dX = stage_vector(X, n, incX, kSlotStageX, s, &sx);
  const F* dY = (term == kDot) ? stage_vector(Yv, n, incY, kSlotStageY, s, &sy) : nullptr;
  if ((n > 0 && !dX) || (term == kDot && n > 0 && !dY)) return (F)kNaN;
  F* out = (F*)workspace(s, kSlotResult, sizeof(F) * (4 * Fmt<F>::max_fold + 8));
  F* host = (F*)pinned_scalars();
  if (!out || !host) return (F)kNaN;
  // out layout: [0] result, [1] scale in, [2] scale out, [8..8+2F) state in, [8+2F..8+4F) state out
  F* d_res = out;
  F* d_state_in = out + 8;
  F* d_state_out = out + 8 + 2 * fold;
  const bool state_dev = state && is_device_pointer(state);
  const F* sin = nullptr;
  F* sout = nullptr;
  if (state) {
    if (state_dev) {
      sin = state;
      sout = state;  // the kernel reads it before writing (last workgroup)
    } else {
      if (!RB_HIP(hipMemcpyAsync(d_state_in, state, sizeof(F) * 2 * fold, hipMemcpyHostToDevice, s))) return (F)kNaN;
      sin = d_state_in;
      sout = d_state_out;
    }
  }
  const F scale_in = scale ? *scale : F(0);
  if (!blas1_of(fold, term, n, dX, incX, dY, incY, sin, scale_in, nullptr, sout, out + 2, d_res, s)) return (F)kNaN;
  if (!RB_HIP(hipMemcpyAsync(host, out, sizeof(F) * 3, hipMemcpyDeviceToHost, s))) return (F)kNaN;
  if (state && !state_dev) {
    if (!RB_HIP(hipMemcpyAsync(host + 8, d_state_out, sizeof(F) * 2 * fold, hipMemcpyDeviceToHost, s)))
      return (F)kNaN;
  }
  if (!RB_HIP(hipStreamSynchronize(s))) return (F)kNaN;
  if (state && !state_dev) memcpy(state, host + 8, sizeof(F) * 2 * fold);
  if (scale) *scale = host[2];
  return host[0];
}

// binnedBLAS_dm* / sms* take primaries/carries with their own strides: gather
// them into a packed host value, run, scatter back.
template <class F, class Body>
void with_packed(int fold, F* pri, int incp, F* car, int incc, Body&& body) {
  F st[2 * Fmt<F>::max_fold];
  for (int i = 0; i < fold; ++i) {
    st[i] = pri[(long)i * incp];
    st[fold + i] = car[(long)i * incc];
  }
  body(st);
  for (int i = 0; i < fold; ++i) {
    pri[(long)i * incp] = st[i];
    car[(long)i * incc] = st[fold + i];
  }
}

template <class F>
bool packed_host(const F* pri, int incp, const F* car, int incc, int fold) {
  return pri + fold == car && incp == 1 && incc == 1;
}

// Device-resident BLAS1 of the *_device entry points.
template <class F>
int device_blas1(int fold, Term term, int64_t N, const F* X, int64_t incX, const F* Y, int64_t incY,
                        const F* sin, const F* scale_in_dev, F* sout, F* scale_out, F* result, void* stream) {
  clear_error();
  RB_STREAM_GUARD(stream);
  if (!check_fold<F>(fold)) return last_error();
  if (N < 0) N = 0;
  // the kernels dereference these directly: refuse host memory up front
  if (N > 0 && (!require_device_accessible(X, "X") || (term == kDot && !require_device_accessible(Y, "Y"))))
    return last_error();
  if (!require_device_accessible(sin, "binned state", true) ||
      !require_device_accessible(scale_in_dev, "scale", true) || !require_device_accessible(result, "result", true))
    return last_error();
  if (!blas1_of(fold, term, N, X, incX, Y, incY, sin, F(0), scale_in_dev, sout, scale_out, result,
                (hipStream_t)stream))
    return last_error();
  return kOk;
}
// Element span (in elements) of a rows x cols matrix with leading dimension ld
// in storage order `order`.
inline int64_t matrix_span(char order, int64_t rows, int64_t cols, int64_t ld) {
  if (rows <= 0 || cols <= 0) return 0;
  const bool rm = (order == 'r' || order == 'R');
  return rm ? (rows - 1) * ld + cols : (cols - 1) * ld + rows;
}

template <class F>
const F* stage_matrix(const F* A, int64_t span, int slot, hipStream_t s, bool* staged) {
  *staged = false;
  if (span <= 0 || is_device_pointer(A)) return A;
  F* d = (F*)workspace(s, slot, sizeof(F) * (size_t)span);
  if (!d) return nullptr;
  if (!RB_HIP(hipMemcpyAsync(d, A, sizeof(F) * (size_t)span, hipMemcpyHostToDevice, s))) return nullptr;
  *staged = true;
  return d;
}

}  // namespace api
}  // namespace rb
