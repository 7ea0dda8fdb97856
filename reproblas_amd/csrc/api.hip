// api.hip -- the C ABI: reproBLAS_* (include/reproBLAS.h), binnedBLAS_*
// (include/binnedBLAS.h) and the HIP additions (include/reproblas_hip.h).
//
// Every bulk reduction runs on the GPU. Host operands are staged into the
// library's per-stream device workspace and results are copied back; device
// operands are used in place. The synchronous calls run on the library
// stream (reproblas_hip_set_stream) and wait for it before returning.
#include <math.h>
#include <string.h>

#include "../../include/binnedBLAS.h"
#include "../../include/reproBLAS.h"
#include "../../include/reproblas_hip.h"
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace rb;

namespace {

constexpr double kNaN = __builtin_nan("");

template <class F = double>
bool check_fold(int fold) {
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
bool blas1_of(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
              const double* sin, double scale_in, const double* scale_in_dev, double* sout, double* scale_out,
              double* result, hipStream_t s) {
  return blas1(fold, term, n, x, incx, y, incy, sin, scale_in, scale_in_dev, sout, scale_out, result, s);
}
bool blas1_of(int fold, Term term, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy,
              const float* sin, float scale_in, const float* scale_in_dev, float* sout, float* scale_out,
              float* result, hipStream_t s) {
  return blas1_s(fold, term, n, x, incx, y, incy, sin, scale_in, scale_in_dev, sout, scale_out, result, s);
}

// Pinned host landing buffer for scalar results of the synchronous calls.
double* pinned_scalars() {
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
int64_t matrix_span(char order, int64_t rows, int64_t cols, int64_t ld) {
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

}  // namespace

extern "C" {

// ------------------------------------------------------------ errors etc. --
int reproblas_hip_last_error(void) { return last_error(); }
const char* reproblas_hip_last_error_string(void) { return last_error_message(); }
void reproblas_hip_clear_error(void) { clear_error(); }
void reproblas_hip_set_stream(void* stream) { set_current_stream((hipStream_t)stream); }
void* reproblas_hip_get_stream(void) { return (void*)current_stream(); }
int reproblas_hip_min_fold(void) { return 2; }
int reproblas_hip_max_fold(void) { return kMaxFold; }

// ------------------------------------------------------------------ BLAS1 --
// reference src/reproBLAS/rdsum.c:20-31
double reproBLAS_rdsum(const int fold, const int N, const double* X, const int incX) {
  return run_blas1<double>(fold, kSum, N, X, incX, nullptr, 0, nullptr, nullptr);
}
// reference src/reproBLAS/rdasum.c:20-31
double reproBLAS_rdasum(const int fold, const int N, const double* X, const int incX) {
  return run_blas1<double>(fold, kASum, N, X, incX, nullptr, 0, nullptr, nullptr);
}
// reference src/reproBLAS/rdnrm2.c:22-34 (scl * sqrt(ddbconv(ssq)))
double reproBLAS_rdnrm2(const int fold, const int N, const double* X, const int incX) {
  return run_blas1<double>(fold, kSSq, N, X, incX, nullptr, 0, nullptr, nullptr);
}
// reference src/reproBLAS/rddot.c:22-33
double reproBLAS_rddot(const int fold, const int N, const double* X, const int incX, const double* Y,
                       const int incY) {
  return run_blas1<double>(fold, kDot, N, X, incX, Y, incY, nullptr, nullptr);
}
// reference src/reproBLAS/dsum.c:20-22 and siblings (fold DIDEFAULTFOLD)
double reproBLAS_dsum(const int N, const double* X, const int incX) { return reproBLAS_rdsum(DIDEFAULTFOLD, N, X, incX); }
double reproBLAS_dasum(const int N, const double* X, const int incX) {
  return reproBLAS_rdasum(DIDEFAULTFOLD, N, X, incX);
}
double reproBLAS_dnrm2(const int N, const double* X, const int incX) {
  return reproBLAS_rdnrm2(DIDEFAULTFOLD, N, X, incX);
}
double reproBLAS_ddot(const int N, const double* X, const int incX, const double* Y, const int incY) {
  return reproBLAS_rddot(DIDEFAULTFOLD, N, X, incX, Y, incY);
}

// ------------------------------------------------------ binned-state BLAS1 --
// reference src/binnedBLAS/dbdsum.c:17-19
void binnedBLAS_dbdsum(const int fold, const int N, const double* X, const int incX, double_binned* Y) {
  run_blas1<double>(fold, kSum, N, X, incX, nullptr, 0, Y, nullptr);
}
void binnedBLAS_dmdsum(const int fold, const int N, const double* X, const int incX, double* priY, const int incpriY,
                       double* carY, const int inccarY) {
  if (packed_host(priY, incpriY, carY, inccarY, fold)) return binnedBLAS_dbdsum(fold, N, X, incX, priY);
  with_packed(fold, priY, incpriY, carY, inccarY, [&](double* st) { binnedBLAS_dbdsum(fold, N, X, incX, st); });
}
// reference src/binnedBLAS/dbdasum.c:16-18
void binnedBLAS_dbdasum(const int fold, const int N, const double* X, const int incX, double_binned* Y) {
  run_blas1<double>(fold, kASum, N, X, incX, nullptr, 0, Y, nullptr);
}
void binnedBLAS_dmdasum(const int fold, const int N, const double* X, const int incX, double* priY,
                        const int incpriY, double* carY, const int inccarY) {
  if (packed_host(priY, incpriY, carY, inccarY, fold)) return binnedBLAS_dbdasum(fold, N, X, incX, priY);
  with_packed(fold, priY, incpriY, carY, inccarY, [&](double* st) { binnedBLAS_dbdasum(fold, N, X, incX, st); });
}
// reference src/binnedBLAS/dbddot.c:19-21
void binnedBLAS_dbddot(const int fold, const int N, const double* X, const int incX, const double* Y,
                       const int incY, double_binned* Z) {
  run_blas1<double>(fold, kDot, N, X, incX, Y, incY, Z, nullptr);
}
void binnedBLAS_dmddot(const int fold, const int N, const double* X, const int incX, const double* Y,
                       const int incY, double* priZ, const int incpriZ, double* carZ, const int inccarZ) {
  if (packed_host(priZ, incpriZ, carZ, inccarZ, fold)) return binnedBLAS_dbddot(fold, N, X, incX, Y, incY, priZ);
  with_packed(fold, priZ, incpriZ, carZ, inccarZ,
              [&](double* st) { binnedBLAS_dbddot(fold, N, X, incX, Y, incY, st); });
}
// reference src/binnedBLAS/dbdssq.c:19-21 (returns the new scale)
double binnedBLAS_dbdssq(const int fold, const int N, const double* X, const int incX, const double scaleY,
                         double_binned* Y) {
  double scale = scaleY;
  run_blas1<double>(fold, kSSq, N, X, incX, nullptr, 0, Y, &scale);
  return scale;
}
double binnedBLAS_dmdssq(const int fold, const int N, const double* X, const int incX, const double scaleY,
                         double* priY, const int incpriY, double* carY, const int inccarY) {
  if (packed_host(priY, incpriY, carY, inccarY, fold)) return binnedBLAS_dbdssq(fold, N, X, incX, scaleY, priY);
  double r = 0.0;
  with_packed(fold, priY, incpriY, carY, inccarY,
              [&](double* st) { r = binnedBLAS_dbdssq(fold, N, X, incX, scaleY, st); });
  return r;
}

// ---------------------------------------------------- device-resident BLAS1 --
int reproblas_hip_rdsum_device(int fold, int64_t N, const double* X, int64_t incX, double* result, void* stream) {
  return device_blas1<double>(fold, kSum, N, X, incX, nullptr, 0, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_rdasum_device(int fold, int64_t N, const double* X, int64_t incX, double* result, void* stream) {
  return device_blas1<double>(fold, kASum, N, X, incX, nullptr, 0, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_rdnrm2_device(int fold, int64_t N, const double* X, int64_t incX, double* result, void* stream) {
  return device_blas1<double>(fold, kSSq, N, X, incX, nullptr, 0, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_rddot_device(int fold, int64_t N, const double* X, int64_t incX, const double* Y, int64_t incY,
                               double* result, void* stream) {
  return device_blas1<double>(fold, kDot, N, X, incX, Y, incY, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_dbdsum_device(int fold, int64_t N, const double* X, int64_t incX, double_binned* Y, void* stream) {
  return device_blas1<double>(fold, kSum, N, X, incX, nullptr, 0, Y, nullptr, Y, nullptr, nullptr, stream);
}
int reproblas_hip_dbdasum_device(int fold, int64_t N, const double* X, int64_t incX, double_binned* Y,
                                 void* stream) {
  return device_blas1<double>(fold, kASum, N, X, incX, nullptr, 0, Y, nullptr, Y, nullptr, nullptr, stream);
}
int reproblas_hip_dbddot_device(int fold, int64_t N, const double* X, int64_t incX, const double* Y, int64_t incY,
                                double_binned* Z, void* stream) {
  return device_blas1<double>(fold, kDot, N, X, incX, Y, incY, Z, nullptr, Z, nullptr, nullptr, stream);
}
int reproblas_hip_dbdssq_device(int fold, int64_t N, const double* X, int64_t incX, double* scale, double_binned* Y,
                                void* stream) {
  return device_blas1<double>(fold, kSSq, N, X, incX, nullptr, 0, Y, scale, Y, scale, nullptr, stream);
}

// -------------------------------------------------------------- gemv/gemm --
// reference src/reproBLAS/rdgemv.c:35-87
void reproBLAS_rdgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                      const double alpha, const double* A, const int lda, const double* X, const int incX,
                      const double beta, double* Y, const int incY) {
  clear_error();
  if (M <= 0 || N <= 0) return;
  if (!check_fold(fold)) return;
  hipStream_t s = current_stream();
  const bool nt = (TransA == 'n' || TransA == 'N');
  const int64_t lenx = nt ? N : M, leny = nt ? M : N;
  bool sa, sx, sy;
  const double* dA = stage_matrix(A, matrix_span(Order, M, N, lda), kSlotStageX, s, &sa);
  const double* dX = stage_vector(X, lenx, incX, kSlotStageY, s, &sx);
  double* dY = (double*)stage_vector(Y, leny, incY, kSlotStageZ, s, &sy);
  if (!dA || !dX || !dY) return;
  if (!rdgemv_device(fold, Order, TransA, M, N, alpha, dA, lda, dX, incX, beta, dY, incY, s)) return;
  if (sy && !unstage_vector(Y, dY, leny, incY, s)) return;
  RB_HIP(hipStreamSynchronize(s));
}
void reproBLAS_dgemv(const char Order, const char TransA, const int M, const int N, const double alpha,
                     const double* A, const int lda, const double* X, const int incX, const double beta, double* Y,
                     const int incY) {
  reproBLAS_rdgemv(DIDEFAULTFOLD, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY);
}

// reference src/reproBLAS/rdgemm.c:41-105
void reproBLAS_rdgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const double alpha, const double* A, const int lda, const double* B,
                      const int ldb, const double beta, double* C, const int ldc) {
  clear_error();
  if (M <= 0 || N <= 0) return;
  if (!check_fold(fold)) return;
  hipStream_t s = current_stream();
  const bool na = (TransA == 'n' || TransA == 'N'), nb = (TransB == 'n' || TransB == 'N');
  bool sa, sb, sc;
  const double* dA = stage_matrix(A, matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa);
  const double* dB = stage_matrix(B, matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb);
  const int64_t cspan = matrix_span(Order, M, N, ldc);
  double* dC = (double*)stage_matrix(C, cspan, kSlotStageZ, s, &sc);
  if ((K > 0 && (!dA || !dB)) || !dC) return;
  if (!rdgemm_device(fold, Order, TransA, TransB, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, s)) return;
  if (sc && !RB_HIP(hipMemcpyAsync(C, dC, sizeof(double) * (size_t)cspan, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}
void reproBLAS_dgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const double alpha, const double* A, const int lda, const double* B, const int ldb,
                     const double beta, double* C, const int ldc) {
  reproBLAS_rdgemm(DIDEFAULTFOLD, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// reference src/binnedBLAS/dbdgemv.ccog:57-162 (Y: binned vector, stride incY)
void binnedBLAS_dbdgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const double alpha, const double* A, const int lda, const double* X, const int incX,
                        double_binned* Y, const int incY) {
  clear_error();
  if (M <= 0 || N <= 0 || alpha == 0.0) return;
  if (!check_fold(fold)) return;
  hipStream_t s = current_stream();
  const bool nt = (TransA == 'n' || TransA == 'N');
  const int64_t lenx = nt ? N : M, leny = nt ? M : N;
  bool sa, sx;
  const double* dA = stage_matrix(A, matrix_span(Order, M, N, lda), kSlotStageX, s, &sa);
  const double* dX = stage_vector(X, lenx, incX, kSlotStageY, s, &sx);
  if (!dA || !dX) return;
  const int64_t ybytes = sizeof(double) * (size_t)(((leny - 1) * (incY < 0 ? -incY : incY) + 1) * 2 * fold);
  const bool ydev = is_device_pointer(Y);
  double* dY = Y;
  if (!ydev) {
    if (incY < 0) {
      set_error(kUnsupported, "binnedBLAS_dbdgemv: negative incY on a host binned vector");
      return;
    }
    dY = (double*)workspace(s, kSlotBinnedOut, ybytes);
    if (!dY || !RB_HIP(hipMemcpyAsync(dY, Y, ybytes, hipMemcpyHostToDevice, s))) return;
  }
  if (!dbdgemv(fold, Order, TransA, M, N, alpha, dA, lda, dX, incX, dY, incY, s)) return;
  if (!ydev && !RB_HIP(hipMemcpyAsync(Y, dY, ybytes, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}

// reference src/binnedBLAS/dbdgemm.ccog:61-420 (C: binned matrix, leading dimension ldc)
void binnedBLAS_dbdgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const double alpha, const double* A, const int lda,
                        const double* B, const int ldb, double_binned* C, const int ldc) {
  clear_error();
  if (M <= 0 || N <= 0 || K <= 0 || alpha == 0.0) return;
  if (!check_fold(fold)) return;
  hipStream_t s = current_stream();
  const bool na = (TransA == 'n' || TransA == 'N'), nb = (TransB == 'n' || TransB == 'N');
  bool sa, sb;
  const double* dA = stage_matrix(A, matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa);
  const double* dB = stage_matrix(B, matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb);
  if (!dA || !dB) return;
  const int64_t cbytes = sizeof(double) * (size_t)matrix_span(Order, M, N, ldc) * 2 * fold;
  const bool cdev = is_device_pointer(C);
  double* dC = C;
  if (!cdev) {
    dC = (double*)workspace(s, kSlotBinnedOut, cbytes);
    if (!dC || !RB_HIP(hipMemcpyAsync(dC, C, cbytes, hipMemcpyHostToDevice, s))) return;
  }
  if (!dbdgemm(fold, Order, TransA, TransB, M, N, K, alpha, dA, lda, dB, ldb, dC, ldc, s)) return;
  if (!cdev && !RB_HIP(hipMemcpyAsync(C, dC, cbytes, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}

int reproblas_hip_rdgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, double alpha,
                                const double* A, int64_t lda, const double* X, int64_t incX, double beta, double* Y,
                                int64_t incY, void* stream) {
  clear_error();
  if (M <= 0 || N <= 0) return kOk;
  if (!check_fold(fold)) return last_error();
  if (!require_device_accessible(Y, "Y") ||
      (alpha != 0.0 && (!require_device_accessible(A, "A") || !require_device_accessible(X, "X"))))
    return last_error();
  if (!rdgemv_device(fold, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY, (hipStream_t)stream))
    return last_error();
  return kOk;
}

int reproblas_hip_rdgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                                double beta, double* C, int64_t ldc, void* stream) {
  clear_error();
  if (M <= 0 || N <= 0) return kOk;
  if (!check_fold(fold)) return last_error();
  if (!require_device_accessible(C, "C") ||
      (alpha != 0.0 && K > 0 && (!require_device_accessible(A, "A") || !require_device_accessible(B, "B"))))
    return last_error();
  if (!rdgemm_device(fold, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                     (hipStream_t)stream))
    return last_error();
  return kOk;
}

int reproblas_hip_gemm_tiling_count(void) { return gemm_tiling_count(); }
int reproblas_hip_gemm_set_tiling(int tiling) {
  clear_error();
  return gemm_set_tiling(tiling) ? kOk : last_error();
}
int reproblas_hip_gemm_get_tiling(void) { return gemm_get_tiling(); }


// ============================================================ single precision
// reference src/reproBLAS/rssum.c, rsasum.c, rsnrm2.c, rsdot.c
float reproBLAS_rssum(const int fold, const int N, const float* X, const int incX) {
  return run_blas1<float>(fold, kSum, N, X, incX, nullptr, 0, nullptr, nullptr);
}
float reproBLAS_rsasum(const int fold, const int N, const float* X, const int incX) {
  return run_blas1<float>(fold, kASum, N, X, incX, nullptr, 0, nullptr, nullptr);
}
float reproBLAS_rsnrm2(const int fold, const int N, const float* X, const int incX) {
  return run_blas1<float>(fold, kSSq, N, X, incX, nullptr, 0, nullptr, nullptr);
}
float reproBLAS_rsdot(const int fold, const int N, const float* X, const int incX, const float* Y, const int incY) {
  return run_blas1<float>(fold, kDot, N, X, incX, Y, incY, nullptr, nullptr);
}
// reference src/reproBLAS/ssum.c and siblings (fold SIDEFAULTFOLD)
float reproBLAS_ssum(const int N, const float* X, const int incX) { return reproBLAS_rssum(SIDEFAULTFOLD, N, X, incX); }
float reproBLAS_sasum(const int N, const float* X, const int incX) {
  return reproBLAS_rsasum(SIDEFAULTFOLD, N, X, incX);
}
float reproBLAS_snrm2(const int N, const float* X, const int incX) {
  return reproBLAS_rsnrm2(SIDEFAULTFOLD, N, X, incX);
}
float reproBLAS_sdot(const int N, const float* X, const int incX, const float* Y, const int incY) {
  return reproBLAS_rsdot(SIDEFAULTFOLD, N, X, incX, Y, incY);
}

// reference src/binnedBLAS/sbssum.c, sbsasum.c, sbsdot.c, sbsssq.c
void binnedBLAS_sbssum(const int fold, const int N, const float* X, const int incX, float_binned* Y) {
  run_blas1<float>(fold, kSum, N, X, incX, nullptr, 0, Y, nullptr);
}
void binnedBLAS_smssum(const int fold, const int N, const float* X, const int incX, float* priY, const int incpriY,
                       float* carY, const int inccarY) {
  if (packed_host(priY, incpriY, carY, inccarY, fold)) return binnedBLAS_sbssum(fold, N, X, incX, priY);
  with_packed(fold, priY, incpriY, carY, inccarY, [&](float* st) { binnedBLAS_sbssum(fold, N, X, incX, st); });
}
void binnedBLAS_sbsasum(const int fold, const int N, const float* X, const int incX, float_binned* Y) {
  run_blas1<float>(fold, kASum, N, X, incX, nullptr, 0, Y, nullptr);
}
void binnedBLAS_smsasum(const int fold, const int N, const float* X, const int incX, float* priY, const int incpriY,
                        float* carY, const int inccarY) {
  if (packed_host(priY, incpriY, carY, inccarY, fold)) return binnedBLAS_sbsasum(fold, N, X, incX, priY);
  with_packed(fold, priY, incpriY, carY, inccarY, [&](float* st) { binnedBLAS_sbsasum(fold, N, X, incX, st); });
}
void binnedBLAS_sbsdot(const int fold, const int N, const float* X, const int incX, const float* Y, const int incY,
                       float_binned* Z) {
  run_blas1<float>(fold, kDot, N, X, incX, Y, incY, Z, nullptr);
}
void binnedBLAS_smsdot(const int fold, const int N, const float* X, const int incX, const float* Y, const int incY,
                       float* priZ, const int incpriZ, float* carZ, const int inccarZ) {
  if (packed_host(priZ, incpriZ, carZ, inccarZ, fold)) return binnedBLAS_sbsdot(fold, N, X, incX, Y, incY, priZ);
  with_packed(fold, priZ, incpriZ, carZ, inccarZ,
              [&](float* st) { binnedBLAS_sbsdot(fold, N, X, incX, Y, incY, st); });
}
float binnedBLAS_sbsssq(const int fold, const int N, const float* X, const int incX, const float scaleY,
                        float_binned* Y) {
  float scale = scaleY;
  run_blas1<float>(fold, kSSq, N, X, incX, nullptr, 0, Y, &scale);
  return scale;
}
float binnedBLAS_smsssq(const int fold, const int N, const float* X, const int incX, const float scaleY, float* priY,
                        const int incpriY, float* carY, const int inccarY) {
  if (packed_host(priY, incpriY, carY, inccarY, fold)) return binnedBLAS_sbsssq(fold, N, X, incX, scaleY, priY);
  float r = 0.0f;
  with_packed(fold, priY, incpriY, carY, inccarY,
              [&](float* st) { r = binnedBLAS_sbsssq(fold, N, X, incX, scaleY, st); });
  return r;
}

int reproblas_hip_rssum_device(int fold, int64_t N, const float* X, int64_t incX, float* result, void* stream) {
  return device_blas1<float>(fold, kSum, N, X, incX, nullptr, 0, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_rsasum_device(int fold, int64_t N, const float* X, int64_t incX, float* result, void* stream) {
  return device_blas1<float>(fold, kASum, N, X, incX, nullptr, 0, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_rsnrm2_device(int fold, int64_t N, const float* X, int64_t incX, float* result, void* stream) {
  return device_blas1<float>(fold, kSSq, N, X, incX, nullptr, 0, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_rsdot_device(int fold, int64_t N, const float* X, int64_t incX, const float* Y, int64_t incY,
                               float* result, void* stream) {
  return device_blas1<float>(fold, kDot, N, X, incX, Y, incY, nullptr, nullptr, nullptr, nullptr, result, stream);
}
int reproblas_hip_sbssum_device(int fold, int64_t N, const float* X, int64_t incX, float_binned* Y, void* stream) {
  return device_blas1<float>(fold, kSum, N, X, incX, nullptr, 0, Y, nullptr, Y, nullptr, nullptr, stream);
}
int reproblas_hip_sbsasum_device(int fold, int64_t N, const float* X, int64_t incX, float_binned* Y, void* stream) {
  return device_blas1<float>(fold, kASum, N, X, incX, nullptr, 0, Y, nullptr, Y, nullptr, nullptr, stream);
}
int reproblas_hip_sbsdot_device(int fold, int64_t N, const float* X, int64_t incX, const float* Y, int64_t incY,
                                float_binned* Z, void* stream) {
  return device_blas1<float>(fold, kDot, N, X, incX, Y, incY, Z, nullptr, Z, nullptr, nullptr, stream);
}
int reproblas_hip_sbsssq_device(int fold, int64_t N, const float* X, int64_t incX, float* scale, float_binned* Y,
                                void* stream) {
  return device_blas1<float>(fold, kSSq, N, X, incX, nullptr, 0, Y, scale, Y, scale, nullptr, stream);
}

// reference src/reproBLAS/rsgemv.c
void reproBLAS_rsgemv(const int fold, const char Order, const char TransA, const int M, const int N, const float alpha,
                      const float* A, const int lda, const float* X, const int incX, const float beta, float* Y,
                      const int incY) {
  clear_error();
  if (M <= 0 || N <= 0) return;
  if (!check_fold<float>(fold)) return;
  hipStream_t s = current_stream();
  const bool nt = (TransA == 'n' || TransA == 'N');
  const int64_t lenx = nt ? N : M, leny = nt ? M : N;
  bool sa, sx, sy;
  const float* dA = stage_matrix(A, matrix_span(Order, M, N, lda), kSlotStageX, s, &sa);
  const float* dX = stage_vector(X, lenx, incX, kSlotStageY, s, &sx);
  float* dY = (float*)stage_vector(Y, leny, incY, kSlotStageZ, s, &sy);
  if (!dA || !dX || !dY) return;
  if (!rsgemv_device(fold, Order, TransA, M, N, alpha, dA, lda, dX, incX, beta, dY, incY, s)) return;
  if (sy && !unstage_vector(Y, dY, leny, incY, s)) return;
  RB_HIP(hipStreamSynchronize(s));
}
void reproBLAS_sgemv(const char Order, const char TransA, const int M, const int N, const float alpha, const float* A,
                     const int lda, const float* X, const int incX, const float beta, float* Y, const int incY) {
  reproBLAS_rsgemv(SIDEFAULTFOLD, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY);
}

// reference src/reproBLAS/rsgemm.c
void reproBLAS_rsgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const float alpha, const float* A, const int lda, const float* B,
                      const int ldb, const float beta, float* C, const int ldc) {
  clear_error();
  if (M <= 0 || N <= 0) return;
  if (!check_fold<float>(fold)) return;
  hipStream_t s = current_stream();
  const bool na = (TransA == 'n' || TransA == 'N'), nb = (TransB == 'n' || TransB == 'N');
  bool sa, sb, sc;
  const float* dA = stage_matrix(A, matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa);
  const float* dB = stage_matrix(B, matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb);
  const int64_t cspan = matrix_span(Order, M, N, ldc);
  float* dC = (float*)stage_matrix(C, cspan, kSlotStageZ, s, &sc);
  if ((K > 0 && (!dA || !dB)) || !dC) return;
  if (!rsgemm_device(fold, Order, TransA, TransB, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, s)) return;
  if (sc && !RB_HIP(hipMemcpyAsync(C, dC, sizeof(float) * (size_t)cspan, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}
void reproBLAS_sgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const float alpha, const float* A, const int lda, const float* B, const int ldb, const float beta,
                     float* C, const int ldc) {
  reproBLAS_rsgemm(SIDEFAULTFOLD, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// reference src/binnedBLAS/sbsgemv.ccog (Y: binned vector, stride incY)
void binnedBLAS_sbsgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const float alpha, const float* A, const int lda, const float* X, const int incX,
                        float_binned* Y, const int incY) {
  clear_error();
  if (M <= 0 || N <= 0 || alpha == 0.0f) return;
  if (!check_fold<float>(fold)) return;
  hipStream_t s = current_stream();
  const bool nt = (TransA == 'n' || TransA == 'N');
  const int64_t lenx = nt ? N : M, leny = nt ? M : N;
  bool sa, sx;
  const float* dA = stage_matrix(A, matrix_span(Order, M, N, lda), kSlotStageX, s, &sa);
  const float* dX = stage_vector(X, lenx, incX, kSlotStageY, s, &sx);
  if (!dA || !dX) return;
  const int64_t ybytes = sizeof(float) * (size_t)(((leny - 1) * (incY < 0 ? -incY : incY) + 1) * 2 * fold);
  const bool ydev = is_device_pointer(Y);
  float* dY = Y;
  if (!ydev) {
    if (incY < 0) {
      set_error(kUnsupported, "binnedBLAS_sbsgemv: negative incY on a host binned vector");
      return;
    }
    dY = (float*)workspace(s, kSlotBinnedOut, ybytes);
    if (!dY || !RB_HIP(hipMemcpyAsync(dY, Y, ybytes, hipMemcpyHostToDevice, s))) return;
  }
  if (!sbsgemv(fold, Order, TransA, M, N, alpha, dA, lda, dX, incX, dY, incY, s)) return;
  if (!ydev && !RB_HIP(hipMemcpyAsync(Y, dY, ybytes, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}

// reference src/binnedBLAS/sbsgemm.ccog (C: binned matrix, leading dimension ldc)
void binnedBLAS_sbsgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const float alpha, const float* A, const int lda, const float* B,
                        const int ldb, float_binned* C, const int ldc) {
  clear_error();
  if (M <= 0 || N <= 0 || K <= 0 || alpha == 0.0f) return;
  if (!check_fold<float>(fold)) return;
  hipStream_t s = current_stream();
  const bool na = (TransA == 'n' || TransA == 'N'), nb = (TransB == 'n' || TransB == 'N');
  bool sa, sb;
  const float* dA = stage_matrix(A, matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa);
  const float* dB = stage_matrix(B, matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb);
  if (!dA || !dB) return;
  const int64_t cbytes = sizeof(float) * (size_t)matrix_span(Order, M, N, ldc) * 2 * fold;
  const bool cdev = is_device_pointer(C);
  float* dC = C;
  if (!cdev) {
    dC = (float*)workspace(s, kSlotBinnedOut, cbytes);
    if (!dC || !RB_HIP(hipMemcpyAsync(dC, C, cbytes, hipMemcpyHostToDevice, s))) return;
  }
  if (!sbsgemm(fold, Order, TransA, TransB, M, N, K, alpha, dA, lda, dB, ldb, dC, ldc, s)) return;
  if (!cdev && !RB_HIP(hipMemcpyAsync(C, dC, cbytes, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}

int reproblas_hip_rsgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, float alpha, const float* A,
                                int64_t lda, const float* X, int64_t incX, float beta, float* Y, int64_t incY,
                                void* stream) {
  clear_error();
  if (M <= 0 || N <= 0) return kOk;
  if (!check_fold<float>(fold)) return last_error();
  if (!require_device_accessible(Y, "Y") ||
      (alpha != 0.0f && (!require_device_accessible(A, "A") || !require_device_accessible(X, "X"))))
    return last_error();
  if (!rsgemv_device(fold, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY, (hipStream_t)stream))
    return last_error();
  return kOk;
}

int reproblas_hip_rsgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                                float* C, int64_t ldc, void* stream) {
  clear_error();
  if (M <= 0 || N <= 0) return kOk;
  if (!check_fold<float>(fold)) return last_error();
  if (!require_device_accessible(C, "C") ||
      (alpha != 0.0f && K > 0 && (!require_device_accessible(A, "A") || !require_device_accessible(B, "B"))))
    return last_error();
  if (!rsgemm_device(fold, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, (hipStream_t)stream))
    return last_error();
  return kOk;
}

}  // extern "C"
