// api_complex.hip -- the C ABI of the complex path: reproBLAS_rz* / rc* /
// rdz* / rsc* and their default-fold forms (include/reproBLAS.h),
// binnedBLAS_zbz* / cbc* / dbz* / sbc* (include/binnedBLAS.h) and the device
// entry points of include/reproblas_hip.h. Complex data are interleaved
// (real, imaginary) pairs; increments count complex elements.
#include <math.h>
#include <string.h>

#include "../../include/binnedBLAS.h"
#include "../../include/reproBLAS.h"
#include "../../include/reproblas_hip.h"
#include "api_helpers.hpp"
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace rb;
using namespace rb::api;

namespace {

// Device view of a strided complex vector: p[2*i*inc + c] = X[2*i*inc + c].
template <class F>
const F* stage_cvector(const F* X, int64_t n, int64_t inc, int slot, hipStream_t s, bool* staged) {
  *staged = false;
  if (n <= 0 || is_device_pointer(X)) return X;
  const int64_t ainc = inc < 0 ? -inc : inc;
  const int64_t span = 2 * ((n - 1) * ainc + 1);
  F* d = (F*)workspace(s, slot, sizeof(F) * (size_t)span);
  if (!d) return nullptr;
  const F* lo = (inc < 0) ? X + 2 * (n - 1) * inc : X;
  if (!RB_HIP(hipMemcpyAsync(d, lo, sizeof(F) * (size_t)span, hipMemcpyHostToDevice, s))) return nullptr;
  *staged = true;
  return (inc < 0) ? d + 2 * (n - 1) * ainc : d;
}

constexpr bool complex_result(ZTerm t) { return t == kZSum || t == kZDotU || t == kZDotC; }

// One complex reduction, synchronously. state: host or device binned value
// (4*fold for complex results, 2*fold for asum/ssq) or null; scale: host
// scale of an ssq state (in/out) or null; out: host results (2 or 1).
template <class F>
void run_zblas1(int fold, ZTerm term, int64_t n, const F* X, int64_t incX, const F* Y, int64_t incY, F* state,
                F* scale, F* out) {
  const int nres = complex_result(term) ? 2 : 1;
  for (int i = 0; i < nres; ++i) out[i] = (F)kNaN;
  clear_error();
  if (!check_fold<F>(fold)) return;
  if (n < 0) n = 0;
  hipStream_t s = current_stream();
  RB_STREAM_GUARD(s);
  const bool dot = (term == kZDotU || term == kZDotC);
  bool sx = false, sy = false;
  const F* dX = stage_cvector(X, n, incX, kSlotStageX, s, &sx);
  const F* dY = dot ? stage_cvector(Y, n, incY, kSlotStageY, s, &sy) : nullptr;
  if ((n > 0 && !dX) || (dot && n > 0 && !dY)) return;
  const int nst = complex_result(term) ? 4 * fold : 2 * fold;
  F* ws = (F*)workspace(s, kSlotResult, sizeof(F) * (8 + 2 * (size_t)nst));
  F* host = (F*)pinned_scalars();
  if (!ws || !host) return;
  // ws: [0..2) results, [2] scale out, [8..8+nst) state in, [8+nst..8+2nst) state out
  const bool state_dev = state && is_device_pointer(state);
  const F* sin = nullptr;
  F* sout = nullptr;
  if (state) {
    if (state_dev) {
      sin = state;
      sout = state;
    } else {
      if (!RB_HIP(hipMemcpyAsync(ws + 8, state, sizeof(F) * nst, hipMemcpyHostToDevice, s))) return;
      sin = ws + 8;
      sout = ws + 8 + nst;
    }
  }
  const F scale_in = scale ? *scale : F(0);
  if (!zblas1(fold, term, n, dX, incX, dY, incY, sin, scale_in, nullptr, sout, ws + 2, ws, s)) return;
  if (!RB_HIP(hipMemcpyAsync(host, ws, sizeof(F) * 3, hipMemcpyDeviceToHost, s))) return;
  if (state && !state_dev && !RB_HIP(hipMemcpyAsync(host + 8, ws + 8 + nst, sizeof(F) * nst, hipMemcpyDeviceToHost, s)))
    return;
  if (!RB_HIP(hipStreamSynchronize(s))) return;
  if (state && !state_dev) memcpy(state, host + 8, sizeof(F) * nst);
  if (scale) *scale = host[2];
  for (int i = 0; i < nres; ++i) out[i] = host[i];
}

template <class F>
F run_zblas1_real(int fold, ZTerm term, int64_t n, const F* X, int64_t incX, F* state, F* scale) {
  F out[2];
  run_zblas1<F>(fold, term, n, X, incX, nullptr, 0, state, scale, out);
  return out[0];
}

// binnedBLAS_zm* / cm*: complex primaries and carries with their own strides
// (complex units: element i of a component at pri[2*i*incpri + c]).
template <class F, class Body>
void with_packed_complex(int fold, F* pri, int incp, F* car, int incc, Body&& body) {
  F st[4 * Fmt<F>::max_fold];
  for (int i = 0; i < fold; ++i)
    for (int c = 0; c < 2; ++c) {
      st[2 * i + c] = pri[2L * i * incp + c];
      st[2 * fold + 2 * i + c] = car[2L * i * incc + c];
    }
  body(st);
  for (int i = 0; i < fold; ++i)
    for (int c = 0; c < 2; ++c) {
      pri[2L * i * incp + c] = st[2 * i + c];
      car[2L * i * incc + c] = st[2 * fold + 2 * i + c];
    }
}

template <class F>
int device_zblas1(int fold, ZTerm term, int64_t N, const F* X, int64_t incX, const F* Y, int64_t incY, F* result,
                  void* stream) {
  clear_error();
  RB_STREAM_GUARD(stream);
  if (!check_fold<F>(fold)) return last_error();
  if (N < 0) N = 0;
  const bool dot = (term == kZDotU || term == kZDotC);
  if (N > 0 && (!require_device_accessible(X, "X") || (dot && !require_device_accessible(Y, "Y"))))
    return last_error();
  if (!require_device_accessible(result, "result")) return last_error();
  if (!zblas1(fold, term, N, X, incX, Y, incY, nullptr, F(0), nullptr, nullptr, nullptr, result,
              (hipStream_t)stream))
    return last_error();
  return kOk;
}

// max |x_c| of one component (c = 0, 1) through the real kernels, and the
// per-component product maxima of zamaxm
template <class F>
F real_amax(int N, const F* X, int inc);
template <>
double real_amax<double>(int N, const double* X, int inc) { return binnedBLAS_damax(N, X, inc); }
template <>
float real_amax<float>(int N, const float* X, int inc) { return binnedBLAS_samax(N, X, inc); }
template <class F>
F real_amaxm(int N, const F* X, int incX, const F* Y, int incY);
template <>
double real_amaxm<double>(int N, const double* X, int incX, const double* Y, int incY) {
  return binnedBLAS_damaxm(N, X, incX, Y, incY);
}
template <>
float real_amaxm<float>(int N, const float* X, int incX, const float* Y, int incY) {
  return binnedBLAS_samaxm(N, X, incX, Y, incY);
}

template <class F>
void complex_amax(int N, const void* Xv, int incX, void* amax) {
  const F* X = (const F*)Xv;
  ((F*)amax)[0] = real_amax<F>(N, X, 2 * incX);
  ((F*)amax)[1] = real_amax<F>(N, X + 1, 2 * incX);
}
template <class F>
void complex_amaxm(int N, const void* Xv, int incX, const void* Yv, int incY, void* amaxm) {
  const F* X = (const F*)Xv;
  const F* Y = (const F*)Yv;
  const F rr = real_amaxm<F>(N, X, 2 * incX, Y, 2 * incY), ii = real_amaxm<F>(N, X + 1, 2 * incX, Y + 1, 2 * incY);
  const F ir = real_amaxm<F>(N, X + 1, 2 * incX, Y, 2 * incY), ri = real_amaxm<F>(N, X, 2 * incX, Y + 1, 2 * incY);
  ((F*)amaxm)[0] = fmax(rr, ii);
  ((F*)amaxm)[1] = fmax(ir, ri);
}

template <class F>
bool unstage_cvector(F* X, const F* d, int64_t n, int64_t inc, hipStream_t s) {
  const int64_t ainc = inc < 0 ? -inc : inc;
  const int64_t span = 2 * ((n - 1) * ainc + 1);
  F* lo = (inc < 0) ? X + 2 * (n - 1) * inc : X;
  const F* dlo = (inc < 0) ? d - 2 * (n - 1) * ainc : d;
  return RB_HIP(hipMemcpyAsync(lo, dlo, sizeof(F) * (size_t)span, hipMemcpyDeviceToHost, s));
}

inline bool notrans(char t) { return t == 'n' || t == 'N'; }

// reference src/reproBLAS/rzgemv.c (rcgemv.c) on host or device operands
template <class F>
void run_rzgemv(int fold, char Order, char TransA, int M, int N, const F* alpha, const F* A, int lda, const F* X,
                int incX, const F* beta, F* Y, int incY) {
  clear_error();
  if (M <= 0 || N <= 0) return;
  if (!check_fold<F>(fold)) return;
  hipStream_t s = current_stream();
  RB_STREAM_GUARD(s);
  const int64_t lenx = notrans(TransA) ? N : M, leny = notrans(TransA) ? M : N;
  bool sa, sx, sy;
  const F* dA = stage_matrix(A, 2 * matrix_span(Order, M, N, lda), kSlotStageX, s, &sa);
  const F* dX = stage_cvector(X, lenx, incX, kSlotStageY, s, &sx);
  F* dY = (F*)stage_cvector(Y, leny, incY, kSlotStageZ, s, &sy);
  if (!dA || !dX || !dY) return;
  if (!rzgemv_device(fold, Order, TransA, M, N, alpha, dA, lda, dX, incX, beta, dY, incY, s)) return;
  if (sy && !unstage_cvector(Y, dY, leny, incY, s)) return;
  RB_HIP(hipStreamSynchronize(s));
}

// reference src/reproBLAS/rzgemm.c (rcgemm.c)
template <class F>
void run_rzgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K, const F* alpha, const F* A,
                int lda, const F* B, int ldb, const F* beta, F* C, int ldc) {
  clear_error();
  if (M <= 0 || N <= 0) return;
  if (!check_fold<F>(fold)) return;
  hipStream_t s = current_stream();
  RB_STREAM_GUARD(s);
  const bool na = notrans(TransA), nb = notrans(TransB);
  bool sa, sb, sc;
  const F* dA = stage_matrix(A, 2 * matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa);
  const F* dB = stage_matrix(B, 2 * matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb);
  const int64_t cspan = 2 * matrix_span(Order, M, N, ldc);
  F* dC = (F*)stage_matrix(C, cspan, kSlotStageZ, s, &sc);
  if ((K > 0 && (!dA || !dB)) || !dC) return;
  if (!rzgemm_device(fold, Order, TransA, TransB, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, s)) return;
  if (sc && !RB_HIP(hipMemcpyAsync(C, dC, sizeof(F) * (size_t)cspan, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}

// reference src/binnedBLAS/zbzgemv.ccog (cbcgemv.ccog): Y complex binned, stride incY
template <class F>
void run_zbzgemv(int fold, char Order, char TransA, int M, int N, const F* alpha, const F* A, int lda, const F* X,
                 int incX, F* Y, int incY) {
  clear_error();
  if (M <= 0 || N <= 0 || (alpha[0] == F(0) && alpha[1] == F(0))) return;
  if (!check_fold<F>(fold)) return;
  hipStream_t s = current_stream();
  RB_STREAM_GUARD(s);
  const int64_t lenx = notrans(TransA) ? N : M, leny = notrans(TransA) ? M : N;
  bool sa, sx;
  const F* dA = stage_matrix(A, 2 * matrix_span(Order, M, N, lda), kSlotStageX, s, &sa);
  const F* dX = stage_cvector(X, lenx, incX, kSlotStageY, s, &sx);
  if (!dA || !dX) return;
  const size_t ybytes = sizeof(F) * (size_t)(((leny - 1) * (incY < 0 ? -incY : incY) + 1) * 4 * fold);
  const bool ydev = is_device_pointer(Y);
  F* dY = Y;
  if (!ydev) {
    if (incY < 0) {
      set_error(kUnsupported, "binnedBLAS_zbzgemv: negative incY on a host binned vector");
      return;
    }
    dY = (F*)workspace(s, kSlotBinnedOut, ybytes);
    if (!dY || !RB_HIP(hipMemcpyAsync(dY, Y, ybytes, hipMemcpyHostToDevice, s))) return;
  }
  if (!zbzgemv(fold, Order, TransA, M, N, alpha, dA, lda, dX, incX, dY, incY, s)) return;
  if (!ydev && !RB_HIP(hipMemcpyAsync(Y, dY, ybytes, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}

// reference src/binnedBLAS/zbzgemm.ccog (cbcgemm.ccog): C complex binned, leading dimension ldc
template <class F>
void run_zbzgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K, const F* alpha, const F* A,
                 int lda, const F* B, int ldb, F* C, int ldc) {
  clear_error();
  if (M <= 0 || N <= 0 || K <= 0 || (alpha[0] == F(0) && alpha[1] == F(0))) return;
  if (!check_fold<F>(fold)) return;
  hipStream_t s = current_stream();
  RB_STREAM_GUARD(s);
  const bool na = notrans(TransA), nb = notrans(TransB);
  bool sa, sb;
  const F* dA = stage_matrix(A, 2 * matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa);
  const F* dB = stage_matrix(B, 2 * matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb);
  if (!dA || !dB) return;
  const size_t cbytes = sizeof(F) * (size_t)matrix_span(Order, M, N, ldc) * 4 * fold;
  const bool cdev = is_device_pointer(C);
  F* dC = C;
  if (!cdev) {
    dC = (F*)workspace(s, kSlotBinnedOut, cbytes);
    if (!dC || !RB_HIP(hipMemcpyAsync(dC, C, cbytes, hipMemcpyHostToDevice, s))) return;
  }
  if (!zbzgemm(fold, Order, TransA, TransB, M, N, K, alpha, dA, lda, dB, ldb, dC, ldc, s)) return;
  if (!cdev && !RB_HIP(hipMemcpyAsync(C, dC, cbytes, hipMemcpyDeviceToHost, s))) return;
  RB_HIP(hipStreamSynchronize(s));
}

template <class F>
int device_rzgemv(int fold, char Order, char TransA, int64_t M, int64_t N, const F* alpha, const F* A, int64_t lda,
                  const F* X, int64_t incX, const F* beta, F* Y, int64_t incY, void* stream) {
  clear_error();
  RB_STREAM_GUARD(stream);
  if (M <= 0 || N <= 0) return kOk;
  if (!check_fold<F>(fold)) return last_error();
  const bool az = alpha[0] == F(0) && alpha[1] == F(0);
  if (!require_device_accessible(Y, "Y") ||
      (!az && (!require_device_accessible(A, "A") || !require_device_accessible(X, "X"))))
    return last_error();
  if (!rzgemv_device(fold, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY, (hipStream_t)stream))
    return last_error();
  return kOk;
}

template <class F>
int device_rzgemm(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K, const F* alpha,
                  const F* A, int64_t lda, const F* B, int64_t ldb, const F* beta, F* C, int64_t ldc, void* stream) {
  clear_error();
  RB_STREAM_GUARD(stream);
  if (M <= 0 || N <= 0) return kOk;
  if (!check_fold<F>(fold)) return last_error();
  const bool az = alpha[0] == F(0) && alpha[1] == F(0);
  if (!require_device_accessible(C, "C") ||
      (!az && K > 0 && (!require_device_accessible(A, "A") || !require_device_accessible(B, "B"))))
    return last_error();
  if (!rzgemm_device(fold, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, (hipStream_t)stream))
    return last_error();
  return kOk;
}

}  // namespace

extern "C" {

// ------------------------------------------------------ double complex --
// reference src/reproBLAS/rzsum_sub.c, rzdotu_sub.c, rzdotc_sub.c, rdzasum.c, rdznrm2.c
void reproBLAS_rzsum_sub(const int fold, const int N, const void* X, int incX, void* sum) {
  run_zblas1<double>(fold, kZSum, N, (const double*)X, incX, nullptr, 0, nullptr, nullptr, (double*)sum);
}
void reproBLAS_rzdotu_sub(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                          void* dotu) {
  run_zblas1<double>(fold, kZDotU, N, (const double*)X, incX, (const double*)Y, incY, nullptr, nullptr,
                     (double*)dotu);
}
void reproBLAS_rzdotc_sub(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                          void* dotc) {
  run_zblas1<double>(fold, kZDotC, N, (const double*)X, incX, (const double*)Y, incY, nullptr, nullptr,
                     (double*)dotc);
}
double reproBLAS_rdzasum(const int fold, const int N, const void* X, const int incX) {
  return run_zblas1_real<double>(fold, kZASum, N, (const double*)X, incX, nullptr, nullptr);
}
double reproBLAS_rdznrm2(const int fold, const int N, const void* X, int incX) {
  return run_zblas1_real<double>(fold, kZSSq, N, (const double*)X, incX, nullptr, nullptr);
}
// reference src/reproBLAS/zsum_sub.c and siblings (fold DIDEFAULTFOLD)
void reproBLAS_zsum_sub(const int N, const void* X, int incX, void* sum) {
  reproBLAS_rzsum_sub(DIDEFAULTFOLD, N, X, incX, sum);
}
double reproBLAS_dzasum(const int N, const void* X, const int incX) {
  return reproBLAS_rdzasum(DIDEFAULTFOLD, N, X, incX);
}
double reproBLAS_dznrm2(const int N, const void* X, int incX) { return reproBLAS_rdznrm2(DIDEFAULTFOLD, N, X, incX); }
void reproBLAS_zdotu_sub(const int N, const void* X, const int incX, const void* Y, const int incY, void* dotu) {
  reproBLAS_rzdotu_sub(DIDEFAULTFOLD, N, X, incX, Y, incY, dotu);
}
void reproBLAS_zdotc_sub(const int N, const void* X, const int incX, const void* Y, const int incY, void* dotc) {
  reproBLAS_rzdotc_sub(DIDEFAULTFOLD, N, X, incX, Y, incY, dotc);
}

// reference src/binnedBLAS/zamax_sub.ccog, zamaxm_sub.ccog (per component)
void binnedBLAS_zamax_sub(const int N, const void* X, const int incX, void* amax) {
  complex_amax<double>(N, X, incX, amax);
}
void binnedBLAS_zamaxm_sub(const int N, const void* X, const int incX, const void* Y, const int incY, void* amaxm) {
  complex_amaxm<double>(N, X, incX, Y, incY, amaxm);
}
// reference src/binnedBLAS/zbzsum.c, zbzdotu.c, zbzdotc.c, dbzasum.c, dbzssq.c
void binnedBLAS_zbzsum(const int fold, const int N, const void* X, const int incX, double_complex_binned* Y) {
  double out[2];
  run_zblas1<double>(fold, kZSum, N, (const double*)X, incX, nullptr, 0, Y, nullptr, out);
}
void binnedBLAS_zbzdotu(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        double_complex_binned* Z) {
  double out[2];
  run_zblas1<double>(fold, kZDotU, N, (const double*)X, incX, (const double*)Y, incY, Z, nullptr, out);
}
void binnedBLAS_zbzdotc(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        double_complex_binned* Z) {
  double out[2];
  run_zblas1<double>(fold, kZDotC, N, (const double*)X, incX, (const double*)Y, incY, Z, nullptr, out);
}
void binnedBLAS_dbzasum(const int fold, const int N, const void* X, const int incX, double_binned* Y) {
  run_zblas1_real<double>(fold, kZASum, N, (const double*)X, incX, Y, nullptr);
}
double binnedBLAS_dbzssq(const int fold, const int N, const void* X, const int incX, const double scaleY,
                         double_binned* Y) {
  double scale = scaleY;
  run_zblas1_real<double>(fold, kZSSq, N, (const double*)X, incX, Y, &scale);
  return scale;
}
// reference src/binnedBLAS/zmzsum.ccog, dmzasum.ccog, dmzssq.ccog, zmzdotu.ccog, zmzdotc.ccog
void binnedBLAS_zmzsum(const int fold, const int N, const void* X, const int incX, double* priY, const int incpriY,
                       double* carY, const int inccarY) {
  with_packed_complex(fold, priY, incpriY, carY, inccarY,
                      [&](double* st) { binnedBLAS_zbzsum(fold, N, X, incX, st); });
}
void binnedBLAS_dmzasum(const int fold, const int N, const void* X, const int incX, double* priY, const int incpriY,
                        double* carY, const int inccarY) {
  with_packed(fold, priY, incpriY, carY, inccarY, [&](double* st) { binnedBLAS_dbzasum(fold, N, X, incX, st); });
}
double binnedBLAS_dmzssq(const int fold, const int N, const void* X, const int incX, const double scaleY,
                         double* priY, const int incpriY, double* carY, const int inccarY) {
  double r = scaleY;
  with_packed(fold, priY, incpriY, carY, inccarY,
              [&](double* st) { r = binnedBLAS_dbzssq(fold, N, X, incX, scaleY, st); });
  return r;
}
void binnedBLAS_zmzdotu(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        double* priZ, const int incpriZ, double* carZ, const int inccarZ) {
  with_packed_complex(fold, priZ, incpriZ, carZ, inccarZ,
                      [&](double* st) { binnedBLAS_zbzdotu(fold, N, X, incX, Y, incY, st); });
}
void binnedBLAS_zmzdotc(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        double* priZ, const int incpriZ, double* carZ, const int inccarZ) {
  with_packed_complex(fold, priZ, incpriZ, carZ, inccarZ,
                      [&](double* st) { binnedBLAS_zbzdotc(fold, N, X, incX, Y, incY, st); });
}

// ------------------------------------------------------- float complex --
// reference src/reproBLAS/rcsum_sub.c, rcdotu_sub.c, rcdotc_sub.c, rscasum.c, rscnrm2.c
void reproBLAS_rcsum_sub(const int fold, const int N, const void* X, const int incX, void* sum) {
  run_zblas1<float>(fold, kZSum, N, (const float*)X, incX, nullptr, 0, nullptr, nullptr, (float*)sum);
}
void reproBLAS_rcdotu_sub(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                          void* dotu) {
  run_zblas1<float>(fold, kZDotU, N, (const float*)X, incX, (const float*)Y, incY, nullptr, nullptr, (float*)dotu);
}
void reproBLAS_rcdotc_sub(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                          void* dotc) {
  run_zblas1<float>(fold, kZDotC, N, (const float*)X, incX, (const float*)Y, incY, nullptr, nullptr, (float*)dotc);
}
float reproBLAS_rscasum(const int fold, const int N, const void* X, const int incX) {
  return run_zblas1_real<float>(fold, kZASum, N, (const float*)X, incX, nullptr, nullptr);
}
float reproBLAS_rscnrm2(const int fold, const int N, const void* X, const int incX) {
  return run_zblas1_real<float>(fold, kZSSq, N, (const float*)X, incX, nullptr, nullptr);
}
// reference src/reproBLAS/csum_sub.c and siblings (fold SIDEFAULTFOLD)
void reproBLAS_csum_sub(const int N, const void* X, const int incX, void* sum) {
  reproBLAS_rcsum_sub(SIDEFAULTFOLD, N, X, incX, sum);
}
float reproBLAS_scasum(const int N, const void* X, const int incX) {
  return reproBLAS_rscasum(SIDEFAULTFOLD, N, X, incX);
}
float reproBLAS_scnrm2(const int N, const void* X, const int incX) {
  return reproBLAS_rscnrm2(SIDEFAULTFOLD, N, X, incX);
}
void reproBLAS_cdotu_sub(const int N, const void* X, const int incX, const void* Y, const int incY, void* dotu) {
  reproBLAS_rcdotu_sub(SIDEFAULTFOLD, N, X, incX, Y, incY, dotu);
}
void reproBLAS_cdotc_sub(const int N, const void* X, const int incX, const void* Y, const int incY, void* dotc) {
  reproBLAS_rcdotc_sub(SIDEFAULTFOLD, N, X, incX, Y, incY, dotc);
}

// reference src/binnedBLAS/camax_sub.ccog, camaxm_sub.ccog
void binnedBLAS_camax_sub(const int N, const void* X, const int incX, void* amax) {
  complex_amax<float>(N, X, incX, amax);
}
void binnedBLAS_camaxm_sub(const int N, const void* X, const int incX, const void* Y, const int incY, void* amaxm) {
  complex_amaxm<float>(N, X, incX, Y, incY, amaxm);
}
// reference src/binnedBLAS/cbcsum.c, cbcdotu.c, cbcdotc.c, sbcasum.c, sbcssq.c
void binnedBLAS_cbcsum(const int fold, const int N, const void* X, const int incX, float_complex_binned* Y) {
  float out[2];
  run_zblas1<float>(fold, kZSum, N, (const float*)X, incX, nullptr, 0, Y, nullptr, out);
}
void binnedBLAS_cbcdotu(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        float_complex_binned* Z) {
  float out[2];
  run_zblas1<float>(fold, kZDotU, N, (const float*)X, incX, (const float*)Y, incY, Z, nullptr, out);
}
void binnedBLAS_cbcdotc(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        float_complex_binned* Z) {
  float out[2];
  run_zblas1<float>(fold, kZDotC, N, (const float*)X, incX, (const float*)Y, incY, Z, nullptr, out);
}
void binnedBLAS_sbcasum(const int fold, const int N, const void* X, const int incX, float_binned* Y) {
  run_zblas1_real<float>(fold, kZASum, N, (const float*)X, incX, Y, nullptr);
}
float binnedBLAS_sbcssq(const int fold, const int N, const void* X, const int incX, const float scaleY,
                        float_binned* Y) {
  float scale = scaleY;
  run_zblas1_real<float>(fold, kZSSq, N, (const float*)X, incX, Y, &scale);
  return scale;
}
// reference src/binnedBLAS/cmcsum.ccog, smcasum.ccog, smcssq.ccog, cmcdotu.ccog, cmcdotc.ccog
void binnedBLAS_cmcsum(const int fold, const int N, const void* X, const int incX, float* priY, const int incpriY,
                       float* carY, const int inccarY) {
  with_packed_complex(fold, priY, incpriY, carY, inccarY,
                      [&](float* st) { binnedBLAS_cbcsum(fold, N, X, incX, st); });
}
void binnedBLAS_smcasum(const int fold, const int N, const void* X, const int incX, float* priY, const int incpriY,
                        float* carY, const int inccarY) {
  with_packed(fold, priY, incpriY, carY, inccarY, [&](float* st) { binnedBLAS_sbcasum(fold, N, X, incX, st); });
}
float binnedBLAS_smcssq(const int fold, const int N, const void* X, const int incX, const float scaleY, float* priY,
                        const int incpriY, float* carY, const int inccarY) {
  float r = scaleY;
  with_packed(fold, priY, incpriY, carY, inccarY,
              [&](float* st) { r = binnedBLAS_sbcssq(fold, N, X, incX, scaleY, st); });
  return r;
}
void binnedBLAS_cmcdotu(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        float* priZ, const int incpriZ, float* carZ, const int inccarZ) {
  with_packed_complex(fold, priZ, incpriZ, carZ, inccarZ,
                      [&](float* st) { binnedBLAS_cbcdotu(fold, N, X, incX, Y, incY, st); });
}
void binnedBLAS_cmcdotc(const int fold, const int N, const void* X, const int incX, const void* Y, const int incY,
                        float* priZ, const int incpriZ, float* carZ, const int inccarZ) {
  with_packed_complex(fold, priZ, incpriZ, carZ, inccarZ,
                      [&](float* st) { binnedBLAS_cbcdotc(fold, N, X, incX, Y, incY, st); });
}

// ------------------------------------------------- device entry points --
int reproblas_hip_rzsum_device(int fold, int64_t N, const double* X, int64_t incX, double* result, void* stream) {
  return device_zblas1<double>(fold, kZSum, N, X, incX, nullptr, 0, result, stream);
}
int reproblas_hip_rzdotu_device(int fold, int64_t N, const double* X, int64_t incX, const double* Y, int64_t incY,
                                double* result, void* stream) {
  return device_zblas1<double>(fold, kZDotU, N, X, incX, Y, incY, result, stream);
}
int reproblas_hip_rzdotc_device(int fold, int64_t N, const double* X, int64_t incX, const double* Y, int64_t incY,
                                double* result, void* stream) {
  return device_zblas1<double>(fold, kZDotC, N, X, incX, Y, incY, result, stream);
}
int reproblas_hip_rdzasum_device(int fold, int64_t N, const double* X, int64_t incX, double* result, void* stream) {
  return device_zblas1<double>(fold, kZASum, N, X, incX, nullptr, 0, result, stream);
}
int reproblas_hip_rdznrm2_device(int fold, int64_t N, const double* X, int64_t incX, double* result, void* stream) {
  return device_zblas1<double>(fold, kZSSq, N, X, incX, nullptr, 0, result, stream);
}
int reproblas_hip_rcsum_device(int fold, int64_t N, const float* X, int64_t incX, float* result, void* stream) {
  return device_zblas1<float>(fold, kZSum, N, X, incX, nullptr, 0, result, stream);
}
int reproblas_hip_rcdotu_device(int fold, int64_t N, const float* X, int64_t incX, const float* Y, int64_t incY,
                                float* result, void* stream) {
  return device_zblas1<float>(fold, kZDotU, N, X, incX, Y, incY, result, stream);
}
int reproblas_hip_rcdotc_device(int fold, int64_t N, const float* X, int64_t incX, const float* Y, int64_t incY,
                                float* result, void* stream) {
  return device_zblas1<float>(fold, kZDotC, N, X, incX, Y, incY, result, stream);
}
int reproblas_hip_rscasum_device(int fold, int64_t N, const float* X, int64_t incX, float* result, void* stream) {
  return device_zblas1<float>(fold, kZASum, N, X, incX, nullptr, 0, result, stream);
}
int reproblas_hip_rscnrm2_device(int fold, int64_t N, const float* X, int64_t incX, float* result, void* stream) {
  return device_zblas1<float>(fold, kZSSq, N, X, incX, nullptr, 0, result, stream);
}

// ------------------------------------------------------------ gemv / gemm --
// reference src/reproBLAS/rzgemv.c, rzgemm.c, rcgemv.c, rcgemm.c and the
// default-fold zgemv.c, zgemm.c, cgemv.c, cgemm.c
void reproBLAS_rzgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                      const void* alpha, const void* A, const int lda, const void* X, const int incX,
                      const void* beta, void* Y, const int incY) {
  run_rzgemv<double>(fold, Order, TransA, M, N, (const double*)alpha, (const double*)A, lda, (const double*)X, incX,
                     (const double*)beta, (double*)Y, incY);
}
void reproBLAS_rzgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const void* alpha, const void* A, const int lda, const void* B,
                      const int ldb, const void* beta, void* C, const int ldc) {
  run_rzgemm<double>(fold, Order, TransA, TransB, M, N, K, (const double*)alpha, (const double*)A, lda,
                     (const double*)B, ldb, (const double*)beta, (double*)C, ldc);
}
void reproBLAS_rcgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                      const void* alpha, const void* A, const int lda, const void* X, const int incX,
                      const void* beta, void* Y, const int incY) {
  run_rzgemv<float>(fold, Order, TransA, M, N, (const float*)alpha, (const float*)A, lda, (const float*)X, incX,
                    (const float*)beta, (float*)Y, incY);
}
void reproBLAS_rcgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const void* alpha, const void* A, const int lda, const void* B,
                      const int ldb, const void* beta, void* C, const int ldc) {
  run_rzgemm<float>(fold, Order, TransA, TransB, M, N, K, (const float*)alpha, (const float*)A, lda,
                    (const float*)B, ldb, (const float*)beta, (float*)C, ldc);
}
void reproBLAS_zgemv(const char Order, const char TransA, const int M, const int N, const void* alpha, const void* A,
                     const int lda, const void* X, const int incX, const void* beta, void* Y, const int incY) {
  reproBLAS_rzgemv(DIDEFAULTFOLD, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY);
}
void reproBLAS_zgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const void* alpha, const void* A, const int lda, const void* B, const int ldb, const void* beta,
                     void* C, const int ldc) {
  reproBLAS_rzgemm(DIDEFAULTFOLD, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}
void reproBLAS_cgemv(const char Order, const char TransA, const int M, const int N, const void* alpha, const void* A,
                     const int lda, const void* X, const int incX, const void* beta, void* Y, const int incY) {
  reproBLAS_rcgemv(SIDEFAULTFOLD, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY);
}
void reproBLAS_cgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const void* alpha, const void* A, const int lda, const void* B, const int ldb, const void* beta,
                     void* C, const int ldc) {
  reproBLAS_rcgemm(SIDEFAULTFOLD, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}
// reference src/binnedBLAS/zbzgemv.ccog, zbzgemm.ccog, cbcgemv.ccog, cbcgemm.ccog
void binnedBLAS_zbzgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const void* alpha, const void* A, const int lda, const void* X, const int incX,
                        double_complex_binned* Y, const int incY) {
  run_zbzgemv<double>(fold, Order, TransA, M, N, (const double*)alpha, (const double*)A, lda, (const double*)X, incX,
                      Y, incY);
}
void binnedBLAS_zbzgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const void* alpha, const void* A, const int lda, const void* B,
                        const int ldb, double_complex_binned* C, const int ldc) {
  run_zbzgemm<double>(fold, Order, TransA, TransB, M, N, K, (const double*)alpha, (const double*)A, lda,
                      (const double*)B, ldb, C, ldc);
}
void binnedBLAS_cbcgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const void* alpha, const void* A, const int lda, const void* X, const int incX,
                        float_complex_binned* Y, const int incY) {
  run_zbzgemv<float>(fold, Order, TransA, M, N, (const float*)alpha, (const float*)A, lda, (const float*)X, incX, Y,
                     incY);
}
void binnedBLAS_cbcgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const void* alpha, const void* A, const int lda, const void* B,
                        const int ldb, float_complex_binned* C, const int ldc) {
  run_zbzgemm<float>(fold, Order, TransA, TransB, M, N, K, (const float*)alpha, (const float*)A, lda,
                     (const float*)B, ldb, C, ldc);
}
int reproblas_hip_rzgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, const double* alpha,
                                const double* A, int64_t lda, const double* X, int64_t incX, const double* beta,
                                double* Y, int64_t incY, void* stream) {
  return device_rzgemv<double>(fold, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY, stream);
}
int reproblas_hip_rzgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                const double* alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                                const double* beta, double* C, int64_t ldc, void* stream) {
  return device_rzgemm<double>(fold, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, stream);
}
int reproblas_hip_rcgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, const float* alpha,
                                const float* A, int64_t lda, const float* X, int64_t incX, const float* beta,
                                float* Y, int64_t incY, void* stream) {
  return device_rzgemv<float>(fold, Order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY, stream);
}
int reproblas_hip_rcgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                const float* alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                                const float* beta, float* C, int64_t ldc, void* stream) {
  return device_rzgemm<float>(fold, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, stream);
}

}  // extern "C"
