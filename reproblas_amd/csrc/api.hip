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
#include "api_helpers.hpp"
#include "runtime.hpp"

using namespace rb;

using namespace rb::api;

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
  RB_STREAM_GUARD(s);
  const bool nt = (TransA == 'n' || TransA == 'N');
  const int64_t lenx = nt ? N : M, leny = nt ? M : N;
  bool sa = false, sx = false, sy;
  // alpha == 0: no product is formed and A, X are never read (rdgemv.c:35-87)
  const bool prod = alpha != 0.0;
  const double* dA = prod ? stage_matrix(A, matrix_span(Order, M, N, lda), kSlotStageX, s, &sa) : nullptr;
  const double* dX = prod ? stage_vector(X, lenx, incX, kSlotStageY, s, &sx) : nullptr;
  double* dY = (double*)stage_vector(Y, leny, incY, kSlotStageZ, s, &sy);
  if ((prod && (!dA || !dX)) || !dY) return;
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
  RB_STREAM_GUARD(s);
  const bool na = (TransA == 'n' || TransA == 'N'), nb = (TransB == 'n' || TransB == 'N');
  bool sa = false, sb = false, sc;
  // alpha == 0 or K <= 0: no product is formed and A, B are never read
  // (dbdgemm.ccog:77-80)
  const bool prod = K > 0 && alpha != 0.0;
  const double* dA =
      prod ? stage_matrix(A, matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa) : nullptr;
  const double* dB =
      prod ? stage_matrix(B, matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb) : nullptr;
  const int64_t cspan = matrix_span(Order, M, N, ldc);
  double* dC = (double*)stage_matrix(C, cspan, kSlotStageZ, s, &sc);
  if ((prod && (!dA || !dB)) || !dC) return;
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
  RB_STREAM_GUARD(s);
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
  RB_STREAM_GUARD(s);
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
  RB_STREAM_GUARD(stream);
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
  RB_STREAM_GUARD(stream);
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
  RB_STREAM_GUARD(s);
  const bool nt = (TransA == 'n' || TransA == 'N');
  const int64_t lenx = nt ? N : M, leny = nt ? M : N;
  bool sa = false, sx = false, sy;
  // alpha == 0: no product is formed and A, X are never read (rdgemv.c:35-87)
  const bool prod = alpha != 0.0f;
  const float* dA = prod ? stage_matrix(A, matrix_span(Order, M, N, lda), kSlotStageX, s, &sa) : nullptr;
  const float* dX = prod ? stage_vector(X, lenx, incX, kSlotStageY, s, &sx) : nullptr;
  float* dY = (float*)stage_vector(Y, leny, incY, kSlotStageZ, s, &sy);
  if ((prod && (!dA || !dX)) || !dY) return;
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
  RB_STREAM_GUARD(s);
  const bool na = (TransA == 'n' || TransA == 'N'), nb = (TransB == 'n' || TransB == 'N');
  bool sa = false, sb = false, sc;
  // alpha == 0 or K <= 0: no product is formed and A, B are never read
  // (dbdgemm.ccog:77-80)
  const bool prod = K > 0 && alpha != 0.0f;
  const float* dA =
      prod ? stage_matrix(A, matrix_span(Order, na ? M : K, na ? K : M, lda), kSlotStageX, s, &sa) : nullptr;
  const float* dB =
      prod ? stage_matrix(B, matrix_span(Order, nb ? K : N, nb ? N : K, ldb), kSlotStageY, s, &sb) : nullptr;
  const int64_t cspan = matrix_span(Order, M, N, ldc);
  float* dC = (float*)stage_matrix(C, cspan, kSlotStageZ, s, &sc);
  if ((prod && (!dA || !dB)) || !dC) return;
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
  RB_STREAM_GUARD(s);
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
  RB_STREAM_GUARD(s);
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
  RB_STREAM_GUARD(stream);
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
  RB_STREAM_GUARD(stream);
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
