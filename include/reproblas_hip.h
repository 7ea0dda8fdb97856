/*
 * reproblas_hip.h -- HIP-specific additions of libreproblas_hip.so.
 *
 * None of these has a counterpart in the reference (ReproBLAS is a CPU
 * library with no streams or device memory); they exist for callers that keep
 * operands resident on the GPU: the PyTorch autograd function, the ragged
 * batch driver, graph capture and the benchmarks. The reference-compatible
 * entry points are in reproBLAS.h, binnedBLAS.h and binned.h.
 *
 * Conventions: every *_device call takes device pointers only, enqueues its
 * work on `stream` (a hipStream_t, NULL for the null stream) and returns
 * without synchronising; results land in device memory. The return value is
 * 0 on success or a nonzero error code (also kept for
 * reproblas_hip_last_error()). Calls that share a stream share the library's
 * per-stream scratch, so calls on one stream are ordered and safe; calls on
 * different streams use different scratch.
 */
#ifndef REPROBLAS_HIP_H_
#define REPROBLAS_HIP_H_
#include <stdint.h>

#include "binned.h"

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define REPROBLAS_HIP_OK 0
#define REPROBLAS_HIP_BAD_FOLD 1
#define REPROBLAS_HIP_BAD_ARGUMENT 2
#define REPROBLAS_HIP_HIP_ERROR 3
#define REPROBLAS_HIP_NO_DEVICE 4
#define REPROBLAS_HIP_UNSUPPORTED 5

int reproblas_hip_last_error(void);
const char *reproblas_hip_last_error_string(void);
void reproblas_hip_clear_error(void);

/* Stream used by the synchronous reproBLAS_/binnedBLAS_ calls (like a BLAS
 * handle's stream). Default: the null stream. */
void reproblas_hip_set_stream(void *stream);
void *reproblas_hip_get_stream(void);

/* Folds accepted by the GPU kernels: [reproblas_hip_min_fold(), reproblas_hip_max_fold()]. */
int reproblas_hip_min_fold(void);
int reproblas_hip_max_fold(void);

/* BLAS1 on device vectors, result (a double) written to device memory. */
int reproblas_hip_rdsum_device(int fold, int64_t N, const double *X, int64_t incX, double *result, void *stream);
int reproblas_hip_rdasum_device(int fold, int64_t N, const double *X, int64_t incX, double *result, void *stream);
int reproblas_hip_rdnrm2_device(int fold, int64_t N, const double *X, int64_t incX, double *result, void *stream);
int reproblas_hip_rddot_device(int fold, int64_t N, const double *X, int64_t incX, const double *Y, int64_t incY,
                               double *result, void *stream);

/* Binned-state BLAS1 on device data: accumulate into the device binned
 * value Y (2*fold doubles), as binnedBLAS_dbdsum/dbdasum/dbddot do. For
 * dbdssq the scale lives in device memory too (*scale in, new scale out). */
int reproblas_hip_dbdsum_device(int fold, int64_t N, const double *X, int64_t incX, double_binned *Y, void *stream);
int reproblas_hip_dbdasum_device(int fold, int64_t N, const double *X, int64_t incX, double_binned *Y, void *stream);
int reproblas_hip_dbddot_device(int fold, int64_t N, const double *X, int64_t incX, const double *Y, int64_t incY,
                                double_binned *Z, void *stream);
int reproblas_hip_dbdssq_device(int fold, int64_t N, const double *X, int64_t incX, double *scale,
                                double_binned *Y, void *stream);

/* gemv / gemm on device operands (reproBLAS_rdgemv / rdgemm semantics). */
int reproblas_hip_rdgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, double alpha,
                                const double *A, int64_t lda, const double *X, int64_t incX, double beta, double *Y,
                                int64_t incY, void *stream);
int reproblas_hip_rdgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                                double beta, double *C, int64_t ldc, void *stream);

/* gemm tiling (the autotuner's knob; every tiling gives identical bits).
 * Returns the number of tilings; set/get select by ordinal. */
int reproblas_hip_gemm_tiling_count(void);
int reproblas_hip_gemm_set_tiling(int tiling);
int reproblas_hip_gemm_get_tiling(void);

#ifdef __cplusplus
}
#endif
#endif
