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
 *
 * Threads: every entry point may be called from several host threads at
 * once. The host side of calls on the same (device, stream) is serialised
 * by a per-stream lock (held from staging through the final synchronize of
 * a synchronous call), so concurrent calls never share a staging or scratch
 * buffer in flight; calls on different streams do not contend. The error
 * state, the stream of the synchronous calls and the gemm tiling are per
 * host thread. A batch plan must not be launched concurrently with itself
 * (its tables and partials are the plan's own).
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

/* Stream used by the synchronous reproBLAS_/binnedBLAS_ calls of the calling
 * host thread (like a BLAS handle's stream). Default: the null stream. */
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

/* reproblas_hip_rddot_device without allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, result and bits as reproblas_hip_rddot_device. Its scratch
 * (the completion ticket and the accumulators through which the workgroups
 * merge) is a fixed 64 KiB device array of the library, one per GPU, kept
 * until the process ends; it does not depend on N, the strides or the fold,
 * so no size is refused. The call allocates nothing (the first call on a GPU
 * looks up that GPU's array and clears its sync block with a one-workgroup
 * kernel on `stream`). Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (different streams, or different host threads)
 * race and must be serialised by the caller. */
int reproblas_hip_rddot_device_noalloc(int fold, int64_t N, const double *X, int64_t incX, const double *Y,
                                       int64_t incY, double *result, void *stream);

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

/* Single-precision counterparts on device data (folds 2..21; float_binned
 * states of 2*fold floats; reproBLAS_rs* / binnedBLAS_sbs* semantics). */
int reproblas_hip_rssum_device(int fold, int64_t N, const float *X, int64_t incX, float *result, void *stream);
int reproblas_hip_rsasum_device(int fold, int64_t N, const float *X, int64_t incX, float *result, void *stream);
int reproblas_hip_rsnrm2_device(int fold, int64_t N, const float *X, int64_t incX, float *result, void *stream);
int reproblas_hip_rsdot_device(int fold, int64_t N, const float *X, int64_t incX, const float *Y, int64_t incY,
                               float *result, void *stream);
int reproblas_hip_sbssum_device(int fold, int64_t N, const float *X, int64_t incX, float_binned *Y, void *stream);
int reproblas_hip_sbsasum_device(int fold, int64_t N, const float *X, int64_t incX, float_binned *Y, void *stream);
int reproblas_hip_sbsdot_device(int fold, int64_t N, const float *X, int64_t incX, const float *Y, int64_t incY,
                                float_binned *Z, void *stream);
int reproblas_hip_sbsssq_device(int fold, int64_t N, const float *X, int64_t incX, float *scale, float_binned *Y,
                                void *stream);
int reproblas_hip_rsgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, float alpha,
                                const float *A, int64_t lda, const float *X, int64_t incX, float beta, float *Y,
                                int64_t incY, void *stream);
int reproblas_hip_rsgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                float alpha, const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
                                float *C, int64_t ldc, void *stream);

/* Complex counterparts on device data (interleaved pairs, increments in
 * complex elements): result holds two values for sum/dotu/dotc, one for
 * asum/nrm2 (reproBLAS_rz* / rc* / rdz* / rsc* semantics). */
int reproblas_hip_rzsum_device(int fold, int64_t N, const double *X, int64_t incX, double *result, void *stream);
int reproblas_hip_rzdotu_device(int fold, int64_t N, const double *X, int64_t incX, const double *Y, int64_t incY,
                                double *result, void *stream);
int reproblas_hip_rzdotc_device(int fold, int64_t N, const double *X, int64_t incX, const double *Y, int64_t incY,
                                double *result, void *stream);
int reproblas_hip_rdzasum_device(int fold, int64_t N, const double *X, int64_t incX, double *result, void *stream);
int reproblas_hip_rdznrm2_device(int fold, int64_t N, const double *X, int64_t incX, double *result, void *stream);
int reproblas_hip_rcsum_device(int fold, int64_t N, const float *X, int64_t incX, float *result, void *stream);
int reproblas_hip_rcdotu_device(int fold, int64_t N, const float *X, int64_t incX, const float *Y, int64_t incY,
                                float *result, void *stream);
int reproblas_hip_rcdotc_device(int fold, int64_t N, const float *X, int64_t incX, const float *Y, int64_t incY,
                                float *result, void *stream);
int reproblas_hip_rscasum_device(int fold, int64_t N, const float *X, int64_t incX, float *result, void *stream);
int reproblas_hip_rscnrm2_device(int fold, int64_t N, const float *X, int64_t incX, float *result, void *stream);

/* complex gemv / gemm on device operands; alpha and beta are host pairs */
int reproblas_hip_rzgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, const double *alpha,
                                const double *A, int64_t lda, const double *X, int64_t incX, const double *beta,
                                double *Y, int64_t incY, void *stream);
int reproblas_hip_rzgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                const double *alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                                const double *beta, double *C, int64_t ldc, void *stream);
int reproblas_hip_rcgemv_device(int fold, char Order, char TransA, int64_t M, int64_t N, const float *alpha,
                                const float *A, int64_t lda, const float *X, int64_t incX, const float *beta,
                                float *Y, int64_t incY, void *stream);
int reproblas_hip_rcgemm_device(int fold, char Order, char TransA, char TransB, int64_t M, int64_t N, int64_t K,
                                const float *alpha, const float *A, int64_t lda, const float *B, int64_t ldb,
                                const float *beta, float *C, int64_t ldc, void *stream);

/* gemm tiling (the autotuner's knob; every tiling gives identical bits) of
 * the calling host thread. Returns the number of tilings; set/get select by
 * ordinal. */
int reproblas_hip_gemm_tiling_count(void);
int reproblas_hip_gemm_set_tiling(int tiling);
int reproblas_hip_gemm_get_tiling(void);

/* Ragged batches (one persistent launch per group, optionally one hipGraph).
 * A plan is built once on the host from host-side sizes and device operand
 * pointers, then launched any number of times; every result is bit-identical
 * to the corresponding single call (reproBLAS_rddot etc. / reproBLAS_rdgemm).
 * Calls that allocate: create and add_* (plan tables); launches do not. */
/* Order-free exchange of binned values between processes: no counterpart
 * function in the reference, whose multi-process reduction is an MPI
 * operator merging pairs in the order MPI chooses (src/binnedMPI/
 * DBDBADD.ccog, binnedMPI_DBDBADD). A value is encoded as a "wire" of
 * REPROBLAS_HIP_DB_WIRE (float: _SB_WIRE) int64 words -- the offsets of its
 * primaries from their bins (in units of the bin's ulp) and its carries, each
 * at its absolute bin index (index + slot), then counts of NaN, +inf and -inf
 * classes -- and its index (REPROBLAS_HIP_EMPTY_INDEX for an empty value).
 * Summing the wires word by word and taking the minimum of the indexes over
 * any set of values (an all-reduce SUM and MIN), then decoding, gives the
 * bits that merging them with binned_dbdbadd gives, whatever the order or
 * grouping: integer addition is associative. For a non-finite result only
 * the top primary is specified (as for binned_dbdbadd, whose other slots
 * then depend on the merge order). Host memory; no GPU needed. */
#define REPROBLAS_HIP_DB_WIRE (2 * 104 + 3)
#define REPROBLAS_HIP_SB_WIRE (2 * 42 + 3)
#define REPROBLAS_HIP_EMPTY_INDEX 1048576
int reproblas_hip_dbencode(int fold, const double_binned *X, int64_t *wire, int64_t *index);
int reproblas_hip_dbdecode(int fold, const int64_t *wire, int64_t index, double_binned *Y);
int reproblas_hip_sbencode(int fold, const float_binned *X, int64_t *wire, int64_t *index);
int reproblas_hip_sbdecode(int fold, const int64_t *wire, int64_t index, float_binned *Y);

typedef struct reproblas_hip_batch reproblas_hip_batch;
#define REPROBLAS_HIP_OP_SUM 0
#define REPROBLAS_HIP_OP_ASUM 1
#define REPROBLAS_HIP_OP_DOT 2
#define REPROBLAS_HIP_OP_NRM2 3
reproblas_hip_batch *reproblas_hip_batch_create(int fold);
/* segment s of X (and Y for DOT) is [offsets[s], offsets[s+1]); offsets is a
 * host array of nseg+1 entries; results (device, nseg doubles); chunk: terms
 * per work item (0 for the default), a tuning knob only */
int reproblas_hip_batch_add_segments(reproblas_hip_batch *b, int op, int64_t nseg, const int64_t *offsets,
                                     const double *X, const double *Y, double *results, int64_t chunk);
/* count gemms of one Order/TransA/TransB; per-gemm sizes, scalars and device
 * pointers in host arrays; tiling as for reproblas_hip_gemm_set_tiling (-1:
 * the current one) */
int reproblas_hip_batch_add_gemms(reproblas_hip_batch *b, char Order, char TransA, char TransB, int count,
                                  const int64_t *M, const int64_t *N, const int64_t *K, const double *alpha,
                                  const double *const *A, const int64_t *lda, const double *const *B,
                                  const int64_t *ldb, const double *beta, double *const *C, const int64_t *ldc,
                                  int tiling);
int reproblas_hip_batch_launch(reproblas_hip_batch *b, void *stream);
int reproblas_hip_batch_graph_launch(reproblas_hip_batch *b, void *stream);
void reproblas_hip_batch_destroy(reproblas_hip_batch *b);

#ifdef __cplusplus
}
#endif
#endif
