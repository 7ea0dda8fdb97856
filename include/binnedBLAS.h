/*
 * binnedBLAS.h -- binned BLAS over double and float vectors and matrices, as
 * exported by libreproblas_hip.so.
 *
 * Drop-in for the real (double and float) part of the reference's
 * include/binnedBLAS.h (ReproBLAS 2.1.0). The dbd* / sbs* calls accumulate
 * into a caller-owned binned value (or vector of them), as the reference's do.
 *
 * Operands may live in device memory (hipMalloc / managed) -- then the work
 * runs in place on the GPU -- or in host memory, which is staged to the GPU
 * and back. The reductions always run on the GPU; the library has no CPU
 * compute path. Calls run on the stream set with reproblas_hip_set_stream()
 * and return after it has finished. Errors are reported through
 * reproblas_hip_last_error().
 */
#ifndef BINNEDBLAS_H_
#define BINNEDBLAS_H_
#include "binned.h"

#ifdef __cplusplus
extern "C" {
#endif

double binnedBLAS_damax(const int N, const double *X, const int incX);              /* binnedBLAS.h:38 */
double binnedBLAS_damaxm(const int N, const double *X, const int incX, const double *Y,
                         const int incY);                                           /* binnedBLAS.h:43 */
void binnedBLAS_dbdsum(const int fold, const int N, const double *X, const int incX,
                       double_binned *Y);                                           /* binnedBLAS.h:47 */
void binnedBLAS_dmdsum(const int fold, const int N, const double *X, const int incX, double *priY,
                       const int incpriY, double *carY, const int inccarY);         /* binnedBLAS.h:48 */
void binnedBLAS_dbdasum(const int fold, const int N, const double *X, const int incX,
                        double_binned *Y);                                          /* binnedBLAS.h:49 */
void binnedBLAS_dmdasum(const int fold, const int N, const double *X, const int incX, double *priY,
                        const int incpriY, double *carY, const int inccarY);        /* binnedBLAS.h:50 */
double binnedBLAS_dbdssq(const int fold, const int N, const double *X, const int incX, const double scaleY,
                         double_binned *Y);                                         /* binnedBLAS.h:51 */
double binnedBLAS_dmdssq(const int fold, const int N, const double *X, const int incX, const double scaleY,
                         double *priY, const int incpriY, double *carY, const int inccarY); /* binnedBLAS.h:52 */
void binnedBLAS_dbddot(const int fold, const int N, const double *X, const int incX, const double *Y,
                       const int incY, double_binned *Z);                           /* binnedBLAS.h:53 */
void binnedBLAS_dmddot(const int fold, const int N, const double *X, const int incX, const double *Y,
                       const int incY, double *priZ, const int incpriZ, double *carZ,
                       const int inccarZ);                                          /* binnedBLAS.h:54 */
void binnedBLAS_dbdgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const double alpha, const double *A, const int lda, const double *X, const int incX,
                        double_binned *Y, const int incY);                          /* binnedBLAS.h:87 */
void binnedBLAS_dbdgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const double alpha, const double *A, const int lda,
                        const double *B, const int ldb, double_binned *C, const int ldc); /* binnedBLAS.h:92 */

/* single precision (float_binned accumulators) */
float binnedBLAS_samax(const int N, const float *X, const int incX);                /* binnedBLAS.h:37 */
float binnedBLAS_samaxm(const int N, const float *X, const int incX, const float *Y,
                        const int incY);                                            /* binnedBLAS.h:42 */
void binnedBLAS_sbssum(const int fold, const int N, const float *X, const int incX,
                       float_binned *Y);                                            /* binnedBLAS.h:67 */
void binnedBLAS_smssum(const int fold, const int N, const float *X, const int incX, float *priY,
                       const int incpriY, float *carY, const int inccarY);          /* binnedBLAS.h:68 */
void binnedBLAS_sbsasum(const int fold, const int N, const float *X, const int incX,
                        float_binned *Y);                                           /* binnedBLAS.h:69 */
void binnedBLAS_smsasum(const int fold, const int N, const float *X, const int incX, float *priY,
                        const int incpriY, float *carY, const int inccarY);         /* binnedBLAS.h:70 */
float binnedBLAS_sbsssq(const int fold, const int N, const float *X, const int incX, const float scaleY,
                        float_binned *Y);                                           /* binnedBLAS.h:71 */
float binnedBLAS_smsssq(const int fold, const int N, const float *X, const int incX, const float scaleY,
                        float *priY, const int incpriY, float *carY, const int inccarY); /* binnedBLAS.h:72 */
void binnedBLAS_sbsdot(const int fold, const int N, const float *X, const int incX, const float *Y,
                       const int incY, float_binned *Z);                            /* binnedBLAS.h:73 */
void binnedBLAS_smsdot(const int fold, const int N, const float *X, const int incX, const float *Y,
                       const int incY, float *manZ, const int incmanZ, float *carZ,
                       const int inccarZ);                                          /* binnedBLAS.h:74 */
void binnedBLAS_sbsgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const float alpha, const float *A, const int lda, const float *X, const int incX,
                        float_binned *Y, const int incY);                           /* binnedBLAS.h:99 */
void binnedBLAS_sbsgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const float alpha, const float *A, const int lda,
                        const float *B, const int ldb, float_binned *C, const int ldc); /* binnedBLAS.h:104 */

/* complex (interleaved pairs; complex binned = 4*fold values, real binned
 * accumulators for the asum/ssq of complex vectors) */
void binnedBLAS_zamax_sub(const int N, const void *X, const int incX, void *amax);                 /* :40 */
void binnedBLAS_zamaxm_sub(const int N, const void *X, const int incX, const void *Y, const int incY,
                           void *amaxm);                                                          /* :45 */
void binnedBLAS_camax_sub(const int N, const void *X, const int incX, void *amax);                 /* :39 */
void binnedBLAS_camaxm_sub(const int N, const void *X, const int incX, const void *Y, const int incY,
                           void *amaxm);                                                          /* :44 */
void binnedBLAS_zbzsum(const int fold, const int N, const void *X, const int incX, double_complex_binned *Y); /* :56 */
void binnedBLAS_zmzsum(const int fold, const int N, const void *X, const int incX, double *priY, const int incpriY,
                       double *carY, const int inccarY);                                          /* :57 */
void binnedBLAS_dbzasum(const int fold, const int N, const void *X, const int incX, double_binned *Y); /* :58 */
void binnedBLAS_dmzasum(const int fold, const int N, const void *X, const int incX, double *priY, const int incpriY,
                        double *carY, const int inccarY);                                         /* :59 */
double binnedBLAS_dbzssq(const int fold, const int N, const void *X, const int incX, const double scaleY,
                         double_binned *Y);                                                       /* :60 */
double binnedBLAS_dmzssq(const int fold, const int N, const void *X, const int incX, const double scaleY,
                         double *priY, const int incpriY, double *carY, const int inccarY);       /* :61 */
void binnedBLAS_zbzdotu(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        double_complex_binned *Z);                                                /* :62 */
void binnedBLAS_zmzdotu(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        double *priZ, const int incpriZ, double *carZ, const int inccarZ);        /* :63 */
void binnedBLAS_zbzdotc(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        double_complex_binned *Z);                                                /* :64 */
void binnedBLAS_zmzdotc(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        double *priZ, const int incpriZ, double *carZ, const int inccarZ);        /* :65 */
void binnedBLAS_cbcsum(const int fold, const int N, const void *X, const int incX, float_complex_binned *Y); /* :76 */
void binnedBLAS_cmcsum(const int fold, const int N, const void *X, const int incX, float *priY, const int incpriY,
                       float *carY, const int inccarY);                                           /* :77 */
void binnedBLAS_sbcasum(const int fold, const int N, const void *X, const int incX, float_binned *Y); /* :78 */
void binnedBLAS_smcasum(const int fold, const int N, const void *X, const int incX, float *priY, const int incpriY,
                        float *carY, const int inccarY);                                          /* :79 */
float binnedBLAS_sbcssq(const int fold, const int N, const void *X, const int incX, const float scaleY,
                        float_binned *Y);                                                         /* :80 */
float binnedBLAS_smcssq(const int fold, const int N, const void *X, const int incX, const float scaleY,
                        float *priY, const int incpriY, float *carY, const int inccarY);          /* :81 */
void binnedBLAS_cbcdotu(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        float_complex_binned *Z);                                                 /* :82 */
void binnedBLAS_cmcdotu(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        float *priZ, const int incpriZ, float *carZ, const int inccarZ);          /* :83 */
void binnedBLAS_cbcdotc(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        float_complex_binned *Z);                                                 /* :84 */
void binnedBLAS_cmcdotc(const int fold, const int N, const void *X, const int incX, const void *Y, const int incY,
                        float *priZ, const int incpriZ, float *carZ, const int inccarZ);          /* :85 */
void binnedBLAS_zbzgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const void *alpha, const void *A, const int lda, const void *X, const int incX,
                        double_complex_binned *Y, const int incY);                                /* :111 */
void binnedBLAS_zbzgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const void *alpha, const void *A, const int lda, const void *B,
                        const int ldb, double_complex_binned *C, const int ldc);                  /* :116 */
void binnedBLAS_cbcgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                        const void *alpha, const void *A, const int lda, const void *X, const int incX,
                        float_complex_binned *Y, const int incY);                                 /* :123 */
void binnedBLAS_cbcgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                        const int N, const int K, const void *alpha, const void *A, const int lda, const void *B,
                        const int ldb, float_complex_binned *C, const int ldc);                   /* :128 */

#ifdef __cplusplus
}
#endif
#endif
