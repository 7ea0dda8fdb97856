/*
 * reproBLAS.h -- reproducible double-precision BLAS, as exported by
 * libreproblas_hip.so.
 *
 * Drop-in for the double-precision part of the reference's
 * include/reproBLAS.h (ReproBLAS 2.1.0): same names, signatures, fold
 * semantics and results, bit for bit. The reproBLAS_r* calls take an explicit
 * fold; the others use DIDEFAULTFOLD (3, reference config.h:14).
 *
 * Operands may be device pointers (the work runs in place on the GPU) or host
 * pointers (staged to the GPU and back). Calls run on the stream set with
 * reproblas_hip_set_stream() and return once their results are ready.
 * Argument conventions follow the reference: Order 'r'/'R' is row-major and
 * anything else column-major; TransA/TransB 'n'/'N' is no-transpose and
 * anything else transpose; negative increments are not pointer-adjusted.
 */
#ifndef REPROBLAS_H_
#define REPROBLAS_H_

#ifdef __cplusplus
extern "C" {
#endif

#ifndef DIDEFAULTFOLD
#define DIDEFAULTFOLD 3 /* reference config.h:14 */
#endif
#ifndef SIDEFAULTFOLD
#define SIDEFAULTFOLD 3 /* reference config.h:22 */
#endif

double reproBLAS_rdsum(const int fold, const int N, const double *X, const int incX);  /* reproBLAS.h:27 */
double reproBLAS_rdasum(const int fold, const int N, const double *X, const int incX); /* reproBLAS.h:28 */
double reproBLAS_rdnrm2(const int fold, const int N, const double *X, const int incX); /* reproBLAS.h:29 */
double reproBLAS_rddot(const int fold, const int N, const double *X, const int incX, const double *Y,
                       const int incY);                                                 /* reproBLAS.h:30 */

void reproBLAS_rdgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                      const double alpha, const double *A, const int lda, const double *X, const int incX,
                      const double beta, double *Y, const int incY); /* reproBLAS.h:49 */
void reproBLAS_rdgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const double alpha, const double *A, const int lda,
                      const double *B, const int ldb, const double beta, double *C,
                      const int ldc); /* reproBLAS.h:54 */

double reproBLAS_dsum(const int N, const double *X, const int incX);  /* reproBLAS.h:97 */
double reproBLAS_dasum(const int N, const double *X, const int incX); /* reproBLAS.h:98 */
double reproBLAS_dnrm2(const int N, const double *X, const int incX); /* reproBLAS.h:99 */
double reproBLAS_ddot(const int N, const double *X, const int incX, const double *Y,
                      const int incY); /* reproBLAS.h:100 */

void reproBLAS_dgemv(const char Order, const char TransA, const int M, const int N, const double alpha,
                     const double *A, const int lda, const double *X, const int incX, const double beta, double *Y,
                     const int incY); /* reproBLAS.h:119 */
void reproBLAS_dgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const double alpha, const double *A, const int lda, const double *B, const int ldb,
                     const double beta, double *C, const int ldc); /* reproBLAS.h:124 */

/* single precision (float_binned accumulators, folds 2..21) */
float reproBLAS_rssum(const int fold, const int N, const float *X, const int incX);  /* reproBLAS.h:34 */
float reproBLAS_rsasum(const int fold, const int N, const float *X, const int incX); /* reproBLAS.h:33 */
float reproBLAS_rsnrm2(const int fold, const int N, const float *X, const int incX); /* reproBLAS.h:35 */
float reproBLAS_rsdot(const int fold, const int N, const float *X, const int incX, const float *Y,
                      const int incY);                                                /* reproBLAS.h:32 */
void reproBLAS_rsgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                      const float alpha, const float *A, const int lda, const float *X, const int incX,
                      const float beta, float *Y, const int incY); /* reproBLAS.h:61 */
void reproBLAS_rsgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const float alpha, const float *A, const int lda, const float *B,
                      const int ldb, const float beta, float *C, const int ldc); /* reproBLAS.h:66 */
float reproBLAS_ssum(const int N, const float *X, const int incX);  /* reproBLAS.h:104 */
float reproBLAS_sasum(const int N, const float *X, const int incX); /* reproBLAS.h:103 */
float reproBLAS_snrm2(const int N, const float *X, const int incX); /* reproBLAS.h:105 */
float reproBLAS_sdot(const int N, const float *X, const int incX, const float *Y,
                     const int incY); /* reproBLAS.h:102 */
void reproBLAS_sgemv(const char Order, const char TransA, const int M, const int N, const float alpha,
                     const float *A, const int lda, const float *X, const int incX, const float beta, float *Y,
                     const int incY); /* reproBLAS.h:130 */
void reproBLAS_sgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const float alpha, const float *A, const int lda, const float *B, const int ldb,
                     const float beta, float *C, const int ldc); /* reproBLAS.h:135 */

/* complex: X, Y, A, B, C point to interleaved (real, imaginary) pairs of
 * double (z) or float (c); alpha, beta, and the *_sub results are such pairs.
 * Complex binned accumulators have independent indices per component. */
void reproBLAS_rzsum_sub(const int fold, const int N, const void *X, int incX, void *sum);         /* :37 */
double reproBLAS_rdzasum(const int fold, const int N, const void *X, const int incX);              /* :38 */
double reproBLAS_rdznrm2(const int fold, const int N, const void *X, int incX);                    /* :39 */
void reproBLAS_rzdotc_sub(const int fold, const int N, const void *X, const int incX, const void *Y,
                          const int incY, void *dotc);                                            /* :40 */
void reproBLAS_rzdotu_sub(const int fold, const int N, const void *X, const int incX, const void *Y,
                          const int incY, void *dotu);                                            /* :41 */
void reproBLAS_rcsum_sub(const int fold, const int N, const void *X, const int incX, void *sum);   /* :43 */
float reproBLAS_rscasum(const int fold, const int N, const void *X, const int incX);               /* :44 */
float reproBLAS_rscnrm2(const int fold, const int N, const void *X, const int incX);               /* :45 */
void reproBLAS_rcdotc_sub(const int fold, const int N, const void *X, const int incX, const void *Y,
                          const int incY, void *dotc);                                            /* :46 */
void reproBLAS_rcdotu_sub(const int fold, const int N, const void *X, const int incX, const void *Y,
                          const int incY, void *dotu);                                            /* :47 */
void reproBLAS_rzgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                      const void *alpha, const void *A, const int lda, const void *X, const int incX,
                      const void *beta, void *Y, const int incY); /* reproBLAS.h:73 */
void reproBLAS_rzgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const void *alpha, const void *A, const int lda, const void *B,
                      const int ldb, const void *beta, void *C, const int ldc); /* reproBLAS.h:78 */
void reproBLAS_rcgemv(const int fold, const char Order, const char TransA, const int M, const int N,
                      const void *alpha, const void *A, const int lda, const void *X, const int incX,
                      const void *beta, void *Y, const int incY); /* reproBLAS.h:85 */
void reproBLAS_rcgemm(const int fold, const char Order, const char TransA, const char TransB, const int M,
                      const int N, const int K, const void *alpha, const void *A, const int lda, const void *B,
                      const int ldb, const void *beta, void *C, const int ldc); /* reproBLAS.h:90 */
void reproBLAS_zsum_sub(const int N, const void *X, int incX, void *sum);                          /* :107 */
double reproBLAS_dzasum(const int N, const void *X, const int incX);                               /* :108 */
double reproBLAS_dznrm2(const int N, const void *X, int incX);                                     /* :109 */
void reproBLAS_zdotc_sub(const int N, const void *X, const int incX, const void *Y, const int incY,
                         void *dotc);                                                             /* :110 */
void reproBLAS_zdotu_sub(const int N, const void *X, const int incX, const void *Y, const int incY,
                         void *dotu);                                                             /* :111 */
void reproBLAS_csum_sub(const int N, const void *X, const int incX, void *sum);                    /* :113 */
float reproBLAS_scasum(const int N, const void *X, const int incX);                                /* :114 */
float reproBLAS_scnrm2(const int N, const void *X, const int incX);                                /* :115 */
void reproBLAS_cdotc_sub(const int N, const void *X, const int incX, const void *Y, const int incY,
                         void *dotc);                                                             /* :116 */
void reproBLAS_cdotu_sub(const int N, const void *X, const int incX, const void *Y, const int incY,
                         void *dotu);                                                             /* :117 */
void reproBLAS_zgemv(const char Order, const char TransA, const int M, const int N, const void *alpha,
                     const void *A, const int lda, const void *X, const int incX, const void *beta, void *Y,
                     const int incY); /* reproBLAS.h:141 */
void reproBLAS_zgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const void *alpha, const void *A, const int lda, const void *B, const int ldb,
                     const void *beta, void *C, const int ldc); /* reproBLAS.h:146 */
void reproBLAS_cgemv(const char Order, const char TransA, const int M, const int N, const void *alpha,
                     const void *A, const int lda, const void *X, const int incX, const void *beta, void *Y,
                     const int incY); /* reproBLAS.h:152 */
void reproBLAS_cgemm(const char Order, const char TransA, const char TransB, const int M, const int N, const int K,
                     const void *alpha, const void *A, const int lda, const void *B, const int ldb,
                     const void *beta, void *C, const int ldc); /* reproBLAS.h:157 */

#ifdef __cplusplus
}
#endif
#endif
