/*
 * oracle.h -- CPU restatement of ReproBLAS 2.1.0's binned FP64 reduction path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness oracle for the
 * HIP library in reproblas_amd/. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the oracle or as the
 * timed CPU baseline. The product never links, loads or calls it.
 *
 * Every symbol is prefixed orc_ so it can never be confused with (or
 * interpose on) the product's reproBLAS_/binnedBLAS_/binned_ symbols.
 *
 * A binned value of fold F is double[2F]: F primaries, then F carries
 * (include/binned.h:46 and src/binned/dbdbadd.c:16-18 of the reference).
 */
#ifndef REPROBLAS_ORACLE_H_
#define REPROBLAS_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* constants (reference include/binned.h:84,105,123,142,184,206) */
#define ORC_DBWIDTH 40
#define ORC_MAXINDEX 51
#define ORC_MAXFOLD 52
#define ORC_ENDURANCE 2048
/* single precision (reference include/binned.h:95,115,131,153) */
#define ORC_SBWIDTH 13
#define ORC_SMAXINDEX 20
#define ORC_SMAXFOLD 21
#define ORC_SENDURANCE 512

/* primitives (reference src/binned/) */
const double *orc_dmbins(int X);
int orc_dindex(double X);
int orc_dmindex(const double *priX);
int orc_dmindex0(const double *priX);
void orc_dmdupdate(int fold, double X, double *priY, int incpriY, double *carY, int inccarY);
void orc_dmddeposit(int fold, double X, double *priY, int incpriY);
void orc_dmrenorm(int fold, double *priX, int incpriX, double *carX, int inccarX);
void orc_dmdmadd(int fold, const double *priX, int incpriX, const double *carX, int inccarX,
                 double *priY, int incpriY, double *carY, int inccarY);
void orc_dbdbadd(int fold, const double *X, double *Y);
void orc_dmdadd(int fold, double X, double *priY, int incpriY, double *carY, int inccarY);
void orc_dbdadd(int fold, double X, double *Y);
void orc_dbdconv(int fold, double X, double *Y);
double orc_ddmconv(int fold, const double *priX, int incpriX, const double *carX, int inccarX);
double orc_ddbconv(int fold, const double *X);
double orc_dscale(double X);
void orc_dmdrescale(int fold, double X, double scaleY, double *priY, int incpriY, double *carY, int inccarY);
double orc_dbdbaddsq(int fold, double scaleX, const double *X, double scaleY, double *Y);
double orc_dbbound(int fold, int N, double X, double S);

/* binned BLAS (reference src/binnedBLAS/) */
double orc_damax(int N, const double *X, int incX);
double orc_damaxm(int N, const double *X, int incX, const double *Y, int incY);
void orc_dbdsum(int fold, int N, const double *X, int incX, double *Y);
void orc_dbdasum(int fold, int N, const double *X, int incX, double *Y);
void orc_dbddot(int fold, int N, const double *X, int incX, const double *Y, int incY, double *Z);
double orc_dbdssq(int fold, int N, const double *X, int incX, double scaleY, double *Y);

/* user API (reference src/reproBLAS/) */
double orc_rdsum(int fold, int N, const double *X, int incX);
double orc_rdasum(int fold, int N, const double *X, int incX);
double orc_rddot(int fold, int N, const double *X, int incX, const double *Y, int incY);
double orc_rdnrm2(int fold, int N, const double *X, int incX);
void orc_rdgemv(int fold, char Order, char TransA, int M, int N, double alpha,
                const double *A, int lda, const double *X, int incX,
                double beta, double *Y, int incY);
void orc_rdgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K,
                double alpha, const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc);

/* binned-state variants of gemv/gemm: accumulate into caller binned arrays
 * (reference binnedBLAS_dbdgemv / binnedBLAS_dbdgemm semantics). */
void orc_dbdgemv(int fold, char Order, char TransA, int M, int N, double alpha,
                 const double *A, int lda, const double *X, int incX, double *Y, int incY);
void orc_dbdgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K,
                 double alpha, const double *A, int lda, const double *B, int ldb,
                 double *C, int ldc);

/* ------------------------------------------------ single precision (oracle_s.c) */
const float *orc_smbins(int X);
int orc_sindex(float X);
int orc_smindex(const float *priX);
int orc_smindex0(const float *priX);
void orc_smsupdate(int fold, float X, float *priY, int incpriY, float *carY, int inccarY);
void orc_smsdeposit(int fold, float X, float *priY, int incpriY);
void orc_smrenorm(int fold, float *priX, int incpriX, float *carX, int inccarX);
void orc_smsmadd(int fold, const float *priX, int incpriX, const float *carX, int inccarX,
                 float *priY, int incpriY, float *carY, int inccarY);
void orc_sbsbadd(int fold, const float *X, float *Y);
void orc_smsadd(int fold, float X, float *priY, int incpriY, float *carY, int inccarY);
void orc_sbsadd(int fold, float X, float *Y);
void orc_sbsconv(int fold, float X, float *Y);
float orc_ssmconv(int fold, const float *priX, int incpriX, const float *carX, int inccarX);
float orc_ssbconv(int fold, const float *X);
float orc_sscale(float X);
void orc_smsrescale(int fold, float X, float scaleY, float *priY, int incpriY, float *carY, int inccarY);
float orc_sbsbaddsq(int fold, float scaleX, const float *X, float scaleY, float *Y);
float orc_sbbound(int fold, int N, float X, float S);
float orc_samax(int N, const float *X, int incX);
float orc_samaxm(int N, const float *X, int incX, const float *Y, int incY);
void orc_sbssum(int fold, int N, const float *X, int incX, float *Y);
void orc_sbsasum(int fold, int N, const float *X, int incX, float *Y);
void orc_sbsdot(int fold, int N, const float *X, int incX, const float *Y, int incY, float *Z);
float orc_sbsssq(int fold, int N, const float *X, int incX, float scaleY, float *Y);
float orc_rssum(int fold, int N, const float *X, int incX);
float orc_rsasum(int fold, int N, const float *X, int incX);
float orc_rsdot(int fold, int N, const float *X, int incX, const float *Y, int incY);
float orc_rsnrm2(int fold, int N, const float *X, int incX);
void orc_sbsgemv(int fold, char Order, char TransA, int M, int N, float alpha, const float *A, int lda,
                 const float *X, int incX, float *Y, int incY);
void orc_rsgemv(int fold, char Order, char TransA, int M, int N, float alpha, const float *A, int lda,
                const float *X, int incX, float beta, float *Y, int incY);
void orc_sbsgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K, float alpha, const float *A,
                 int lda, const float *B, int ldb, float *C, int ldc);
void orc_rsgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K, float alpha, const float *A,
                int lda, const float *B, int ldb, float beta, float *C, int ldc);


/* ------------------------------------------------ complex (oracle_z.c) ----
 * Complex binned values of fold F are R[4F]: F interleaved (re, im)
 * primaries, then F interleaved carries. Complex scalars are R[2]. */
#define ORC_ZDECL(R, z, d, rz, rd)                                                                            \
  void orc_##z##b##z##badd(int fold, const R *X, R *Y);                                                     \
  void orc_##z##b##z##conv(int fold, const R *X, R *Y);                                                     \
  void orc_##z##z##bconv(int fold, const R *X, R *out);                                                     \
  void orc_##z##b##z##sum(int fold, int N, const R *X, int incX, R *Y);                                     \
  void orc_##z##b##z##dotu(int fold, int N, const R *X, int incX, const R *Y, int incY, R *Z);              \
  void orc_##z##b##z##dotc(int fold, int N, const R *X, int incX, const R *Y, int incY, R *Z);              \
  void orc_##d##b##z##asum(int fold, int N, const R *X, int incX, R *Y);                                    \
  R orc_##d##b##z##ssq(int fold, int N, const R *X, int incX, R scaleY, R *Y);                              \
  void orc_r##z##sum(int fold, int N, const R *X, int incX, R *sum);                                        \
  void orc_r##z##dotu(int fold, int N, const R *X, int incX, const R *Y, int incY, R *dot);                 \
  void orc_r##z##dotc(int fold, int N, const R *X, int incX, const R *Y, int incY, R *dot);                 \
  R orc_##rd##asum(int fold, int N, const R *X, int incX);                                                  \
  R orc_##rd##nrm2(int fold, int N, const R *X, int incX);                                                  \
  void orc_##z##b##z##gemv(int fold, char Order, char TransA, int M, int N, const R *alpha, const R *A,     \
                           int lda, const R *X, int incX, R *Y, int incY);                                  \
  void orc_r##z##gemv(int fold, char Order, char TransA, int M, int N, const R *alpha, const R *A, int lda, \
                      const R *X, int incX, const R *beta, R *Y, int incY);                                 \
  void orc_##z##b##z##gemm(int fold, char Order, char TransA, char TransB, int M, int N, int K,             \
                           const R *alpha, const R *A, int lda, const R *B, int ldb, R *C, int ldc);        \
  void orc_r##z##gemm(int fold, char Order, char TransA, char TransB, int M, int N, int K, const R *alpha,  \
                      const R *A, int lda, const R *B, int ldb, const R *beta, R *C, int ldc);              \
  void orc_##z##amax(int N, const R *X, int incX, R *amax);                                                 \
  void orc_##z##amaxm(int N, const R *X, int incX, const R *Y, int incY, R *amaxm);
ORC_ZDECL(double, z, d, rz, rdz)
ORC_ZDECL(float, c, s, rc, rsc)
#undef ORC_ZDECL

#ifdef __cplusplus
}
#endif
#endif
