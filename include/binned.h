/*
 * binned.h -- the binned double-precision type and its primitives, as
 * exported by libreproblas_hip.so.
 *
 * Drop-in for the double-precision part of the reference's include/binned.h
 * (ReproBLAS 2.1.0): same type, constants, names, signatures and layout.
 * A double_binned of fold F is double[2F]: F primaries followed by F carries.
 *
 * These primitives act on a few dozen bytes and run on the host, on host
 * memory, exactly where the reference runs them; binned_dbdbaddv also accepts
 * device arrays and then runs on the GPU. The bulk reductions are in
 * binnedBLAS.h / reproBLAS.h.
 *
 * Each declaration cites the reference interface it replaces
 * (include/binned.h:line of ReproBLAS 2.1.0).
 */
#ifndef BINNED_H_
#define BINNED_H_
#include <float.h>
#include <stddef.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/binned.h:46 */
typedef double double_binned;

/* reference include/binned.h:84 -- bin width in bits */
#define DBWIDTH 40
/* reference include/binned.h:105 -- maximum index (51) */
#define binned_DBMAXINDEX (((DBL_MAX_EXP - DBL_MIN_EXP + DBL_MANT_DIG - 1) / DBWIDTH) - 1)
/* reference include/binned.h:123 -- maximum fold (52) */
#define binned_DBMAXFOLD (binned_DBMAXINDEX + 1)
/* reference include/binned.h:142 -- deposits between renormalisations (2048) */
#define binned_DBENDURANCE (1 << (DBL_MANT_DIG - DBWIDTH - 2))
/* reference include/binned.h:163 */
#define binned_DBCAPACITY (binned_DBENDURANCE * (1.0 / DBL_EPSILON - 1.0))
/* reference include/binned.h:184,206 */
#define binned_DMCOMPRESSION (1.0 / (1 << (DBL_MANT_DIG - DBWIDTH + 1)))
#define binned_DMEXPANSION (1.0 * (1 << (DBL_MANT_DIG - DBWIDTH + 1)))

size_t binned_dbsize(const int fold);                           /* binned.h:219 */
double_binned *binned_dballoc(const int fold);                  /* binned.h:224 */
int binned_dbnum(const int fold);                               /* binned.h:229 */
double binned_dbbound(const int fold, const int N, const double X, const double S); /* binned.h:234 */
const double *binned_dmbins(const int X);                       /* binned.h:237 */
int binned_dindex(const double X);                              /* binned.h:240 */
int binned_dmindex(const double *priX);                         /* binned.h:241 */
int binned_dmindex0(const double *priX);                        /* binned.h:242 */
int binned_dmdenorm(const int fold, const double *priX);        /* binned.h:248 */
void binned_dbprint(const int fold, const double_binned *X);    /* binned.h:253 */
void binned_dmprint(const int fold, const double *priX, const int incpriX, const double *carX,
                    const int inccarX);                         /* binned.h:254 */
void binned_dbdbset(const int fold, const double_binned *X, double_binned *Y); /* binned.h:262 */
void binned_dmdmset(const int fold, const double *priX, const int incpriX, const double *carX, const int inccarX,
                    double *priY, const int incpriY, double *carY, const int inccarY); /* binned.h:263 */
void binned_dbsetzero(const int fold, double_binned *X);        /* binned.h:275 */
void binned_dmsetzero(const int fold, double *priX, const int incpriX, double *carX,
                      const int inccarX);                       /* binned.h:276 */
void binned_dbdbadd(const int fold, const double_binned *X, double_binned *Y); /* binned.h:284 */
void binned_dmdmadd(const int fold, const double *priX, const int incpriX, const double *carX, const int inccarX,
                    double *priY, const int incpriY, double *carY, const int inccarY); /* binned.h:285 */
/* binned.h:293 -- host or device arrays (both operands must be of one kind) */
void binned_dbdbaddv(const int fold, const int N, const double_binned *X, const int incX, double_binned *Y,
                     const int incY);
void binned_dbdadd(const int fold, const double X, double_binned *Y); /* binned.h:298 */
void binned_dmdadd(const int fold, const double X, double *priY, const int incpriY, double *carY,
                   const int inccarY);                          /* binned.h:299 */
void binned_dbdupdate(const int fold, const double X, double_binned *Y); /* binned.h:307 */
void binned_dmdupdate(const int fold, const double X, double *priY, const int incpriY, double *carY,
                      const int inccarY);                       /* binned.h:308 */
void binned_dbddeposit(const int fold, const double X, double_binned *Y); /* binned.h:320 */
void binned_dmddeposit(const int fold, const double X, double *priY, const int incpriY); /* binned.h:321 */
void binned_dbrenorm(const int fold, double_binned *X);         /* binned.h:329 */
void binned_dmrenorm(const int fold, double *priX, const int incpriX, double *carX,
                     const int inccarX);                        /* binned.h:330 */
void binned_dbdconv(const int fold, const double X, double_binned *Y); /* binned.h:338 */
void binned_dmdconv(const int fold, const double X, double *priY, const int incpriY, double *carY,
                    const int inccarY);                         /* binned.h:339 */
double binned_ddbconv(const int fold, const double_binned *X);  /* binned.h:346 */
double binned_ddmconv(const int fold, const double *priX, const int incpriX, const double *carX,
                      const int inccarX);                       /* binned.h:347 */
void binned_dbnegate(const int fold, double_binned *X);         /* binned.h:355 */
void binned_dmnegate(const int fold, double *priX, const int incpriX, double *carX,
                     const int inccarX);                        /* binned.h:356 */
double binned_dscale(const double X);                           /* binned.h:364 */
void binned_dmdrescale(const int fold, const double X, const double scaleY, double *priY, const int incpriY,
                       double *carY, const int inccarY);        /* binned.h:367 */
double binned_dmdmaddsq(const int fold, const double scaleX, const double *priX, const int incpriX,
                        const double *carX, const int inccarX, const double scaleY, double *priY,
                        const int incpriY, double *carY, const int inccarY); /* binned.h:372 */
double binned_dbdbaddsq(const int fold, const double scaleX, const double_binned *X, const double scaleY,
                        double_binned *Y);                      /* binned.h:373 */
double binned_ufp(const double X);                              /* binned.h:377 */

#ifdef __cplusplus
}
#endif
#endif
