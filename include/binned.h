/*
 * binned.h -- the binned double- and single-precision types and their
 * primitives, as exported by libreproblas_hip.so.
 *
 * Drop-in for the real (double and float) part of the reference's
 * include/binned.h (ReproBLAS 2.1.0): same types, constants, names,
 * signatures and layout. A double_binned of fold F is double[2F], a
 * float_binned float[2F]: F primaries followed by F carries.
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
/* reference include/binned.h:55 -- double[4F]: F interleaved (re, im)
 * primaries, then F interleaved carries */
typedef double double_complex_binned;

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

/* ------------------------------------------------------------ float_binned
 * Single precision: float[2F] (F primaries, then F carries), folds 2..21. */
/* reference include/binned.h:64 */
typedef float float_binned;
/* reference include/binned.h:73 -- float[4F], laid out as double_complex_binned */
typedef float float_complex_binned;
/* reference include/binned.h:95 -- bin width in bits */
#define SBWIDTH 13
/* reference include/binned.h:115 -- maximum index (20) */
#define binned_SBMAXINDEX (((FLT_MAX_EXP - FLT_MIN_EXP + FLT_MANT_DIG - 1) / SBWIDTH) - 1)
/* reference include/binned.h:131 -- maximum fold (21) */
#define binned_SBMAXFOLD (binned_SBMAXINDEX + 1)
/* reference include/binned.h:153 -- deposits between renormalisations (512) */
#define binned_SBENDURANCE (1 << (FLT_MANT_DIG - SBWIDTH - 2))
/* reference include/binned.h:173 */
#define binned_SBCAPACITY (binned_SBENDURANCE * (1.0 / FLT_EPSILON - 1.0))
/* reference include/binned.h:195,217 */
#define binned_SMCOMPRESSION (1.0 / (1 << (FLT_MANT_DIG - SBWIDTH + 1)))
#define binned_SMEXPANSION (1.0 * (1 << (FLT_MANT_DIG - SBWIDTH + 1)))

size_t binned_sbsbze(const int fold);                           /* binned.h:221 */
float_binned *binned_sballoc(const int fold);                   /* binned.h:226 */
int binned_sbnum(const int fold);                               /* binned.h:231 */
float binned_sbbound(const int fold, const int N, const float X, const float S); /* binned.h:235 */
const float *binned_smbins(const int X);                        /* binned.h:238 */
int binned_sindex(const float X);                               /* binned.h:244 */
int binned_smindex(const float *priX);                          /* binned.h:245 */
int binned_smindex0(const float *priX);                         /* binned.h:246 */
int binned_smdenorm(const int fold, const float *priX);         /* binned.h:250 */
void binned_sbprint(const int fold, const float_binned *X);     /* binned.h:257 */
void binned_smprint(const int fold, const float *priX, const int incpriX, const float *carX,
                    const int inccarX);                         /* binned.h:258 */
void binned_sbsbset(const int fold, const float_binned *X, float_binned *Y); /* binned.h:268 */
void binned_smsmset(const int fold, const float *priX, const int incpriX, const float *carX, const int inccarX,
                    float *priY, const int incpriY, float *carY, const int inccarY); /* binned.h:269 */
void binned_sbsetzero(const int fold, float_binned *X);         /* binned.h:279 */
void binned_smsetzero(const int fold, float *priX, const int incpriX, float *carX,
                      const int inccarX);                       /* binned.h:280 */
void binned_sbsbadd(const int fold, const float_binned *X, float_binned *Y); /* binned.h:288 */
void binned_smsmadd(const int fold, const float *priX, const int incpriX, const float *carX, const int inccarX,
                    float *priY, const int incpriY, float *carY, const int inccarY); /* binned.h:289 */
/* binned.h:295 -- host or device arrays (both operands must be of one kind) */
void binned_sbsbaddv(const int fold, const int N, const float_binned *X, const int incX, float_binned *Y,
                     const int incY);
void binned_sbsadd(const int fold, const float X, float_binned *Y); /* binned.h:302 */
void binned_smsadd(const int fold, const float X, float *priY, const int incpriY, float *carY,
                   const int inccarY);                          /* binned.h:303 */
void binned_sbsupdate(const int fold, const float X, float_binned *Y); /* binned.h:313 */
void binned_smsupdate(const int fold, const float X, float *priY, const int incpriY, float *carY,
                      const int inccarY);                       /* binned.h:314 */
void binned_sbsdeposit(const int fold, const float X, float_binned *Y); /* binned.h:324 */
void binned_smsdeposit(const int fold, const float X, float *priY, const int incpriY); /* binned.h:325 */
void binned_sbrenorm(const int fold, float_binned *X);          /* binned.h:333 */
void binned_smrenorm(const int fold, float *priX, const int incpriX, float *carX,
                     const int inccarX);                        /* binned.h:334 */
void binned_sbsconv(const int fold, const float X, float_binned *Y); /* binned.h:342 */
void binned_smsconv(const int fold, const float X, float *priY, const int incpriY, float *carY,
                    const int inccarY);                         /* binned.h:343 */
float binned_ssbconv(const int fold, const float_binned *X);    /* binned.h:350 */
float binned_ssmconv(const int fold, const float *priX, const int incpriX, const float *carX,
                     const int inccarX);                        /* binned.h:351 */
void binned_sbnegate(const int fold, float_binned *X);          /* binned.h:359 */
void binned_smnegate(const int fold, float *priX, const int incpriX, float *carX,
                     const int inccarX);                        /* binned.h:360 */
float binned_sscale(const float X);                             /* binned.h:365 */
void binned_smsrescale(const int fold, const float X, const float scaleY, float *priY, const int incpriY,
                       float *carY, const int inccarY);         /* binned.h:369 */
float binned_smsmaddsq(const int fold, const float scaleX, const float *priX, const int incpriX,
                       const float *carX, const int inccarX, const float scaleY, float *priY,
                       const int incpriY, float *carY, const int inccarY); /* binned.h:374 */
float binned_sbsbaddsq(const int fold, const float scaleX, const float_binned *X, const float scaleY,
                       float_binned *Y);                        /* binned.h:375 */
float binned_ufpf(const float X);                               /* binned.h:378 */

/* ------------------------------------------------------------ complex
 * Each complex primitive applies the real one to both components (strides
 * doubled), as the reference's do. X of the *add/*update/*deposit/*conv
 * calls points to an interleaved (re, im) pair. */
size_t binned_zbsize(const int fold);                               /* binned.h:220 */
size_t binned_cbsize(const int fold);                               /* binned.h:222 */
double_complex_binned *binned_zballoc(const int fold);              /* binned.h:225 */
float_complex_binned *binned_cballoc(const int fold);               /* binned.h:227 */
int binned_zbnum(const int fold);                                   /* binned.h:230 */
int binned_cbnum(const int fold);                                   /* binned.h:232 */
int binned_zmdenorm(const int fold, const double *priX);            /* binned.h:249 */
int binned_cmdenorm(const int fold, const float *priX);             /* binned.h:251 */
void binned_zbprint(const int fold, const double_complex_binned *X); /* binned.h:255 */
void binned_zmprint(const int fold, const double *priX, const int incpriX, const double *carX,
                    const int inccarX);                             /* binned.h:256 */
void binned_cbprint(const int fold, const float_complex_binned *X); /* binned.h:259 */
void binned_cmprint(const int fold, const float *priX, const int incpriX, const float *carX,
                    const int inccarX);                             /* binned.h:260 */
void binned_zbzbset(const int fold, const double_complex_binned *X, double_complex_binned *Y); /* binned.h:264 */
void binned_zmzmset(const int fold, const double *priX, const int incpriX, const double *carX, const int inccarX,
                    double *priY, const int incpriY, double *carY, const int inccarY); /* binned.h:265 */
void binned_zbdbset(const int fold, const double_binned *X, double_complex_binned *Y); /* binned.h:266 */
void binned_zmdmset(const int fold, const double *priX, const int incpriX, const double *carX, const int inccarX,
                    double *priY, const int incpriY, double *carY, const int inccarY); /* binned.h:267 */
void binned_cbcbset(const int fold, const float_complex_binned *X, float_complex_binned *Y); /* binned.h:270 */
void binned_cmcmset(const int fold, const float *priX, const int incpriX, const float *carX, const int inccarX,
                    float *priY, const int incpriY, float *carY, const int inccarY); /* binned.h:271 */
void binned_cbsbset(const int fold, const float_binned *X, float_complex_binned *Y); /* binned.h:272 */
void binned_cmsmset(const int fold, const float *priX, const int incpriX, const float *carX, const int inccarX,
                    float *priY, const int incpriY, float *carY, const int inccarY); /* binned.h:273 */
void binned_zbsetzero(const int fold, double_complex_binned *X);    /* binned.h:277 */
void binned_zmsetzero(const int fold, double *priX, const int incpriX, double *carX,
                      const int inccarX);                           /* binned.h:278 */
void binned_cbsetzero(const int fold, float_complex_binned *X);     /* binned.h:281 */
void binned_cmsetzero(const int fold, float *priX, const int incpriX, float *carX,
                      const int inccarX);                           /* binned.h:282 */
void binned_zbzbadd(const int fold, const double_complex_binned *X, double_complex_binned *Y); /* binned.h:286 */
void binned_zmzmadd(const int fold, const double *priX, const int incpriX, const double *carX, const int inccarX,
                    double *priY, const int incpriY, double *carY, const int inccarY); /* binned.h:287 */
void binned_cbcbadd(const int fold, const float_complex_binned *X, float_complex_binned *Y); /* binned.h:290 */
void binned_cmcmadd(const int fold, const float *priX, const int incpriX, const float *carX, const int inccarX,
                    float *priY, const int incpriY, float *carY, const int inccarY); /* binned.h:291 */
/* binned.h:294,296 -- host or device arrays (both operands of one kind) */
void binned_zbzbaddv(const int fold, const int N, const double_complex_binned *X, const int incX,
                     double_complex_binned *Y, const int incY);
void binned_cbcbaddv(const int fold, const int N, const float_complex_binned *X, const int incX,
                     float_complex_binned *Y, const int incY);
void binned_zbzadd(const int fold, const void *X, double_complex_binned *Y); /* binned.h:300 */
void binned_zmzadd(const int fold, const void *X, double *priY, const int incpriY, double *carY,
                   const int inccarY);                              /* binned.h:301 */
void binned_cbcadd(const int fold, const void *X, float_complex_binned *Y); /* binned.h:304 */
void binned_cmcadd(const int fold, const void *X, float *priY, const int incpriY, float *carY,
                   const int inccarY);                              /* binned.h:305 */
void binned_zbzupdate(const int fold, const void *X, double_complex_binned *Y); /* binned.h:309 */
void binned_zmzupdate(const int fold, const void *X, double *priY, const int incpriY, double *carY,
                      const int inccarY);                           /* binned.h:310 */
void binned_zbdupdate(const int fold, const double X, double_complex_binned *Y); /* binned.h:311 */
void binned_zmdupdate(const int fold, const double X, double *priY, const int incpriY, double *carY,
                      const int inccarY);                           /* binned.h:312 */
void binned_cbcupdate(const int fold, const void *X, float_complex_binned *Y); /* binned.h:315 */
void binned_cmcupdate(const int fold, const void *X, float *priY, const int incpriY, float *carY,
                      const int inccarY);                           /* binned.h:316 */
void binned_cbsupdate(const int fold, const float X, float_complex_binned *Y); /* binned.h:317 */
void binned_cmsupdate(const int fold, const float X, float *priY, const int incpriY, float *carY,
                      const int inccarY);                           /* binned.h:318 */
void binned_zbzdeposit(const int fold, const void *X, double_complex_binned *Y); /* binned.h:322 */
void binned_zmzdeposit(const int fold, const void *X, double *priY, const int incpriY); /* binned.h:323 */
void binned_cbcdeposit(const int fold, const void *X, float_complex_binned *Y); /* binned.h:326 */
void binned_cmcdeposit(const int fold, const void *X, float *priY, const int incpriY); /* binned.h:327 */
void binned_zbrenorm(const int fold, double_complex_binned *X);     /* binned.h:331 */
void binned_zmrenorm(const int fold, double *priX, const int incpriX, double *carX,
                     const int inccarX);                            /* binned.h:332 */
void binned_cbrenorm(const int fold, float_complex_binned *X);      /* binned.h:335 */
void binned_cmrenorm(const int fold, float *priX, const int incpriX, float *carX,
                     const int inccarX);                            /* binned.h:336 */
void binned_zbzconv(const int fold, const void *X, double_complex_binned *Y); /* binned.h:340 */
void binned_zmzconv(const int fold, const void *X, double *priY, const int incpriY, double *carY,
                    const int inccarY);                             /* binned.h:341 */
void binned_cbcconv(const int fold, const void *X, float_complex_binned *Y); /* binned.h:344 */
void binned_cmcconv(const int fold, const void *X, float *priY, const int incpriY, float *carY,
                    const int inccarY);                             /* binned.h:345 */
void binned_zzbconv_sub(const int fold, const double_complex_binned *X, void *conv); /* binned.h:348 */
void binned_zzmconv_sub(const int fold, const double *priX, const int incpriX, const double *carX,
                        const int inccarX, void *conv);             /* binned.h:349 */
void binned_ccbconv_sub(const int fold, const float_complex_binned *X, void *conv); /* binned.h:352 */
void binned_ccmconv_sub(const int fold, const float *priX, const int incpriX, const float *carX,
                        const int inccarX, void *conv);             /* binned.h:353 */
void binned_zbnegate(const int fold, double_complex_binned *X);     /* binned.h:357 */
void binned_zmnegate(const int fold, double *priX, const int incpriX, double *carX,
                     const int inccarX);                            /* binned.h:358 */
void binned_cbnegate(const int fold, float_complex_binned *X);      /* binned.h:361 */
void binned_cmnegate(const int fold, float *priX, const int incpriX, float *carX,
                     const int inccarX);                            /* binned.h:362 */
void binned_zmdrescale(const int fold, const double X, const double scaleY, double *priY, const int incpriY,
                       double *carY, const int inccarY);            /* binned.h:368 */
void binned_cmsrescale(const int fold, const float X, const float scaleY, float *priY, const int incpriY,
                       float *carY, const int inccarY);             /* binned.h:370 */

#ifdef __cplusplus
}
#endif
#endif
