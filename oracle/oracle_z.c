/*
 * oracle_z.c -- CPU restatement of ReproBLAS 2.1.0's complex binned path:
 * double complex ("z": zbz*, dbz*, rz*, rdz*) and float complex ("c":
 * cbc*, sbc*, rc*, rsc*). TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The code is oracle_cplx.inc, instantiated once per precision on top of
 * the real primitives of oracle.c / oracle_s.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* -------------------------------------------------------- double complex */
#define R double
#define ZF(name) zd_##name
#define R_UPDATE orc_dmdupdate
#define R_DEPOSIT orc_dmddeposit
#define R_RENORM orc_dmrenorm
#define R_MADD orc_dmdmadd
#define R_ADD orc_dmdadd
#define R_CONV orc_ddmconv
#define R_RESCALE orc_dmdrescale
#define R_SCALE orc_dscale
#define R_ENDURANCE ORC_ENDURANCE
#define R_MAXFOLD ORC_MAXFOLD
#define R_NRM2_FINISH(scl, v) ((scl) * sqrt(v))
#define NB_SUM 2048 /* zmzsum.ccog:36 */
#define NB_DOT 2048 /* zmzdotu.ccog:36, zmzdotc.ccog:36 */
#define NB_ABS 1024 /* dmzasum.ccog:36, dmzssq.ccog:36 */
#define Z_BZBADD orc_zbzbadd
#define Z_BZCONV orc_zbzconv
#define Z_ZBCONV orc_zzbconv
#define Z_BZSUM orc_zbzsum
#define Z_BZDOTU orc_zbzdotu
#define Z_BZDOTC orc_zbzdotc
#define R_BZASUM orc_dbzasum
#define R_BZSSQ orc_dbzssq
#define Z_RSUM orc_rzsum
#define Z_RDOTU orc_rzdotu
#define Z_RDOTC orc_rzdotc
#define R_RASUM orc_rdzasum
#define R_RNRM2 orc_rdznrm2
#define Z_BZGEMV orc_zbzgemv
#define Z_RGEMV orc_rzgemv
#define Z_BZGEMM orc_zbzgemm
#define Z_RGEMM orc_rzgemm
#define Z_AMAX orc_zamax
#define Z_AMAXM orc_zamaxm
#include "oracle_cplx.inc"
#undef R
#undef ZF
#undef R_UPDATE
#undef R_DEPOSIT
#undef R_RENORM
#undef R_MADD
#undef R_ADD
#undef R_CONV
#undef R_RESCALE
#undef R_SCALE
#undef R_ENDURANCE
#undef R_MAXFOLD
#undef R_NRM2_FINISH
#undef NB_SUM
#undef NB_DOT
#undef NB_ABS
#undef Z_BZBADD
#undef Z_BZCONV
#undef Z_ZBCONV
#undef Z_BZSUM
#undef Z_BZDOTU
#undef Z_BZDOTC
#undef R_BZASUM
#undef R_BZSSQ
#undef Z_RSUM
#undef Z_RDOTU
#undef Z_RDOTC
#undef R_RASUM
#undef R_RNRM2
#undef Z_BZGEMV
#undef Z_RGEMV
#undef Z_BZGEMM
#undef Z_RGEMM
#undef Z_AMAX
#undef Z_AMAXM

/* --------------------------------------------------------- float complex */
#define R float
#define ZF(name) zs_##name
#define R_UPDATE orc_smsupdate
#define R_DEPOSIT orc_smsdeposit
#define R_RENORM orc_smrenorm
#define R_MADD orc_smsmadd
#define R_ADD orc_smsadd
#define R_CONV orc_ssmconv
#define R_RESCALE orc_smsrescale
#define R_SCALE orc_sscale
#define R_ENDURANCE ORC_SENDURANCE
#define R_MAXFOLD ORC_SMAXFOLD
/* reference src/reproBLAS/rscnrm2.c: float scl times double sqrt, rounded */
#define R_NRM2_FINISH(scl, v) ((float)((double)(scl) * sqrt((double)(v))))
#define NB_SUM 512 /* cmcsum.ccog:36 */
#define NB_DOT 512 /* cmcdotu.ccog:36, cmcdotc.ccog:36 */
#define NB_ABS 256 /* smcasum.ccog:36, smcssq.ccog:36 */
#define Z_BZBADD orc_cbcbadd
#define Z_BZCONV orc_cbcconv
#define Z_ZBCONV orc_ccbconv
#define Z_BZSUM orc_cbcsum
#define Z_BZDOTU orc_cbcdotu
#define Z_BZDOTC orc_cbcdotc
#define R_BZASUM orc_sbcasum
#define R_BZSSQ orc_sbcssq
#define Z_RSUM orc_rcsum
#define Z_RDOTU orc_rcdotu
#define Z_RDOTC orc_rcdotc
#define R_RASUM orc_rscasum
#define R_RNRM2 orc_rscnrm2
#define Z_BZGEMV orc_cbcgemv
#define Z_RGEMV orc_rcgemv
#define Z_BZGEMM orc_cbcgemm
#define Z_RGEMM orc_rcgemm
#define Z_AMAX orc_camax
#define Z_AMAXM orc_camaxm
#include "oracle_cplx.inc"
