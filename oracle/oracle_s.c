/*
 * oracle_s.c -- scalar CPU restatement of ReproBLAS 2.1.0's binned FP32 path
 * (float_binned = float[2*fold]: primaries then carries).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Same rules as oracle.c: compiled
 * with -ffp-contract=off, each function names the reference file:line it
 * restates, deposits follow the scalar (SISD) sections of the reference's
 * generated kernels. Float arithmetic stays float wherever the reference's
 * does; conversion accumulates in double as the reference's ssmconv does.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define MINI(a, b) ((a) < (b) ? (a) : (b))

/* 2^-(24-13+1) and 2^(24-13+1): reference include/binned.h:195,217 */
static const float SCOMPRESSION = 1.0f / (1 << 12);
static const float SEXPANSION = 1.0f * (1 << 12);

static inline uint32_t fbits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float ffrom(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
/* biased exponent field: reference src/common/common.h:51-55 */
static inline int fexp(float x) { return (int)((fbits(x) >> 23) & 0xff); }
/* exponent all ones: reference src/common/common.h:65-69 */
static inline int fnan_or_inf(float x) { return fexp(x) == 0xff; }
static inline float fset_lsb(float x) { return ffrom(fbits(x) | 1u); }

/* ---------------------------------------------------------------- bins --- */
/* reference src/binned/smbins.c: bins[0] = 0.75*2^128, bins[i] =
 * 0.75*2^(128+24-13+1-13i) for 1 <= i <= 20, then bins[20] repeated. */
static float sbin_table[ORC_SMAXINDEX + ORC_SMAXFOLD];
static int sbin_ready = 0;

const float *orc_smbins(int X) {
  if (!sbin_ready) {
    int i;
    sbin_table[0] = ldexpf(0.75f, 128);
    for (i = 1; i <= ORC_SMAXINDEX; i++) sbin_table[i] = ldexpf(0.75f, 128 + 24 - ORC_SBWIDTH + 1 - i * ORC_SBWIDTH);
    for (; i < ORC_SMAXINDEX + ORC_SMAXFOLD; i++) sbin_table[i] = sbin_table[ORC_SMAXINDEX];
    sbin_ready = 1;
  }
  return sbin_table + X;
}

/* --------------------------------------------------------------- index --- */
/* reference src/binned/sindex.c (active version); EXPF_BIAS = 126 */
int orc_sindex(float X) {
  int e = fexp(X);
  if (e == 0) {
    if (X == 0.0f) return ORC_SMAXINDEX;
    frexpf(X, &e);
    return MINI((128 - e) / ORC_SBWIDTH, ORC_SMAXINDEX);
  }
  return ((128 + 126) - e) / ORC_SBWIDTH;
}

/* reference src/binned/smindex.c */
int orc_smindex(const float *priX) { return ((128 + 24 - ORC_SBWIDTH + 1 + 126) - fexp(priX[0])) / ORC_SBWIDTH; }

/* reference src/binned/smindex0.c */
int orc_smindex0(const float *priX) { return fexp(priX[0]) == 128 + 126; }

/* -------------------------------------------------------------- update --- */
/* reference src/binned/smsupdate.c */
void orc_smsupdate(int fold, float X, float *priY, int incpriY, float *carY, int inccarY) {
  int i, shift, xi;
  const float *b;
  if (fnan_or_inf(priY[0])) return;
  xi = orc_sindex(X);
  if (priY[0] == 0.0f) {
    b = orc_smbins(xi);
    for (i = 0; i < fold; i++) { priY[i * incpriY] = b[i]; carY[i * inccarY] = 0.0f; }
    return;
  }
  shift = orc_smindex(priY) - xi;
  if (shift <= 0) return;
  for (i = fold - 1; i >= shift; i--) {
    priY[i * incpriY] = priY[(i - shift) * incpriY];
    carY[i * inccarY] = carY[(i - shift) * inccarY];
  }
  b = orc_smbins(xi);
  for (i = 0; i < MINI(shift, fold); i++) { priY[i * incpriY] = b[i]; carY[i * inccarY] = 0.0f; }
}

/* ------------------------------------------------------------- deposit --- */
/* reference src/binned/smsdeposit.c and the scalar loops of
 * src/binnedBLAS/smssum.ccog (no NaN/Inf test inside the loops). */
static inline void sdeposit_one(int fold, int top, float x, float *s, int inc) {
  int j = 0;
  float q, m;
  if (top) {
    m = s[0];
    q = m + fset_lsb(x * SCOMPRESSION);
    s[0] = q;
    m = (m - q) * (SEXPANSION * 0.5f);
    x = (x + m) + m;
    j = 1;
  }
  for (; j < fold - 1; j++) {
    m = s[j * inc];
    q = m + fset_lsb(x);
    s[j * inc] = q;
    x = x + (m - q);
  }
  s[j * inc] = s[j * inc] + fset_lsb(x);
}

/* reference src/binned/smsdeposit.c (with its NaN/Inf short cut) */
void orc_smsdeposit(int fold, float X, float *priY, int incpriY) {
  if (fnan_or_inf(X) || fnan_or_inf(priY[0])) { priY[0] += X; return; }
  sdeposit_one(fold, orc_smindex0(priY), X, priY, incpriY);
}

/* -------------------------------------------------------------- renorm --- */
/* reference src/binned/smrenorm.c (active version): bits 21 and 22 */
void orc_smrenorm(int fold, float *priX, int incpriX, float *carX, int inccarX) {
  int i;
  if (priX[0] == 0.0f || fnan_or_inf(priX[0])) return;
  for (i = 0; i < fold; i++) {
    uint32_t u = fbits(priX[i * incpriX]);
    carX[i * inccarX] += (float)((int)((u >> 21) & 3) - 2);
    u &= ~(1u << 21);
    u |= 1u << 22;
    priX[i * incpriX] = ffrom(u);
  }
}

/* ----------------------------------------------------------------- add --- */
/* reference src/binned/smsmadd.c */
void orc_smsmadd(int fold, const float *priX, int incpriX, const float *carX, int inccarX,
                 float *priY, int incpriY, float *carY, int inccarY) {
  int i, shift, xi, yi;
  const float *b;
  if (priX[0] == 0.0f) return;
  if (priY[0] == 0.0f) {
    for (i = 0; i < fold; i++) { priY[i * incpriY] = priX[i * incpriX]; carY[i * inccarY] = carX[i * inccarX]; }
    return;
  }
  if (fnan_or_inf(priX[0]) || fnan_or_inf(priY[0])) { priY[0] += priX[0]; return; }
  xi = orc_smindex(priX);
  yi = orc_smindex(priY);
  shift = yi - xi;
  if (shift > 0) {
    b = orc_smbins(yi);
    for (i = fold - 1; i >= shift; i--) {
      priY[i * incpriY] = priX[i * incpriX] + (priY[(i - shift) * incpriY] - b[i - shift]);
      carY[i * inccarY] = carX[i * inccarX] + carY[(i - shift) * inccarY];
    }
    for (i = 0; i < shift && i < fold; i++) { priY[i * incpriY] = priX[i * incpriX]; carY[i * inccarY] = carX[i * inccarX]; }
  } else {
    b = orc_smbins(xi);
    for (i = -shift; i < fold; i++) {
      priY[i * incpriY] += priX[(i + shift) * incpriX] - b[i + shift];
      carY[i * inccarY] += carX[(i + shift) * inccarX];
    }
  }
  orc_smrenorm(fold, priY, incpriY, carY, inccarY);
}

/* reference src/binned/sbsbadd.c */
void orc_sbsbadd(int fold, const float *X, float *Y) { orc_smsmadd(fold, X, 1, X + fold, 1, Y, 1, Y + fold, 1); }

/* reference src/binned/smsadd.c, sbsadd.c */
void orc_smsadd(int fold, float X, float *priY, int incpriY, float *carY, int inccarY) {
  orc_smsupdate(fold, X, priY, incpriY, carY, inccarY);
  orc_smsdeposit(fold, X, priY, incpriY);
  orc_smrenorm(fold, priY, incpriY, carY, inccarY);
}
void orc_sbsadd(int fold, float X, float *Y) { orc_smsadd(fold, X, Y, 1, Y + fold, 1); }

/* reference src/binned/smsconv.c, sbsconv.c */
void orc_sbsconv(int fold, float X, float *Y) {
  memset(Y, 0, 2 * (size_t)fold * sizeof(float));
  orc_smsadd(fold, X, Y, 1, Y + fold, 1);
}

/* ------------------------------------------------------------- convert --- */
/* reference src/binned/ssmconv.c: accumulated in double in a fixed order,
 * rounded to float once at the end. */
float orc_ssmconv(int fold, const float *priX, int incpriX, const float *carX, int inccarX) {
  int i;
  const float *b;
  double Y = 0.0;
  if (fnan_or_inf(priX[0])) return priX[0];
  if (priX[0] == 0.0f) return 0.0f;
  b = orc_smbins(orc_smindex(priX));
  if (orc_smindex(priX) == 0) {
    Y += (double)carX[0] * (double)(b[0] / 6.0) * (double)SEXPANSION;
    Y += (double)carX[inccarX] * (double)(b[1] / 6.0);
    Y += (double)(priX[0] - b[0]) * (double)SEXPANSION;
    i = 2;
  } else {
    Y += (double)carX[0] * (double)(b[0] / 6.0);
    i = 1;
  }
  for (; i < fold; i++) {
    Y += (double)carX[i * inccarX] * (double)(b[i] / 6.0);
    Y += (double)(priX[(i - 1) * incpriX] - b[i - 1]);
  }
  Y += (double)(priX[(fold - 1) * incpriX] - b[fold - 1]);
  return (float)Y;
}
float orc_ssbconv(int fold, const float *X) { return orc_ssmconv(fold, X, 1, X + fold, 1); }

/* --------------------------------------------------------------- scale --- */
/* reference src/binned/sscale.c (C '%' truncates toward zero) */
float orc_sscale(float X) {
  int e = fexp(X);
  e = e < ORC_SBWIDTH ? ORC_SBWIDTH : e;
  e -= (e - 126 - 1) % ORC_SBWIDTH;
  e -= 126;
  return ldexpf(0.5f, e);
}

/* reference src/binned/smsrescale.c */
void orc_smsrescale(int fold, float X, float scaleY, float *priY, int incpriY, float *carY, int inccarY) {
  int i;
  float r;
  if (X == scaleY || X == 0.0f || scaleY == 0.0f) return;
  r = X / scaleY;
  r *= r;
  for (i = 0; i < fold; i++) {
    priY[i * incpriY] /= r;
    if (priY[i * incpriY] == 0.0f) {
      orc_smsupdate(fold - i, 0.0f, priY + i * incpriY, incpriY, carY + i * inccarY, inccarY);
      return;
    }
  }
}

/* reference src/binned/smsmaddsq.c via sbsbaddsq.c */
float orc_sbsbaddsq(int fold, float scaleX, const float *X, float scaleY, float *Y) {
  if (scaleX > scaleY) {
    orc_smsrescale(fold, scaleX, scaleY, Y, 1, Y + fold, 1);
    orc_sbsbadd(fold, X, Y);
    return scaleX;
  }
  if (scaleX == scaleY) {
    orc_sbsbadd(fold, X, Y);
    return scaleX;
  }
  {
    float tmp[2 * ORC_SMAXFOLD];
    memcpy(tmp, X, 2 * (size_t)fold * sizeof(float));
    orc_smsrescale(fold, scaleY, scaleX, tmp, 1, tmp + fold, 1);
    orc_sbsbadd(fold, tmp, Y);
    return scaleY;
  }
}

/* reference src/binned/sbbound.c */
float orc_sbbound(int fold, int N, float X, float S) {
  return (float)(MAXI(fabs((double)X), ldexp(0.5, FLT_MIN_EXP - 1)) * ldexp(0.5, (1 - fold) * ORC_SBWIDTH + 1) * N +
                 ((7.0 * FLT_EPSILON) / (1.0 - 6.0 * sqrt((double)FLT_EPSILON) - 7.0 * FLT_EPSILON)) * fabs((double)S));
}

/* ---------------------------------------------------------- max scans --- */
/* reference src/binnedBLAS/samax.ccog (scalar section) */
float orc_samax(int N, const float *X, int incX) {
  float m = 0.0f, v;
  int i;
  for (i = 0; i < N; i++) { v = fabsf(X[(long)i * incX]); m = (m > v ? m : v); }
  return m;
}

/* reference src/binnedBLAS/samaxm.ccog (scalar section) */
float orc_samaxm(int N, const float *X, int incX, const float *Y, int incY) {
  float m = 0.0f, v;
  int i;
  for (i = 0; i < N; i++) { v = fabsf(X[(long)i * incX] * Y[(long)i * incY]); m = (m > v ? m : v); }
  return m;
}

/* ------------------------------------------------------ block drivers --- */
enum { S_SUM, S_ASUM, S_DOT };

static inline float sterm_of(int kind, const float *X, long ix, const float *Y, long iy) {
  switch (kind) {
    case S_SUM: return X[ix];
    case S_ASUM: return fabsf(X[ix]);
    default: return X[ix] * Y[iy];
  }
}

/* reference src/binnedBLAS/smssum.ccog, smsasum.ccog, smsdot.ccog: blocks of
 * N_block_MAX = 512 (= binned_SBENDURANCE) elements, same shape as the
 * double drivers. */
static void sdrive(int kind, int fold, int N, const float *X, int incX, const float *Y, int incY, float *pri,
                   float *car) {
  int i, j, nb, deposits = 0;
  for (i = 0; i < N; i += nb) {
    float amax;
    const float *Xb = X + (long)i * incX;
    const float *Yb = (kind == S_DOT) ? Y + (long)i * incY : NULL;
    nb = MINI(N - i, 512);
    amax = (kind == S_DOT) ? orc_samaxm(nb, Xb, incX, Yb, incY) : orc_samax(nb, Xb, incX);
    if (isinf(amax) || isinf(pri[0]))
      for (j = 0; j < nb; j++) pri[0] += sterm_of(kind, Xb, (long)j * incX, Yb, (long)j * incY);
    if (isnan(pri[0])) return;
    if (isinf(pri[0])) continue;
    if (deposits + nb > ORC_SENDURANCE) { orc_smrenorm(fold, pri, 1, car, 1); deposits = 0; }
    orc_smsupdate(fold, amax, pri, 1, car, 1);
    {
      int top = orc_smindex0(pri);
      for (j = 0; j < nb; j++) sdeposit_one(fold, top, sterm_of(kind, Xb, (long)j * incX, Yb, (long)j * incY), pri, 1);
    }
    deposits += nb;
  }
  orc_smrenorm(fold, pri, 1, car, 1);
}

/* reference src/binnedBLAS/sbssum.c, sbsasum.c, sbsdot.c */
void orc_sbssum(int fold, int N, const float *X, int incX, float *Y) { sdrive(S_SUM, fold, N, X, incX, NULL, 0, Y, Y + fold); }
void orc_sbsasum(int fold, int N, const float *X, int incX, float *Y) { sdrive(S_ASUM, fold, N, X, incX, NULL, 0, Y, Y + fold); }
void orc_sbsdot(int fold, int N, const float *X, int incX, const float *Y, int incY, float *Z) {
  sdrive(S_DOT, fold, N, X, incX, Y, incY, Z, Z + fold);
}

/* reference src/binnedBLAS/smsssq.ccog and sbsssq.c; deposit of
 * fl(fl(x*(1/scl))^2) in float (the generated loops' scale_mask_inv). */
float orc_sbsssq(int fold, int N, const float *X, int incX, float scaleY, float *Yb) {
  float *pri = Yb, *car = Yb + fold;
  float scl = scaleY;
  int i, j, nb, deposits = 0;
  for (i = 0; i < N; i += nb) {
    float amax, nscl;
    const float *Xb = X + (long)i * incX;
    nb = MINI(N - i, 512);
    amax = orc_samax(nb, Xb, incX);
    if (isinf(amax) || isinf(pri[0]))
      for (j = 0; j < nb; j++) pri[0] += fabsf(Xb[(long)j * incX]);
    if (isnan(pri[0])) return orc_sscale(1.0f);
    if (isinf(pri[0])) continue;
    if (deposits + nb > ORC_SENDURANCE) { orc_smrenorm(fold, pri, 1, car, 1); deposits = 0; }
    nscl = orc_sscale(amax);
    if (nscl > scl) {
      if (scl > 0.0f) orc_smsrescale(fold, nscl, scl, pri, 1, car, 1);
      scl = nscl;
    }
    amax /= scl;
    amax = amax * amax;
    orc_smsupdate(fold, amax, pri, 1, car, 1);
    {
      int top = orc_smindex0(pri);
      float inv = (float)(1.0 / scl);
      for (j = 0; j < nb; j++) {
        float t = Xb[(long)j * incX] * inv;
        sdeposit_one(fold, top, t * t, pri, 1);
      }
    }
    deposits += nb;
  }
  orc_smrenorm(fold, pri, 1, car, 1);
  return fnan_or_inf(pri[0]) ? orc_sscale(1.0f) : scl;
}

/* ------------------------------------------------------------ user API --- */
/* reference src/reproBLAS/rssum.c, rsasum.c, rsdot.c */
float orc_rssum(int fold, int N, const float *X, int incX) {
  float st[2 * ORC_SMAXFOLD] = {0};
  orc_sbssum(fold, N, X, incX, st);
  return orc_ssbconv(fold, st);
}
float orc_rsasum(int fold, int N, const float *X, int incX) {
  float st[2 * ORC_SMAXFOLD] = {0};
  orc_sbsasum(fold, N, X, incX, st);
  return orc_ssbconv(fold, st);
}
float orc_rsdot(int fold, int N, const float *X, int incX, const float *Y, int incY) {
  float st[2 * ORC_SMAXFOLD] = {0};
  orc_sbsdot(fold, N, X, incX, Y, incY, st);
  return orc_ssbconv(fold, st);
}
/* reference src/reproBLAS/rsnrm2.c: scl * sqrt(conv) in double, rounded to
 * float by the assignment */
float orc_rsnrm2(int fold, int N, const float *X, int incX) {
  float st[2 * ORC_SMAXFOLD] = {0};
  float scl = orc_sbsssq(fold, N, X, incX, 0.0f, st);
  return (float)(scl * sqrt(orc_ssbconv(fold, st)));
}

static inline float smat_at(char order, const float *A, int ld, int r, int c) {
  return (order == 'r' || order == 'R') ? A[(long)r * ld + c] : A[(long)c * ld + r];
}
static inline int s_notrans(char t) { return t == 'n' || t == 'N'; }

/* reference src/binnedBLAS/sbsgemv.ccog: alpha on X, one binned dot per
 * output (the reference's blocked sbsdot calls equal it bit for bit). */
void orc_sbsgemv(int fold, char Order, char TransA, int M, int N, float alpha, const float *A, int lda,
                 const float *X, int incX, float *Y, int incY) {
  int opM = s_notrans(TransA) ? M : N, opN = s_notrans(TransA) ? N : M;
  int i, j;
  float *xs, *row;
  if (M == 0 || N == 0 || alpha == 0.0f) return;
  xs = (float *)malloc(sizeof(float) * (size_t)opN);
  row = (float *)malloc(sizeof(float) * (size_t)opN);
  for (j = 0; j < opN; j++) xs[j] = (alpha == 1.0f) ? X[(long)j * incX] : X[(long)j * incX] * alpha;
  for (i = 0; i < opM; i++) {
    for (j = 0; j < opN; j++) row[j] = s_notrans(TransA) ? smat_at(Order, A, lda, i, j) : smat_at(Order, A, lda, j, i);
    orc_sbsdot(fold, opN, row, 1, xs, 1, Y + (long)i * incY * 2 * fold);
  }
  free(xs);
  free(row);
}

/* reference src/reproBLAS/rsgemv.c */
void orc_rsgemv(int fold, char Order, char TransA, int M, int N, float alpha, const float *A, int lda,
                const float *X, int incX, float beta, float *Y, int incY) {
  int opM = s_notrans(TransA) ? M : N, i;
  float *YI;
  if (M == 0 || N == 0) return;
  YI = (float *)calloc((size_t)opM * 2 * fold, sizeof(float));
  if (beta != 0.0f)
    for (i = 0; i < opM; i++)
      orc_sbsconv(fold, beta == 1.0f ? Y[(long)i * incY] : Y[(long)i * incY] * beta, YI + (long)i * 2 * fold);
  orc_sbsgemv(fold, Order, TransA, M, N, alpha, A, lda, X, incX, YI, 1);
  for (i = 0; i < opM; i++) Y[(long)i * incY] = orc_ssbconv(fold, YI + (long)i * 2 * fold);
  free(YI);
}

/* reference src/binnedBLAS/sbsgemm.ccog: alpha on A, one binned dot per
 * output; C binned with leading dimension ldc. */
void orc_sbsgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K, float alpha, const float *A,
                 int lda, const float *B, int ldb, float *C, int ldc) {
  int i, j, k, rowmaj = (Order == 'r' || Order == 'R');
  float *arow, *bcol;
  if (M == 0 || N == 0 || K == 0 || alpha == 0.0f) return;
  arow = (float *)malloc(sizeof(float) * (size_t)K);
  bcol = (float *)malloc(sizeof(float) * (size_t)K);
  for (i = 0; i < M; i++) {
    for (k = 0; k < K; k++) {
      float a = s_notrans(TransA) ? smat_at(Order, A, lda, i, k) : smat_at(Order, A, lda, k, i);
      arow[k] = (alpha == 1.0f) ? a : a * alpha;
    }
    for (j = 0; j < N; j++) {
      long off = rowmaj ? ((long)i * ldc + j) : ((long)j * ldc + i);
      for (k = 0; k < K; k++) bcol[k] = s_notrans(TransB) ? smat_at(Order, B, ldb, k, j) : smat_at(Order, B, ldb, j, k);
      orc_sbsdot(fold, K, arow, 1, bcol, 1, C + off * 2 * fold);
    }
  }
  free(arow);
  free(bcol);
}

/* reference src/reproBLAS/rsgemm.c */
void orc_rsgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K, float alpha, const float *A,
                int lda, const float *B, int ldb, float beta, float *C, int ldc) {
  int i, j, rowmaj = (Order == 'r' || Order == 'R');
  int ldci = rowmaj ? N : M;
  float *CI;
  if (M == 0 || N == 0) return;
  CI = (float *)calloc((size_t)M * N * 2 * fold, sizeof(float));
  if (beta != 0.0f)
    for (i = 0; i < M; i++)
      for (j = 0; j < N; j++) {
        long src = rowmaj ? ((long)i * ldc + j) : ((long)j * ldc + i);
        long dst = rowmaj ? ((long)i * ldci + j) : ((long)j * ldci + i);
        orc_sbsconv(fold, beta == 1.0f ? C[src] : C[src] * beta, CI + dst * 2 * fold);
      }
  orc_sbsgemm(fold, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, CI, ldci);
  for (i = 0; i < M; i++)
    for (j = 0; j < N; j++) {
      long src = rowmaj ? ((long)i * ldci + j) : ((long)j * ldci + i);
      long dst = rowmaj ? ((long)i * ldc + j) : ((long)j * ldc + i);
      C[dst] = orc_ssbconv(fold, CI + src * 2 * fold);
    }
  free(CI);
}
