/*
 * oracle.c -- scalar CPU restatement of ReproBLAS 2.1.0's binned FP64 path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Compiled with -ffp-contract=off so
 * no a*b+c is ever fused, matching the reference's -std=c99 build
 * (reference config.mk:22).
 *
 * Each function names the reference file:line whose behaviour it restates.
 * The deposit loops follow the scalar (SISD) sections of the reference's
 * generated kernels; the binned result does not depend on lane layout, so the
 * AVX/SSE sections need no separate restatement.
 *
 * Pinning (tests/test_oracle_*.py): bit-equality with the compiled reference
 * (oracle/_ref, built from the reference's own sources by oracle/Makefile)
 * over random, ill-conditioned and special-value inputs; closed-form sums;
 * brute force with exact rationals on tiny inputs; the committed golden
 * fixtures in tests/golden/.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define MINI(a, b) ((a) < (b) ? (a) : (b))

/* 2^-(53-40+1) and 2^(53-40+1): reference include/binned.h:184,206 */
static const double COMPRESSION = 1.0 / (1 << 14);
static const double EXPANSION = 1.0 * (1 << 14);

static inline uint64_t as_bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
/* biased exponent field: reference src/common/common.h:43-47 */
static inline int biased_exp(double x) { return (int)((as_bits(x) >> 52) & 0x7ff); }
/* exponent all ones: reference src/common/common.h:59-63 */
static inline int nan_or_inf(double x) { return biased_exp(x) == 0x7ff; }
/* force the least significant mantissa bit to one (the "|1" of the deposit) */
static inline double set_lsb(double x) { return from_bits(as_bits(x) | 1ull); }

/* ---------------------------------------------------------------- bins --- */
/* reference src/binned/dmbins.c:23-39: bins[0] = 1.5*2^1023,
 * bins[i] = 0.75*2^(1038-40i) for 1 <= i <= 51, then bins[51] repeated. */
static double bin_table[ORC_MAXINDEX + ORC_MAXFOLD];
static int bin_table_ready = 0;

const double *orc_dmbins(int X) {
  if (!bin_table_ready) {
    int i;
    bin_table[0] = 2.0 * ldexp(0.75, 1023);
    for (i = 1; i <= ORC_MAXINDEX; i++) bin_table[i] = ldexp(0.75, 1038 - ORC_DBWIDTH * i);
    for (; i < ORC_MAXINDEX + ORC_MAXFOLD; i++) bin_table[i] = bin_table[ORC_MAXINDEX];
    bin_table_ready = 1;
  }
  return bin_table + X;
}

/* --------------------------------------------------------------- index --- */
/* reference src/binned/dindex.c:19-41 */
int orc_dindex(double X) {
  int e = biased_exp(X);
  if (e == 0) {
    if (X == 0.0) return ORC_MAXINDEX;
    frexp(X, &e);
    return MINI((1024 - e) / ORC_DBWIDTH, ORC_MAXINDEX);
  }
  return (2046 - e) / ORC_DBWIDTH;
}

/* reference src/binned/dmindex.c:22-38 */
int orc_dmindex(const double *priX) { return (2060 - biased_exp(priX[0])) / ORC_DBWIDTH; }

/* reference src/binned/dmindex0.c:19-31 */
int orc_dmindex0(const double *priX) { return biased_exp(priX[0]) == 2046; }

/* -------------------------------------------------------------- update --- */
/* reference src/binned/dmdupdate.c:24-56 */
void orc_dmdupdate(int fold, double X, double *priY, int incpriY, double *carY, int inccarY) {
  int i, shift, xi;
  const double *b;
  if (nan_or_inf(priY[0])) return;
  xi = orc_dindex(X);
  if (priY[0] == 0.0) {
    b = orc_dmbins(xi);
    for (i = 0; i < fold; i++) { priY[i * incpriY] = b[i]; carY[i * inccarY] = 0.0; }
    return;
  }
  shift = orc_dmindex(priY) - xi;
  if (shift <= 0) return;
  for (i = fold - 1; i >= shift; i--) {
    priY[i * incpriY] = priY[(i - shift) * incpriY];
    carY[i * inccarY] = carY[(i - shift) * inccarY];
  }
  b = orc_dmbins(xi);
  for (i = 0; i < MINI(shift, fold); i++) { priY[i * incpriY] = b[i]; carY[i * inccarY] = 0.0; }
}

/* ------------------------------------------------------------- deposit --- */
/* One element into primaries that already admit it. Non-top path and top
 * (index 0, compressed) path as in reference src/binned/dmddeposit.c:35-70
 * and the generated scalar loops of src/binnedBLAS/dmdsum.ccog:5969-6192.
 * No NaN/Inf test: the generated loops have none (the block drivers route
 * infinities through their own path). */
static inline void deposit_one(int fold, int top, double x, double *s, int inc) {
  int j = 0;
  double q, m;
  if (top) {
    m = s[0];
    q = m + set_lsb(x * COMPRESSION);
    s[0] = q;
    m = (m - q) * (EXPANSION * 0.5);
    x = (x + m) + m;
    j = 1;
  }
  for (; j < fold - 1; j++) {
    m = s[j * inc];
    q = m + set_lsb(x);
    s[j * inc] = q;
    x = x + (m - q);
  }
  s[j * inc] = s[j * inc] + set_lsb(x);
}

/* reference src/binned/dmddeposit.c:24-71 (with its NaN/Inf short cut) */
void orc_dmddeposit(int fold, double X, double *priY, int incpriY) {
  if (nan_or_inf(X) || nan_or_inf(priY[0])) { priY[0] += X; return; }
  deposit_one(fold, orc_dmindex0(priY), X, priY, incpriY);
}

/* -------------------------------------------------------------- renorm --- */
/* reference src/binned/dmrenorm.c:70-85 (the active version) */
void orc_dmrenorm(int fold, double *priX, int incpriX, double *carX, int inccarX) {
  int i;
  if (priX[0] == 0.0 || nan_or_inf(priX[0])) return;
  for (i = 0; i < fold; i++) {
    uint64_t u = as_bits(priX[i * incpriX]);
    carX[i * inccarX] += (double)((int)((u >> 50) & 3) - 2);
    u &= ~(1ull << 50);
    u |= 1ull << 51;
    priX[i * incpriX] = from_bits(u);
  }
}

/* ----------------------------------------------------------------- add --- */
/* reference src/binned/dmdmadd.c:27-74 */
void orc_dmdmadd(int fold, const double *priX, int incpriX, const double *carX, int inccarX,
                 double *priY, int incpriY, double *carY, int inccarY) {
  int i, shift, xi, yi;
  const double *b;
  if (priX[0] == 0.0) return;
  if (priY[0] == 0.0) {
    for (i = 0; i < fold; i++) { priY[i * incpriY] = priX[i * incpriX]; carY[i * inccarY] = carX[i * inccarX]; }
    return;
  }
  if (nan_or_inf(priX[0]) || nan_or_inf(priY[0])) { priY[0] += priX[0]; return; }
  xi = orc_dmindex(priX);
  yi = orc_dmindex(priY);
  shift = yi - xi;
  if (shift > 0) {
    /* X sits in larger bins: Y's slots move down by shift, losing its lowest */
    b = orc_dmbins(yi);
    for (i = fold - 1; i >= shift; i--) {
      priY[i * incpriY] = priX[i * incpriX] + (priY[(i - shift) * incpriY] - b[i - shift]);
      carY[i * inccarY] = carX[i * inccarX] + carY[(i - shift) * inccarY];
    }
    for (i = 0; i < shift && i < fold; i++) { priY[i * incpriY] = priX[i * incpriX]; carY[i * inccarY] = carX[i * inccarX]; }
  } else {
    b = orc_dmbins(xi);
    for (i = -shift; i < fold; i++) {
      priY[i * incpriY] += priX[(i + shift) * incpriX] - b[i + shift];
      carY[i * inccarY] += carX[(i + shift) * inccarX];
    }
  }
  orc_dmrenorm(fold, priY, incpriY, carY, inccarY);
}

/* reference src/binned/dbdbadd.c:16-18 */
void orc_dbdbadd(int fold, const double *X, double *Y) { orc_dmdmadd(fold, X, 1, X + fold, 1, Y, 1, Y + fold, 1); }

/* reference src/binned/dmdadd.c:20-24 */
void orc_dmdadd(int fold, double X, double *priY, int incpriY, double *carY, int inccarY) {
  orc_dmdupdate(fold, X, priY, incpriY, carY, inccarY);
  orc_dmddeposit(fold, X, priY, incpriY);
  orc_dmrenorm(fold, priY, incpriY, carY, inccarY);
}
void orc_dbdadd(int fold, double X, double *Y) { orc_dmdadd(fold, X, Y, 1, Y + fold, 1); }

/* reference src/binned/dmdconv.c:18-21, dbdconv.c */
void orc_dbdconv(int fold, double X, double *Y) {
  memset(Y, 0, 2 * (size_t)fold * sizeof(double));
  orc_dmdadd(fold, X, Y, 1, Y + fold, 1);
}

/* ------------------------------------------------------------- convert --- */
/* reference src/binned/ddmconv.c:23-81: fixed summation order; scaled by
 * 2^-66 while the index is at most 3 so the top bins cannot overflow. */
double orc_ddmconv(int fold, const double *priX, int incpriX, const double *carX, int inccarX) {
  int i = 0, xi, nscaled;
  const double *b;
  double Y = 0.0, sdown, sup;
  if (nan_or_inf(priX[0])) return priX[0];
  if (priX[0] == 0.0) return 0.0;
  xi = orc_dmindex(priX);
  b = orc_dmbins(xi);
  if (xi <= (3 * 53) / ORC_DBWIDTH) {
    sdown = ldexp(0.5, 1 - (2 * 53 - ORC_DBWIDTH));
    sup = ldexp(0.5, 1 + (2 * 53 - ORC_DBWIDTH));
    nscaled = MAXI(MINI(fold, (3 * 53) / ORC_DBWIDTH - xi), 0);
    if (xi == 0) {
      Y += carX[0] * ((b[0] / 6.0) * sdown * EXPANSION);
      Y += carX[inccarX] * ((b[1] / 6.0) * sdown);
      Y += (priX[0] - b[0]) * sdown * EXPANSION;
      i = 2;
    } else {
      Y += carX[0] * ((b[0] / 6.0) * sdown);
      i = 1;
    }
    for (; i < nscaled; i++) {
      Y += carX[i * inccarX] * ((b[i] / 6.0) * sdown);
      Y += (priX[(i - 1) * incpriX] - b[i - 1]) * sdown;
    }
    if (i == fold) {
      Y += (priX[(fold - 1) * incpriX] - b[fold - 1]) * sdown;
      return Y * sup;
    }
    if (isinf(Y * sup)) return Y * sup;
    Y *= sup;
    for (; i < fold; i++) {
      Y += carX[i * inccarX] * (b[i] / 6.0);
      Y += priX[(i - 1) * incpriX] - b[i - 1];
    }
    Y += priX[(fold - 1) * incpriX] - b[fold - 1];
    return Y;
  }
  Y += carX[0] * (b[0] / 6.0);
  for (i = 1; i < fold; i++) {
    Y += carX[i * inccarX] * (b[i] / 6.0);
    Y += priX[(i - 1) * incpriX] - b[i - 1];
  }
  Y += priX[(fold - 1) * incpriX] - b[fold - 1];
  return Y;
}
double orc_ddbconv(int fold, const double *X) { return orc_ddmconv(fold, X, 1, X + fold, 1); }

/* --------------------------------------------------------------- scale --- */
/* reference src/binned/dscale.c:25-31 (C '%' truncates toward zero) */
double orc_dscale(double X) {
  int e = biased_exp(X);
  e = e < ORC_DBWIDTH ? ORC_DBWIDTH : e;
  e -= (e - 1022 - 1) % ORC_DBWIDTH;
  e -= 1022;
  return ldexp(0.5, e);
}

/* reference src/binned/dmdrescale.c:20-37 */
void orc_dmdrescale(int fold, double X, double scaleY, double *priY, int incpriY, double *carY, int inccarY) {
  int i;
  double r;
  if (X == scaleY || X == 0.0 || scaleY == 0.0) return;
  r = X / scaleY;
  r *= r;
  for (i = 0; i < fold; i++) {
    priY[i * incpriY] /= r;
    if (priY[i * incpriY] == 0.0) {
      orc_dmdupdate(fold - i, 0.0, priY + i * incpriY, incpriY, carY + i * inccarY, inccarY);
      return;
    }
  }
}

/* reference src/binned/dmdmaddsq.c:26-41 via dbdbaddsq.c:19-21 */
double orc_dbdbaddsq(int fold, double scaleX, const double *X, double scaleY, double *Y) {
  if (scaleX > scaleY) {
    orc_dmdrescale(fold, scaleX, scaleY, Y, 1, Y + fold, 1);
    orc_dbdbadd(fold, X, Y);
    return scaleX;
  }
  if (scaleX == scaleY) {
    orc_dbdbadd(fold, X, Y);
    return scaleX;
  }
  {
    double tmp[2 * ORC_MAXFOLD];
    memcpy(tmp, X, 2 * (size_t)fold * sizeof(double));
    orc_dmdrescale(fold, scaleY, scaleX, tmp, 1, tmp + fold, 1);
    orc_dbdbadd(fold, tmp, Y);
    return scaleY;
  }
}

/* reference src/binned/dbbound.c:21-23 */
double orc_dbbound(int fold, int N, double X, double S) {
  return MAXI(fabs(X), ldexp(0.5, -1021 - 1)) * ldexp(0.5, (1 - fold) * ORC_DBWIDTH + 1) * N +
         ((7.0 * 2.220446049250313080847e-16) /
          (1.0 - 6.0 * sqrt(2.220446049250313080847e-16) - 7.0 * 2.220446049250313080847e-16)) * fabs(S);
}

/* ---------------------------------------------------------- max scans --- */
/* reference src/binnedBLAS/damax.ccog:282-300 (scalar section) */
double orc_damax(int N, const double *X, int incX) {
  double m = 0.0, v;
  int i;
  for (i = 0; i < N; i++) { v = fabs(X[(long)i * incX]); m = (m > v ? m : v); }
  return m;
}

/* reference src/binnedBLAS/damaxm.ccog:355-366 (scalar section) */
double orc_damaxm(int N, const double *X, int incX, const double *Y, int incY) {
  double m = 0.0, v;
  int i;
  for (i = 0; i < N; i++) { v = fabs(X[(long)i * incX] * Y[(long)i * incY]); m = (m > v ? m : v); }
  return m;
}

/* ------------------------------------------------------ block drivers --- */
/* The four reference drivers share one shape: blocks of N_block elements;
 * per block a max scan, an infinity path that adds plain terms into the top
 * primary, the endurance renorm, an index update, then deposits; a final
 * renorm. Term kinds: */
enum { T_SUM, T_ASUM, T_DOT };

static inline double term_of(int kind, const double *X, long ix, const double *Y, long iy) {
  switch (kind) {
    case T_SUM: return X[ix];
    case T_ASUM: return fabs(X[ix]);
    default: return X[ix] * Y[iy];
  }
}

/* reference src/binnedBLAS/dmdsum.ccog:57-89,6202 (N_block 2048, :36),
 * dmdasum.ccog:57-85 (N_block 2048), dmddot.ccog:59-88,12071 (N_block 1024) */
static void drive(int kind, int fold, int N, const double *X, int incX, const double *Y, int incY,
                  double *pri, double *car) {
  const int nb_max = (kind == T_DOT) ? 1024 : 2048;
  int i, j, nb, deposits = 0;
  for (i = 0; i < N; i += nb) {
    double amax;
    const double *Xb = X + (long)i * incX;
    const double *Yb = (kind == T_DOT) ? Y + (long)i * incY : NULL;
    nb = MINI(N - i, nb_max);
    amax = (kind == T_DOT) ? orc_damaxm(nb, Xb, incX, Yb, incY) : orc_damax(nb, Xb, incX);
    if (isinf(amax) || isinf(pri[0]))
      for (j = 0; j < nb; j++) pri[0] += term_of(kind, Xb, (long)j * incX, Yb, (long)j * incY);
    if (isnan(pri[0])) return;
    if (isinf(pri[0])) continue;
    if (deposits + nb > ORC_ENDURANCE) { orc_dmrenorm(fold, pri, 1, car, 1); deposits = 0; }
    orc_dmdupdate(fold, amax, pri, 1, car, 1);
    {
      int top = orc_dmindex0(pri);
      for (j = 0; j < nb; j++) deposit_one(fold, top, term_of(kind, Xb, (long)j * incX, Yb, (long)j * incY), pri, 1);
    }
    deposits += nb;
  }
  orc_dmrenorm(fold, pri, 1, car, 1);
}

/* reference src/binnedBLAS/dbdsum.c:17-19 */
void orc_dbdsum(int fold, int N, const double *X, int incX, double *Y) { drive(T_SUM, fold, N, X, incX, NULL, 0, Y, Y + fold); }
/* reference src/binnedBLAS/dbdasum.c:16-18 */
void orc_dbdasum(int fold, int N, const double *X, int incX, double *Y) { drive(T_ASUM, fold, N, X, incX, NULL, 0, Y, Y + fold); }
/* reference src/binnedBLAS/dbddot.c:19-21 */
void orc_dbddot(int fold, int N, const double *X, int incX, const double *Y, int incY, double *Z) {
  drive(T_DOT, fold, N, X, incX, Y, incY, Z, Z + fold);
}

/* reference src/binnedBLAS/dmdssq.ccog:60-101 and its tail; deposit of
 * fl(fl(x*(1/scl))^2) from src/binnedBLAS/depositSSq.py:17-27 */
double orc_dbdssq(int fold, int N, const double *X, int incX, double scaleY, double *Yb) {
  double *pri = Yb, *car = Yb + fold;
  double scl = scaleY;
  int i, j, nb, deposits = 0;
  for (i = 0; i < N; i += nb) {
    double amax, nscl;
    const double *Xb = X + (long)i * incX;
    nb = MINI(N - i, 2048);
    amax = orc_damax(nb, Xb, incX);
    if (isinf(amax) || isinf(pri[0]))
      for (j = 0; j < nb; j++) pri[0] += fabs(Xb[(long)j * incX]);
    if (isnan(pri[0])) return orc_dscale(1.0);
    if (isinf(pri[0])) continue;
    if (deposits + nb > ORC_ENDURANCE) { orc_dmrenorm(fold, pri, 1, car, 1); deposits = 0; }
    nscl = orc_dscale(amax);
    if (nscl > scl) {
      if (scl > 0.0) orc_dmdrescale(fold, nscl, scl, pri, 1, car, 1);
      scl = nscl;
    }
    amax /= scl;
    amax = amax * amax;
    orc_dmdupdate(fold, amax, pri, 1, car, 1);
    {
      int top = orc_dmindex0(pri);
      double inv = 1.0 / scl;
      for (j = 0; j < nb; j++) {
        double t = Xb[(long)j * incX] * inv;
        deposit_one(fold, top, t * t, pri, 1);
      }
    }
    deposits += nb;
  }
  orc_dmrenorm(fold, pri, 1, car, 1);
  return nan_or_inf(pri[0]) ? orc_dscale(1.0) : scl;
}

/* ------------------------------------------------------------ user API --- */
/* reference src/reproBLAS/rdsum.c:20-31 */
double orc_rdsum(int fold, int N, const double *X, int incX) {
  double st[2 * ORC_MAXFOLD] = {0};
  orc_dbdsum(fold, N, X, incX, st);
  return orc_ddbconv(fold, st);
}
/* reference src/reproBLAS/rdasum.c:20-31 */
double orc_rdasum(int fold, int N, const double *X, int incX) {
  double st[2 * ORC_MAXFOLD] = {0};
  orc_dbdasum(fold, N, X, incX, st);
  return orc_ddbconv(fold, st);
}
/* reference src/reproBLAS/rddot.c:22-33 */
double orc_rddot(int fold, int N, const double *X, int incX, const double *Y, int incY) {
  double st[2 * ORC_MAXFOLD] = {0};
  orc_dbddot(fold, N, X, incX, Y, incY, st);
  return orc_ddbconv(fold, st);
}
/* reference src/reproBLAS/rdnrm2.c:22-34 */
double orc_rdnrm2(int fold, int N, const double *X, int incX) {
  double st[2 * ORC_MAXFOLD] = {0};
  double scl = orc_dbdssq(fold, N, X, incX, 0.0, st);
  return scl * sqrt(orc_ddbconv(fold, st));
}

/* element (r, c) of op(A) where A is M-by-N in storage order `order` */
static inline double mat_at(char order, const double *A, int ld, int r, int c) {
  return (order == 'r' || order == 'R') ? A[(long)r * ld + c] : A[(long)c * ld + r];
}
static inline int is_notrans(char t) { return t == 'n' || t == 'N'; }

/* reference src/binnedBLAS/dbdgemv.ccog:57-162: y_i (+)= sum_j fl(A_ij * fl(alpha*x_j));
 * alpha is applied to X (:86-95, :110-119). One binned dot per output, which
 * the reference's blocked calls equal bit for bit (its tests/checks/
 * wrap_rdgemv.h:16-60 is the same per-element formulation). */
void orc_dbdgemv(int fold, char Order, char TransA, int M, int N, double alpha,
                 const double *A, int lda, const double *X, int incX, double *Y, int incY) {
  int opM = is_notrans(TransA) ? M : N, opN = is_notrans(TransA) ? N : M;
  int i, j;
  double *xs, *row;
  if (M == 0 || N == 0 || alpha == 0.0) return;
  xs = (double *)malloc(sizeof(double) * (size_t)opN);
  row = (double *)malloc(sizeof(double) * (size_t)opN);
  for (j = 0; j < opN; j++) xs[j] = (alpha == 1.0) ? X[(long)j * incX] : X[(long)j * incX] * alpha;
  for (i = 0; i < opM; i++) {
    for (j = 0; j < opN; j++)
      row[j] = is_notrans(TransA) ? mat_at(Order, A, lda, i, j) : mat_at(Order, A, lda, j, i);
    orc_dbddot(fold, opN, row, 1, xs, 1, Y + (long)i * incY * 2 * fold);
  }
  free(xs);
  free(row);
}

/* reference src/reproBLAS/rdgemv.c:35-87 */
void orc_rdgemv(int fold, char Order, char TransA, int M, int N, double alpha,
                const double *A, int lda, const double *X, int incX,
                double beta, double *Y, int incY) {
  int opM = is_notrans(TransA) ? M : N, i;
  double *YI;
  if (M == 0 || N == 0) return;
  YI = (double *)calloc((size_t)opM * 2 * fold, sizeof(double));
  if (beta != 0.0)
    for (i = 0; i < opM; i++)
      orc_dbdconv(fold, beta == 1.0 ? Y[(long)i * incY] : Y[(long)i * incY] * beta, YI + (long)i * 2 * fold);
  orc_dbdgemv(fold, Order, TransA, M, N, alpha, A, lda, X, incX, YI, 1);
  for (i = 0; i < opM; i++) Y[(long)i * incY] = orc_ddbconv(fold, YI + (long)i * 2 * fold);
  free(YI);
}

/* reference src/binnedBLAS/dbdgemm.ccog:61-420: C_ij (+)= sum_k fl(fl(alpha*A_ik) * B_kj);
 * alpha is applied to A (:92-104 and every other branch). C is binned,
 * laid out like the reference's CI buffer with leading dimension ldc. */
void orc_dbdgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K,
                 double alpha, const double *A, int lda, const double *B, int ldb,
                 double *C, int ldc) {
  int i, j, k, rowmaj = (Order == 'r' || Order == 'R');
  double *arow, *bcol;
  if (M == 0 || N == 0 || K == 0 || alpha == 0.0) return;
  arow = (double *)malloc(sizeof(double) * (size_t)K);
  bcol = (double *)malloc(sizeof(double) * (size_t)K);
  for (i = 0; i < M; i++) {
    for (k = 0; k < K; k++) {
      double a = is_notrans(TransA) ? mat_at(Order, A, lda, i, k) : mat_at(Order, A, lda, k, i);
      arow[k] = (alpha == 1.0) ? a : a * alpha;
    }
    for (j = 0; j < N; j++) {
      long off = rowmaj ? ((long)i * ldc + j) : ((long)j * ldc + i);
      for (k = 0; k < K; k++)
        bcol[k] = is_notrans(TransB) ? mat_at(Order, B, ldb, k, j) : mat_at(Order, B, ldb, j, k);
      orc_dbddot(fold, K, arow, 1, bcol, 1, C + off * 2 * fold);
    }
  }
  free(arow);
  free(bcol);
}

/* reference src/reproBLAS/rdgemm.c:41-105 */
void orc_rdgemm(int fold, char Order, char TransA, char TransB, int M, int N, int K,
                double alpha, const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc) {
  int i, j, rowmaj = (Order == 'r' || Order == 'R');
  int ldci = rowmaj ? N : M;
  double *CI;
  if (M == 0 || N == 0) return;
  CI = (double *)calloc((size_t)M * N * 2 * fold, sizeof(double));
  if (beta != 0.0)
    for (i = 0; i < M; i++)
      for (j = 0; j < N; j++) {
        long src = rowmaj ? ((long)i * ldc + j) : ((long)j * ldc + i);
        long dst = rowmaj ? ((long)i * ldci + j) : ((long)j * ldci + i);
        orc_dbdconv(fold, beta == 1.0 ? C[src] : C[src] * beta, CI + dst * 2 * fold);
      }
  orc_dbdgemm(fold, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, CI, ldci);
  for (i = 0; i < M; i++)
    for (j = 0; j < N; j++) {
      long src = rowmaj ? ((long)i * ldci + j) : ((long)j * ldci + i);
      long dst = rowmaj ? ((long)i * ldc + j) : ((long)j * ldc + i);
      C[dst] = orc_ddbconv(fold, CI + src * 2 * fold);
    }
  free(CI);
}
