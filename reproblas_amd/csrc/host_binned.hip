// host_binned.hip -- the binned_* primitives of include/binned.h.
//
// These act on one binned value (a few dozen bytes) in host memory and run on
// the host, as the reference's do; they share their arithmetic with the
// kernels through binned_core.hpp (compiled for both sides with
// -ffp-contract=off). binned_dbdbaddv on device arrays runs on the GPU.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../../include/binned.h"
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace rb;

namespace {

constexpr int CAP = kMaxFold;

struct Local {
  double p[CAP];
  double c[CAP];
};

inline void gather(int fold, const double* pri, int incp, const double* car, int incc, Local& v) {
  for (int i = 0; i < fold; ++i) {
    v.p[i] = pri[(long)i * incp];
    v.c[i] = car[(long)i * incc];
  }
  for (int i = fold; i < CAP; ++i) v.p[i] = v.c[i] = 0.0;
}
inline void scatter(int fold, const Local& v, double* pri, int incp, double* car, int incc) {
  for (int i = 0; i < fold; ++i) {
    pri[(long)i * incp] = v.p[i];
    car[(long)i * incc] = v.c[i];
  }
}

struct BinTable {
  double v[kMaxIndex + kMaxFold];
  BinTable() {
    for (int i = 0; i < kMaxIndex + kMaxFold; ++i) v[i] = bin(i);
  }
};
const BinTable g_bins;  // built at load time: no lazy initialisation race

// the reference's UFP returns float (src/common/common.h:29-34), so
// binned_ufp narrows through float
inline double ufp_narrowed(double x) { return (double)(float)from_bits(bits(x) & (0x7ffull << 52)); }

}  // namespace

extern "C" {

// reference src/binned/dbsize.c, dballoc.c, dbnum.c
size_t binned_dbsize(const int fold) { return 2 * (size_t)fold * sizeof(double); }
double_binned* binned_dballoc(const int fold) { return (double_binned*)malloc(binned_dbsize(fold)); }
int binned_dbnum(const int fold) { return 2 * fold; }

// reference src/binned/dbbound.c:21-23
double binned_dbbound(const int fold, const int N, const double X, const double S) {
  const double a = fabs(X) > ldexp(0.5, DBL_MIN_EXP - 1) ? fabs(X) : ldexp(0.5, DBL_MIN_EXP - 1);
  return a * ldexp(0.5, (1 - fold) * DBWIDTH + 1) * N +
         ((7.0 * DBL_EPSILON) / (1.0 - 6.0 * sqrt(DBL_EPSILON) - 7.0 * DBL_EPSILON)) * fabs(S);
}

// reference src/binned/dmbins.c:23-39
const double* binned_dmbins(const int X) { return g_bins.v + X; }
// reference src/binned/dindex.c, dmindex.c, dmindex0.c
int binned_dindex(const double X) { return dindex(X); }
int binned_dmindex(const double* priX) { return dmindex(priX[0]); }
int binned_dmindex0(const double* priX) { return dmindex0(priX[0]) ? 1 : 0; }
// reference src/binned/dmdenorm.c:20-22
int binned_dmdenorm(const int fold, const double* priX) {
  return biased_exp(priX[0]) - DBWIDTH * fold < DBL_MIN_EXP + DBL_MANT_DIG + (DBL_MAX_EXP - 2) + DBWIDTH;
}
// reference src/binned/ufp.c
double binned_ufp(const double X) { return ufp_narrowed(X); }

// reference src/binned/dmprint.c
void binned_dmprint(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX) {
  for (int i = 0; i < fold; ++i) {
    const double p = priX[(long)i * incpriX], c = carX[(long)i * inccarX];
    const double M = ufp_narrowed(p);
    printf("(2^%d: %g #%g =%g)\n", (int)log2(M) + 1, p - 1.5 * M, c, ((c - 6) * 0.25 * M + p));
  }
}
void binned_dbprint(const int fold, const double_binned* X) { binned_dmprint(fold, X, 1, X + fold, 1); }

// reference src/binned/dmdmset.c, dbdbset.c, dmsetzero.c, dbsetzero.c
void binned_dmdmset(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX,
                    double* priY, const int incpriY, double* carY, const int inccarY) {
  for (int i = 0; i < fold; ++i) {
    priY[(long)i * incpriY] = priX[(long)i * incpriX];
    carY[(long)i * inccarY] = carX[(long)i * inccarX];
  }
}
void binned_dbdbset(const int fold, const double_binned* X, double_binned* Y) {
  binned_dmdmset(fold, X, 1, X + fold, 1, Y, 1, Y + fold, 1);
}
void binned_dmsetzero(const int fold, double* priX, const int incpriX, double* carX, const int inccarX) {
  for (int i = 0; i < fold; ++i) {
    priX[(long)i * incpriX] = 0.0;
    carX[(long)i * inccarX] = 0.0;
  }
}
void binned_dbsetzero(const int fold, double_binned* X) { memset(X, 0, binned_dbsize(fold)); }

// reference src/binned/dmdmadd.c:27-74
void binned_dmdmadd(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX,
                    double* priY, const int incpriY, double* carY, const int inccarY) {
  Local x, y;
  gather(fold, priX, incpriX, carX, inccarX, x);
  gather(fold, priY, incpriY, carY, inccarY, y);
  add(fold, y.p, y.c, x.p, x.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
}
void binned_dbdbadd(const int fold, const double_binned* X, double_binned* Y) {
  binned_dmdmadd(fold, X, 1, X + fold, 1, Y, 1, Y + fold, 1);
}

// reference src/binned/dbdbaddv.c:18-23; device arrays run on the GPU
void binned_dbdbaddv(const int fold, const int N, const double_binned* X, const int incX, double_binned* Y,
                     const int incY) {
  clear_error();
  if (N <= 0) return;
  const bool dx = is_device_pointer(X), dy = is_device_pointer(Y);
  if (dx && dy) {
    hipStream_t s = current_stream();
    if (dbdbaddv(fold, N, X, incX, Y, incY, s)) RB_HIP(hipStreamSynchronize(s));
    return;
  }
  if (dx != dy) {
    set_error(kBadArgument, "binned_dbdbaddv: X and Y must both be host or both be device arrays");
    return;
  }
  for (long i = 0; i < N; ++i)
    binned_dbdbadd(fold, X + i * incX * 2 * fold, Y + i * incY * 2 * fold);
}

// reference src/binned/dmdupdate.c:24-56
void binned_dmdupdate(const int fold, const double X, double* priY, const int incpriY, double* carY,
                      const int inccarY) {
  Local y;
  gather(fold, priY, incpriY, carY, inccarY, y);
  update(fold, X, y.p, y.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
}
void binned_dbdupdate(const int fold, const double X, double_binned* Y) {
  binned_dmdupdate(fold, X, Y, 1, Y + fold, 1);
}

// reference src/binned/dmddeposit.c:24-71
void binned_dmddeposit(const int fold, const double X, double* priY, const int incpriY) {
  if (nan_or_inf(X) || nan_or_inf(priY[0])) {
    priY[0] += X;
    return;
  }
  Local y;
  gather(fold, priY, incpriY, priY, incpriY, y);
  deposit(fold, dmindex0(y.p[0]), X, y.p);
  for (int i = 0; i < fold; ++i) priY[(long)i * incpriY] = y.p[i];
}
void binned_dbddeposit(const int fold, const double X, double_binned* Y) { binned_dmddeposit(fold, X, Y, 1); }

// reference src/binned/dmrenorm.c:70-85
void binned_dmrenorm(const int fold, double* priX, const int incpriX, double* carX, const int inccarX) {
  Local x;
  gather(fold, priX, incpriX, carX, inccarX, x);
  renorm(fold, x.p, x.c);
  scatter(fold, x, priX, incpriX, carX, inccarX);
}
void binned_dbrenorm(const int fold, double_binned* X) { binned_dmrenorm(fold, X, 1, X + fold, 1); }

// reference src/binned/dmdadd.c:20-24
void binned_dmdadd(const int fold, const double X, double* priY, const int incpriY, double* carY,
                   const int inccarY) {
  binned_dmdupdate(fold, X, priY, incpriY, carY, inccarY);
  binned_dmddeposit(fold, X, priY, incpriY);
  binned_dmrenorm(fold, priY, incpriY, carY, inccarY);
}
void binned_dbdadd(const int fold, const double X, double_binned* Y) { binned_dmdadd(fold, X, Y, 1, Y + fold, 1); }

// reference src/binned/dmdconv.c:18-21, dbdconv.c
void binned_dmdconv(const int fold, const double X, double* priY, const int incpriY, double* carY,
                    const int inccarY) {
  binned_dmsetzero(fold, priY, incpriY, carY, inccarY);
  binned_dmdadd(fold, X, priY, incpriY, carY, inccarY);
}
void binned_dbdconv(const int fold, const double X, double_binned* Y) { binned_dmdconv(fold, X, Y, 1, Y + fold, 1); }

// reference src/binned/ddmconv.c:23-81
double binned_ddmconv(const int fold, const double* priX, const int incpriX, const double* carX,
                      const int inccarX) {
  Local x;
  gather(fold, priX, incpriX, carX, inccarX, x);
  return to_double(fold, x.p, x.c);
}
double binned_ddbconv(const int fold, const double_binned* X) { return binned_ddmconv(fold, X, 1, X + fold, 1); }

// reference src/binned/dmnegate.c
void binned_dmnegate(const int fold, double* priX, const int incpriX, double* carX, const int inccarX) {
  if (priX[0] == 0.0) return;
  const int I = dmindex(priX[0]);
  for (int i = 0; i < fold; ++i) {
    const double b = bin(I + i);
    priX[(long)i * incpriX] = b - (priX[(long)i * incpriX] - b);
    carX[(long)i * inccarX] = -carX[(long)i * inccarX];
  }
}
void binned_dbnegate(const int fold, double_binned* X) { binned_dmnegate(fold, X, 1, X + fold, 1); }

// reference src/binned/dscale.c:25-31
double binned_dscale(const double X) { return dscale(X); }

// reference src/binned/dmdrescale.c:20-37
void binned_dmdrescale(const int fold, const double X, const double scaleY, double* priY, const int incpriY,
                       double* carY, const int inccarY) {
  Local y;
  gather(fold, priY, incpriY, carY, inccarY, y);
  rescale(fold, X, scaleY, y.p, y.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
}

// reference src/binned/dmdmaddsq.c:26-41
double binned_dmdmaddsq(const int fold, const double scaleX, const double* priX, const int incpriX,
                        const double* carX, const int inccarX, const double scaleY, double* priY,
                        const int incpriY, double* carY, const int inccarY) {
  Local x, y;
  gather(fold, priX, incpriX, carX, inccarX, x);
  gather(fold, priY, incpriY, carY, inccarY, y);
  const double s = add_scaled(fold, scaleX, x.p, x.c, scaleY, y.p, y.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
  return s;
}
double binned_dbdbaddsq(const int fold, const double scaleX, const double_binned* X, const double scaleY,
                        double_binned* Y) {
  return binned_dmdmaddsq(fold, scaleX, X, 1, X + fold, 1, scaleY, Y, 1, Y + fold, 1);
}

}  // extern "C"
