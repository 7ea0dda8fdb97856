// host_binned.hip -- the binned_* primitives of include/binned.h (double_binned
// and float_binned).
//
// These act on one binned value (a few dozen bytes) in host memory and run on
// the host, as the reference's do; they share their arithmetic with the
// kernels through binned_core.hpp (compiled for both sides with
// -ffp-contract=off). binned_dbdbaddv on device arrays runs on the GPU.
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include "../../include/binned.h"
#include "../../include/reproblas_hip.h"
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace rb;

namespace {

constexpr int CAP = kMaxFold;
constexpr int SCAP = Fmt<float>::max_fold;

template <class F, int N>
struct LocalT {
  F p[N];
  F c[N];
};
using Local = LocalT<double, CAP>;
using LocalS = LocalT<float, SCAP>;

template <class F, int N>
inline void gather(int fold, const F* pri, int incp, const F* car, int incc, LocalT<F, N>& v) {
  for (int i = 0; i < fold; ++i) {
    v.p[i] = pri[(long)i * incp];
    v.c[i] = car[(long)i * incc];
  }
  for (int i = fold; i < N; ++i) v.p[i] = v.c[i] = F(0);
}
template <class F, int N>
inline void scatter(int fold, const LocalT<F, N>& v, F* pri, int incp, F* car, int incc) {
  for (int i = 0; i < fold; ++i) {
    pri[(long)i * incp] = v.p[i];
    car[(long)i * incc] = v.c[i];
  }
}

template <class F>
struct BinTable {
  F v[Fmt<F>::max_index + Fmt<F>::max_fold];
  BinTable() {
    for (int i = 0; i < Fmt<F>::max_index + Fmt<F>::max_fold; ++i) v[i] = bin<F>(i);
  }
};
const BinTable<double> g_bins;  // built at load time: no lazy initialisation race
const BinTable<float> g_sbins;

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

}  // extern "C"

// ---- order-free exchange (integer form of binned values) ----
// Between renormalisations a primary stays in its bin's binade, so its offset
// from the bin in units of the bin's ulp is a difference of bit patterns;
// renorm keeps the low mbits-2 bits of an offset in the primary and moves the
// floor of the rest into the carry. The renormalised sum of values is
// therefore a function of the per-bin sums of offsets and carries alone (at
// the lowest index; slots of higher-index values below that window are the
// ones a raise drops), the same as dmdmadd's pairwise adds give.
namespace {
template <class F>
constexpr int wire_bins() { return Fmt<F>::max_index + Fmt<F>::max_fold + 1; }
static_assert(wire_bins<double>() == 104 && wire_bins<float>() == 42, "wire sizes of reproblas_hip.h");

template <class F>
int encode(int fold, const F* X, int64_t* wire, int64_t* index) {
  clear_error();
  if (fold < 2 || fold > Fmt<F>::max_fold) {
    set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
    return last_error();
  }
  if (!X || !wire || !index) {
    set_error(kBadArgument, "binned encode: null pointer");
    return last_error();
  }
  constexpr int NB = wire_bins<F>();
  for (int i = 0; i < 2 * NB + 3; ++i) wire[i] = 0;
  const F p0 = X[0];
  if (nan_or_inf(p0)) {
    wire[2 * NB + (p0 != p0 ? 0 : (p0 > F(0) ? 1 : 2))] = 1;
    *index = REPROBLAS_HIP_EMPTY_INDEX;
    return kOk;
  }
  if (p0 == F(0)) {
    *index = REPROBLAS_HIP_EMPTY_INDEX;
    return kOk;
  }
  const int I = dmindex(p0);
  for (int j = 0; j < fold; ++j) {
    wire[I + j] = (int64_t)bits(X[j]) - (int64_t)bits(bin<F>(I + j));
    wire[NB + I + j] = (int64_t)X[fold + j];
  }
  *index = I;
  return kOk;
}

template <class F>
int decode(int fold, const int64_t* wire, int64_t index, F* Y) {
  clear_error();
  if (fold < 2 || fold > Fmt<F>::max_fold) {
    set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
    return last_error();
  }
  constexpr int NB = wire_bins<F>();
  if (!wire || !Y || index < 0 || (index > Fmt<F>::max_index && index != REPROBLAS_HIP_EMPTY_INDEX)) {
    set_error(kBadArgument, "binned decode: null pointer or index %lld outside [0, %d]", (long long)index,
              Fmt<F>::max_index);
    return last_error();
  }
  using U = typename Fmt<F>::U;
  constexpr int lo = Lay<F>::mbits - 2;
  if (index == REPROBLAS_HIP_EMPTY_INDEX) {
    for (int j = 0; j < 2 * fold; ++j) Y[j] = F(0);
  } else {
    const int I = (int)index;
    for (int j = 0; j < fold; ++j) {
      const int64_t d = wire[I + j];
      Y[j] = fp_from_bits<F>((U)(bits(bin<F>(I + j)) + (U)(d & (((int64_t)1 << lo) - 1))));
      Y[fold + j] = (F)(wire[NB + I + j] + (d >> lo));
    }
  }
  const int64_t nan = wire[2 * NB], pinf = wire[2 * NB + 1], ninf = wire[2 * NB + 2];
  if (nan || (pinf && ninf)) {
    Y[0] = (F)NAN;
  } else if (pinf) {
    Y[0] = (F)INFINITY;
  } else if (ninf) {
    Y[0] = -(F)INFINITY;
  }
  return kOk;
}
}  // namespace

extern "C" {
int reproblas_hip_dbencode(int fold, const double_binned* X, int64_t* wire, int64_t* index) {
  return encode<double>(fold, X, wire, index);
}
int reproblas_hip_dbdecode(int fold, const int64_t* wire, int64_t index, double_binned* Y) {
  return decode<double>(fold, wire, index, Y);
}
int reproblas_hip_sbencode(int fold, const float_binned* X, int64_t* wire, int64_t* index) {
  return encode<float>(fold, X, wire, index);
}
int reproblas_hip_sbdecode(int fold, const int64_t* wire, int64_t index, float_binned* Y) {
  return decode<float>(fold, wire, index, Y);
}

// reference src/binned/dbdbaddv.c:18-23; device arrays run on the GPU
void binned_dbdbaddv(const int fold, const int N, const double_binned* X, const int incX, double_binned* Y,
                     const int incY) {
  clear_error();
  if (N <= 0) return;
  const bool dx = is_device_pointer(X), dy = is_device_pointer(Y);
  if (dx && dy) {
    hipStream_t s = current_stream();
    RB_STREAM_GUARD(s);
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


// ------------------------------------------------------------ float_binned --
// reference src/binned/sbsbze.c, sballoc.c, sbnum.c
size_t binned_sbsbze(const int fold) { return 2 * (size_t)fold * sizeof(float); }
float_binned* binned_sballoc(const int fold) { return (float_binned*)malloc(binned_sbsbze(fold)); }
int binned_sbnum(const int fold) { return 2 * fold; }

// reference src/binned/sbbound.c
float binned_sbbound(const int fold, const int N, const float X, const float S) {
  const double a = fabs((double)X) > ldexp(0.5, FLT_MIN_EXP - 1) ? fabs((double)X) : ldexp(0.5, FLT_MIN_EXP - 1);
  return (float)(a * ldexp(0.5, (1 - fold) * SBWIDTH + 1) * N +
                 ((7.0 * FLT_EPSILON) / (1.0 - 6.0 * sqrt((double)FLT_EPSILON) - 7.0 * FLT_EPSILON)) *
                     fabs((double)S));
}

// reference src/binned/smbins.c, sindex.c, smindex.c, smindex0.c
const float* binned_smbins(const int X) { return g_sbins.v + X; }
int binned_sindex(const float X) { return dindex(X); }
int binned_smindex(const float* priX) { return dmindex(priX[0]); }
int binned_smindex0(const float* priX) { return dmindex0(priX[0]) ? 1 : 0; }
// reference src/binned/smdenorm.c
int binned_smdenorm(const int fold, const float* priX) {
  return biased_exp(priX[0]) - fold * SBWIDTH < FLT_MIN_EXP + FLT_MANT_DIG + (FLT_MAX_EXP - 2) + SBWIDTH;
}
// reference src/binned/ufpf.c (UFPF: the exponent field kept)
float binned_ufpf(const float X) { return fp_from_bits<float>(bits(X) & (0xffu << 23)); }

// reference src/binned/smprint.c
void binned_smprint(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX) {
  for (int i = 0; i < fold; ++i) {
    const float p = priX[(long)i * incpriX], c = carX[(long)i * inccarX];
    const float M = binned_ufpf(p);
    printf("(2^%d: %g #%g =%g)\n", (int)log2f(M) + 1, p - 1.5 * M, c, ((c - 6) * 0.25 * M + p));
  }
}
void binned_sbprint(const int fold, const float_binned* X) { binned_smprint(fold, X, 1, X + fold, 1); }

// reference src/binned/smsmset.c, sbsbset.c, smsetzero.c, sbsetzero.c
void binned_smsmset(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX,
                    float* priY, const int incpriY, float* carY, const int inccarY) {
  for (int i = 0; i < fold; ++i) {
    priY[(long)i * incpriY] = priX[(long)i * incpriX];
    carY[(long)i * inccarY] = carX[(long)i * inccarX];
  }
}
void binned_sbsbset(const int fold, const float_binned* X, float_binned* Y) {
  binned_smsmset(fold, X, 1, X + fold, 1, Y, 1, Y + fold, 1);
}
void binned_smsetzero(const int fold, float* priX, const int incpriX, float* carX, const int inccarX) {
  for (int i = 0; i < fold; ++i) {
    priX[(long)i * incpriX] = 0.0f;
    carX[(long)i * inccarX] = 0.0f;
  }
}
void binned_sbsetzero(const int fold, float_binned* X) { memset(X, 0, binned_sbsbze(fold)); }

// reference src/binned/smsmadd.c, sbsbadd.c
void binned_smsmadd(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX,
                    float* priY, const int incpriY, float* carY, const int inccarY) {
  LocalS x, y;
  gather(fold, priX, incpriX, carX, inccarX, x);
  gather(fold, priY, incpriY, carY, inccarY, y);
  add(fold, y.p, y.c, x.p, x.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
}
void binned_sbsbadd(const int fold, const float_binned* X, float_binned* Y) {
  binned_smsmadd(fold, X, 1, X + fold, 1, Y, 1, Y + fold, 1);
}

// reference src/binned/sbsbaddv.c; device arrays run on the GPU
void binned_sbsbaddv(const int fold, const int N, const float_binned* X, const int incX, float_binned* Y,
                     const int incY) {
  clear_error();
  if (N <= 0) return;
  const bool dx = is_device_pointer(X), dy = is_device_pointer(Y);
  if (dx && dy) {
    hipStream_t s = current_stream();
    RB_STREAM_GUARD(s);
    if (sbsbaddv(fold, N, X, incX, Y, incY, s)) RB_HIP(hipStreamSynchronize(s));
    return;
  }
  if (dx != dy) {
    set_error(kBadArgument, "binned_sbsbaddv: X and Y must both be host or both be device arrays");
    return;
  }
  for (long i = 0; i < N; ++i)
    binned_sbsbadd(fold, X + i * incX * 2 * fold, Y + i * incY * 2 * fold);
}

// reference src/binned/smsupdate.c, sbsupdate.c
void binned_smsupdate(const int fold, const float X, float* priY, const int incpriY, float* carY, const int inccarY) {
  LocalS y;
  gather(fold, priY, incpriY, carY, inccarY, y);
  update(fold, X, y.p, y.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
}
void binned_sbsupdate(const int fold, const float X, float_binned* Y) {
  binned_smsupdate(fold, X, Y, 1, Y + fold, 1);
}

// reference src/binned/smsdeposit.c, sbsdeposit.c
void binned_smsdeposit(const int fold, const float X, float* priY, const int incpriY) {
  if (nan_or_inf(X) || nan_or_inf(priY[0])) {
    priY[0] += X;
    return;
  }
  LocalS y;
  gather(fold, priY, incpriY, priY, incpriY, y);
  deposit(fold, dmindex0(y.p[0]), X, y.p);
  for (int i = 0; i < fold; ++i) priY[(long)i * incpriY] = y.p[i];
}
void binned_sbsdeposit(const int fold, const float X, float_binned* Y) { binned_smsdeposit(fold, X, Y, 1); }

// reference src/binned/smrenorm.c, sbrenorm.c
void binned_smrenorm(const int fold, float* priX, const int incpriX, float* carX, const int inccarX) {
  LocalS x;
  gather(fold, priX, incpriX, carX, inccarX, x);
  renorm(fold, x.p, x.c);
  scatter(fold, x, priX, incpriX, carX, inccarX);
}
void binned_sbrenorm(const int fold, float_binned* X) { binned_smrenorm(fold, X, 1, X + fold, 1); }

// reference src/binned/smsadd.c, sbsadd.c
void binned_smsadd(const int fold, const float X, float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsupdate(fold, X, priY, incpriY, carY, inccarY);
  binned_smsdeposit(fold, X, priY, incpriY);
  binned_smrenorm(fold, priY, incpriY, carY, inccarY);
}
void binned_sbsadd(const int fold, const float X, float_binned* Y) { binned_smsadd(fold, X, Y, 1, Y + fold, 1); }

// reference src/binned/smsconv.c, sbsconv.c
void binned_smsconv(const int fold, const float X, float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsetzero(fold, priY, incpriY, carY, inccarY);
  binned_smsadd(fold, X, priY, incpriY, carY, inccarY);
}
void binned_sbsconv(const int fold, const float X, float_binned* Y) { binned_smsconv(fold, X, Y, 1, Y + fold, 1); }

// reference src/binned/ssmconv.c, ssbconv.c
float binned_ssmconv(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX) {
  LocalS x;
  gather(fold, priX, incpriX, carX, inccarX, x);
  return convert(fold, x.p, x.c);
}
float binned_ssbconv(const int fold, const float_binned* X) { return binned_ssmconv(fold, X, 1, X + fold, 1); }

// reference src/binned/smnegate.c, sbnegate.c
void binned_smnegate(const int fold, float* priX, const int incpriX, float* carX, const int inccarX) {
  if (priX[0] == 0.0f) return;
  const int I = dmindex(priX[0]);
  for (int i = 0; i < fold; ++i) {
    const float b = bin<float>(I + i);
    priX[(long)i * incpriX] = b - (priX[(long)i * incpriX] - b);
    carX[(long)i * inccarX] = -carX[(long)i * inccarX];
  }
}
void binned_sbnegate(const int fold, float_binned* X) { binned_smnegate(fold, X, 1, X + fold, 1); }

// reference src/binned/sscale.c
float binned_sscale(const float X) { return dscale(X); }

// reference src/binned/smsrescale.c
void binned_smsrescale(const int fold, const float X, const float scaleY, float* priY, const int incpriY, float* carY,
                       const int inccarY) {
  LocalS y;
  gather(fold, priY, incpriY, carY, inccarY, y);
  rescale(fold, X, scaleY, y.p, y.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
}

// reference src/binned/smsmaddsq.c, sbsbaddsq.c
float binned_smsmaddsq(const int fold, const float scaleX, const float* priX, const int incpriX, const float* carX,
                       const int inccarX, const float scaleY, float* priY, const int incpriY, float* carY,
                       const int inccarY) {
  LocalS x, y;
  gather(fold, priX, incpriX, carX, inccarX, x);
  gather(fold, priY, incpriY, carY, inccarY, y);
  const float sc = add_scaled(fold, scaleX, x.p, x.c, scaleY, y.p, y.c);
  scatter(fold, y, priY, incpriY, carY, inccarY);
  return sc;
}
float binned_sbsbaddsq(const int fold, const float scaleX, const float_binned* X, const float scaleY,
                       float_binned* Y) {
  return binned_smsmaddsq(fold, scaleX, X, 1, X + fold, 1, scaleY, Y, 1, Y + fold, 1);
}

}  // extern "C"
