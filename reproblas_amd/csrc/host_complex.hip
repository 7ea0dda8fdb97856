// host_complex.hip -- the complex binned primitives of include/binned.h
// (double_complex_binned, float_complex_binned) on host memory. Each applies
// the real primitive of its precision to both components with doubled
// strides, as the reference's src/binned/z*.c and c*.c do; the few places
// where the reference deviates from that (zzmconv_sub's carry stride,
// zmdrescale's zero test) are reproduced as written there.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/binned.h"
#include "kernels.hpp"
#include "runtime.hpp"

extern "C" {

// ---------------------------------------------------------------- double complex (double) --
// reference src/binned/zbsize.c, zballoc.c, zbnum.c
size_t binned_zbsize(const int fold) { return 4 * (size_t)fold * sizeof(double); }
double_complex_binned* binned_zballoc(const int fold) { return (double_complex_binned*)malloc(binned_zbsize(fold)); }
int binned_zbnum(const int fold) { return 4 * fold; }
// reference src/binned/zmdenorm.c (its second test reads priX + 1 with the real stride)
int binned_zmdenorm(const int fold, const double* priX) {
  return binned_dmdenorm(fold, priX) || binned_dmdenorm(fold, priX + 1);
}
// reference src/binned/zmprint.c, zbprint.c
void binned_zmprint(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX) {
  for (int i = 0; i < fold; i++, priX += incpriX, carX += inccarX) {
    double M = binned_ufp(priX[0]);
    printf("(2^%d: %g #%g =%g", (int)log2(M) + 1, (double)(priX[0] - 1.5 * M), (double)carX[0],
           (double)((carX[0] - 6) * 0.25 * M + priX[0]));
    M = binned_ufp(priX[1]);
    printf("| 2^%d: %g #%g =%g)\n", (int)log2(M) + 1, (double)(priX[1] - 1.5 * M), (double)carX[1],
           (double)((carX[1] - 6) * 0.25 * M + priX[1]));
  }
}
void binned_zbprint(const int fold, const double_complex_binned* X) { binned_zmprint(fold, X, 2, X + 2 * fold, 2); }
// reference src/binned/zmzmset.c, zbzbset.c, zmdmset.c, zbdbset.c
void binned_zmzmset(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX,
                    double* priY, const int incpriY, double* carY, const int inccarY) {
  binned_dmdmset(fold, priX, 2 * incpriX, carX, 2 * inccarX, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_dmdmset(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_zbzbset(const int fold, const double_complex_binned* X, double_complex_binned* Y) {
  binned_zmzmset(fold, X, 1, X + 2 * fold, 1, Y, 1, Y + 2 * fold, 1);
}
void binned_zmdmset(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX,
                    double* priY, const int incpriY, double* carY, const int inccarY) {
  binned_dmdmset(fold, priX, incpriX, carX, inccarX, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_dmsetzero(fold, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_zbdbset(const int fold, const double_binned* X, double_complex_binned* Y) {
  binned_zmdmset(fold, X, 1, X + fold, 1, Y, 1, Y + 2 * fold, 1);
}
// reference src/binned/zmsetzero.c, zbsetzero.c
void binned_zmsetzero(const int fold, double* priX, const int incpriX, double* carX, const int inccarX) {
  binned_dmsetzero(fold, priX, 2 * incpriX, carX, 2 * inccarX);
  binned_dmsetzero(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX);
}
void binned_zbsetzero(const int fold, double_complex_binned* X) { memset(X, 0, binned_zbsize(fold)); }
// reference src/binned/zmzmadd.c, zbzbadd.c
void binned_zmzmadd(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX,
                    double* priY, const int incpriY, double* carY, const int inccarY) {
  binned_dmdmadd(fold, priX, 2 * incpriX, carX, 2 * inccarX, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_dmdmadd(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_zbzbadd(const int fold, const double_complex_binned* X, double_complex_binned* Y) {
  binned_zmzmadd(fold, X, 1, X + 2 * fold, 1, Y, 1, Y + 2 * fold, 1);
}
// reference src/binned/zbzbaddv.c: host arrays element by element; device
// arrays in one kernel
void binned_zbzbaddv(const int fold, const int N, const double_complex_binned* X, const int incX, double_complex_binned* Y, const int incY) {
  if (N <= 0) return;
  if (rb::is_device_pointer(X) && rb::is_device_pointer(Y)) {
    rb::clear_error();
    hipStream_t s = rb::current_stream();
    RB_STREAM_GUARD(s);
    if (rb::zaddv(fold, N, X, incX, Y, incY, s)) RB_HIP(hipStreamSynchronize(s));
    return;
  }
  for (int i = 0; i < N; i++)
    binned_zbzbadd(fold, X + (long)i * incX * 4 * fold, Y + (long)i * incY * 4 * fold);
}
// reference src/binned/zmzupdate.c, zbzupdate.c, zmdupdate.c, zbdupdate.c
void binned_zmzupdate(const int fold, const void* X, double* priY, const int incpriY, double* carY, const int inccarY) {
  binned_dmdupdate(fold, ((const double*)X)[0], priY, 2 * incpriY, carY, 2 * inccarY);
  binned_dmdupdate(fold, ((const double*)X)[1], priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_zbzupdate(const int fold, const void* X, double_complex_binned* Y) {
  binned_zmzupdate(fold, X, Y, 1, Y + 2 * fold, 1);
}
void binned_zmdupdate(const int fold, const double X, double* priY, const int incpriY, double* carY, const int inccarY) {
  binned_dmdupdate(fold, X, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_dmdupdate(fold, X, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_zbdupdate(const int fold, const double X, double_complex_binned* Y) {
  binned_zmdupdate(fold, X, Y, 1, Y + 2 * fold, 1);
}
// reference src/binned/zmzdeposit.c: with its non-finite short cuts the
// component loops deposit exactly as the real deposit does per component
void binned_zmzdeposit(const int fold, const void* X, double* priY, const int incpriY) {
  binned_dmddeposit(fold, ((const double*)X)[0], priY, 2 * incpriY);
  binned_dmddeposit(fold, ((const double*)X)[1], priY + 1, 2 * incpriY);
}
void binned_zbzdeposit(const int fold, const void* X, double_complex_binned* Y) { binned_zmzdeposit(fold, X, Y, 1); }
// reference src/binned/zmrenorm.c, zbrenorm.c
void binned_zmrenorm(const int fold, double* priX, const int incpriX, double* carX, const int inccarX) {
  binned_dmrenorm(fold, priX, 2 * incpriX, carX, 2 * inccarX);
  binned_dmrenorm(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX);
}
void binned_zbrenorm(const int fold, double_complex_binned* X) { binned_zmrenorm(fold, X, 1, X + 2 * fold, 1); }
// reference src/binned/zmzadd.c, zbzadd.c
void binned_zmzadd(const int fold, const void* X, double* priY, const int incpriY, double* carY, const int inccarY) {
  binned_zmzupdate(fold, X, priY, incpriY, carY, inccarY);
  binned_zmzdeposit(fold, X, priY, incpriY);
  binned_zmrenorm(fold, priY, incpriY, carY, inccarY);
}
void binned_zbzadd(const int fold, const void* X, double_complex_binned* Y) { binned_zmzadd(fold, X, Y, 1, Y + 2 * fold, 1); }
// reference src/binned/zmzconv.c, zbzconv.c
void binned_zmzconv(const int fold, const void* X, double* priY, const int incpriY, double* carY, const int inccarY) {
  binned_dmdconv(fold, ((const double*)X)[0], priY, 2 * incpriY, carY, 2 * inccarY);
  binned_dmdconv(fold, ((const double*)X)[1], priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_zbzconv(const int fold, const void* X, double_complex_binned* Y) { binned_zmzconv(fold, X, Y, 1, Y + 2 * fold, 1); }
// reference src/binned/zzmconv_sub.c (carries read with stride inccarX + 1,
// as the reference reads them), zzbconv_sub.c
void binned_zzmconv_sub(const int fold, const double* priX, const int incpriX, const double* carX, const int inccarX,
                        void* conv) {
  ((double*)conv)[0] = binned_ddmconv(fold, priX, 2 * incpriX, carX, inccarX + 1);
  ((double*)conv)[1] = binned_ddmconv(fold, priX + 1, 2 * incpriX, carX + 1, inccarX + 1);
}
void binned_zzbconv_sub(const int fold, const double_complex_binned* X, void* conv) {
  binned_zzmconv_sub(fold, X, 1, X + 2 * fold, 1, conv);
}
// reference src/binned/zmnegate.c, zbnegate.c
void binned_zmnegate(const int fold, double* priX, const int incpriX, double* carX, const int inccarX) {
  binned_dmnegate(fold, priX, 2 * incpriX, carX, 2 * inccarX);
  binned_dmnegate(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX);
}
void binned_zbnegate(const int fold, double_complex_binned* X) { binned_zmnegate(fold, X, 1, X + 2 * fold, 1); }
// reference src/binned/zmdrescale.c, including its zero test on
// priY[i * incpriY] and the update it then applies
void binned_zmdrescale(const int fold, const double X, const double scaleY, double* priY, const int incpriY, double* carY,
                       const int inccarY) {
  if (X == scaleY || X == 0.0 || scaleY == 0.0) return;
  double r = X / scaleY;
  r *= r;
  for (int i = 0; i < fold; i++) {
    priY[i * 2 * incpriY] /= r;
    priY[i * 2 * incpriY + 1] /= r;
    if (priY[i * incpriY] == 0.0) {
      binned_zmdupdate(fold - i, 0.0, priY + i * 2 * incpriY, incpriY, carY + i * 2 * inccarY, inccarY);
      return;
    }
  }
}

// ---------------------------------------------------------------- float complex (float) --
// reference src/binned/cbsize.c, cballoc.c, cbnum.c
size_t binned_cbsize(const int fold) { return 4 * (size_t)fold * sizeof(float); }
float_complex_binned* binned_cballoc(const int fold) { return (float_complex_binned*)malloc(binned_cbsize(fold)); }
int binned_cbnum(const int fold) { return 4 * fold; }
// reference src/binned/cmdenorm.c (its second test reads priX + 1 with the real stride)
int binned_cmdenorm(const int fold, const float* priX) {
  return binned_smdenorm(fold, priX) || binned_smdenorm(fold, priX + 1);
}
// reference src/binned/cmprint.c, cbprint.c
void binned_cmprint(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX) {
  for (int i = 0; i < fold; i++, priX += incpriX, carX += inccarX) {
    double M = binned_ufpf(priX[0]);
    printf("(2^%d: %g #%g =%g", (int)log2(M) + 1, (double)(priX[0] - 1.5 * M), (double)carX[0],
           (double)((carX[0] - 6) * 0.25 * M + priX[0]));
    M = binned_ufpf(priX[1]);
    printf("| 2^%d: %g #%g =%g)\n", (int)log2(M) + 1, (double)(priX[1] - 1.5 * M), (double)carX[1],
           (double)((carX[1] - 6) * 0.25 * M + priX[1]));
  }
}
void binned_cbprint(const int fold, const float_complex_binned* X) { binned_cmprint(fold, X, 2, X + 2 * fold, 2); }
// reference src/binned/cmcmset.c, cbcbset.c, cmsmset.c, cbsbset.c
void binned_cmcmset(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX,
                    float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsmset(fold, priX, 2 * incpriX, carX, 2 * inccarX, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_smsmset(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_cbcbset(const int fold, const float_complex_binned* X, float_complex_binned* Y) {
  binned_cmcmset(fold, X, 1, X + 2 * fold, 1, Y, 1, Y + 2 * fold, 1);
}
void binned_cmsmset(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX,
                    float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsmset(fold, priX, incpriX, carX, inccarX, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_smsetzero(fold, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_cbsbset(const int fold, const float_binned* X, float_complex_binned* Y) {
  binned_cmsmset(fold, X, 1, X + fold, 1, Y, 1, Y + 2 * fold, 1);
}
// reference src/binned/cmsetzero.c, cbsetzero.c
void binned_cmsetzero(const int fold, float* priX, const int incpriX, float* carX, const int inccarX) {
  binned_smsetzero(fold, priX, 2 * incpriX, carX, 2 * inccarX);
  binned_smsetzero(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX);
}
void binned_cbsetzero(const int fold, float_complex_binned* X) { memset(X, 0, binned_cbsize(fold)); }
// reference src/binned/cmcmadd.c, cbcbadd.c
void binned_cmcmadd(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX,
                    float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsmadd(fold, priX, 2 * incpriX, carX, 2 * inccarX, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_smsmadd(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_cbcbadd(const int fold, const float_complex_binned* X, float_complex_binned* Y) {
  binned_cmcmadd(fold, X, 1, X + 2 * fold, 1, Y, 1, Y + 2 * fold, 1);
}
// reference src/binned/cbcbaddv.c: host arrays element by element; device
// arrays in one kernel
void binned_cbcbaddv(const int fold, const int N, const float_complex_binned* X, const int incX, float_complex_binned* Y, const int incY) {
  if (N <= 0) return;
  if (rb::is_device_pointer(X) && rb::is_device_pointer(Y)) {
    rb::clear_error();
    hipStream_t s = rb::current_stream();
    RB_STREAM_GUARD(s);
    if (rb::zaddv(fold, N, X, incX, Y, incY, s)) RB_HIP(hipStreamSynchronize(s));
    return;
  }
  for (int i = 0; i < N; i++)
    binned_cbcbadd(fold, X + (long)i * incX * 4 * fold, Y + (long)i * incY * 4 * fold);
}
// reference src/binned/cmcupdate.c, cbcupdate.c, cmsupdate.c, cbsupdate.c
void binned_cmcupdate(const int fold, const void* X, float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsupdate(fold, ((const float*)X)[0], priY, 2 * incpriY, carY, 2 * inccarY);
  binned_smsupdate(fold, ((const float*)X)[1], priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_cbcupdate(const int fold, const void* X, float_complex_binned* Y) {
  binned_cmcupdate(fold, X, Y, 1, Y + 2 * fold, 1);
}
void binned_cmsupdate(const int fold, const float X, float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsupdate(fold, X, priY, 2 * incpriY, carY, 2 * inccarY);
  binned_smsupdate(fold, X, priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_cbsupdate(const int fold, const float X, float_complex_binned* Y) {
  binned_cmsupdate(fold, X, Y, 1, Y + 2 * fold, 1);
}
// reference src/binned/cmcdeposit.c: with its non-finite short cuts the
// component loops deposit exactly as the real deposit does per component
void binned_cmcdeposit(const int fold, const void* X, float* priY, const int incpriY) {
  binned_smsdeposit(fold, ((const float*)X)[0], priY, 2 * incpriY);
  binned_smsdeposit(fold, ((const float*)X)[1], priY + 1, 2 * incpriY);
}
void binned_cbcdeposit(const int fold, const void* X, float_complex_binned* Y) { binned_cmcdeposit(fold, X, Y, 1); }
// reference src/binned/cmrenorm.c, cbrenorm.c
void binned_cmrenorm(const int fold, float* priX, const int incpriX, float* carX, const int inccarX) {
  binned_smrenorm(fold, priX, 2 * incpriX, carX, 2 * inccarX);
  binned_smrenorm(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX);
}
void binned_cbrenorm(const int fold, float_complex_binned* X) { binned_cmrenorm(fold, X, 1, X + 2 * fold, 1); }
// reference src/binned/cmcadd.c, cbcadd.c
void binned_cmcadd(const int fold, const void* X, float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_cmcupdate(fold, X, priY, incpriY, carY, inccarY);
  binned_cmcdeposit(fold, X, priY, incpriY);
  binned_cmrenorm(fold, priY, incpriY, carY, inccarY);
}
void binned_cbcadd(const int fold, const void* X, float_complex_binned* Y) { binned_cmcadd(fold, X, Y, 1, Y + 2 * fold, 1); }
// reference src/binned/cmcconv.c, cbcconv.c
void binned_cmcconv(const int fold, const void* X, float* priY, const int incpriY, float* carY, const int inccarY) {
  binned_smsconv(fold, ((const float*)X)[0], priY, 2 * incpriY, carY, 2 * inccarY);
  binned_smsconv(fold, ((const float*)X)[1], priY + 1, 2 * incpriY, carY + 1, 2 * inccarY);
}
void binned_cbcconv(const int fold, const void* X, float_complex_binned* Y) { binned_cmcconv(fold, X, Y, 1, Y + 2 * fold, 1); }
// reference src/binned/ccmconv_sub.c (carries read with stride inccarX + 1,
// as the reference reads them), ccbconv_sub.c
void binned_ccmconv_sub(const int fold, const float* priX, const int incpriX, const float* carX, const int inccarX,
                        void* conv) {
  ((float*)conv)[0] = binned_ssmconv(fold, priX, 2 * incpriX, carX, inccarX + 1);
  ((float*)conv)[1] = binned_ssmconv(fold, priX + 1, 2 * incpriX, carX + 1, inccarX + 1);
}
void binned_ccbconv_sub(const int fold, const float_complex_binned* X, void* conv) {
  binned_ccmconv_sub(fold, X, 1, X + 2 * fold, 1, conv);
}
// reference src/binned/cmnegate.c, cbnegate.c
void binned_cmnegate(const int fold, float* priX, const int incpriX, float* carX, const int inccarX) {
  binned_smnegate(fold, priX, 2 * incpriX, carX, 2 * inccarX);
  binned_smnegate(fold, priX + 1, 2 * incpriX, carX + 1, 2 * inccarX);
}
void binned_cbnegate(const int fold, float_complex_binned* X) { binned_cmnegate(fold, X, 1, X + 2 * fold, 1); }
// reference src/binned/cmsrescale.c, including its zero test on
// priY[i * incpriY] and the update it then applies
void binned_cmsrescale(const int fold, const float X, const float scaleY, float* priY, const int incpriY, float* carY,
                       const int inccarY) {
  if (X == scaleY || X == 0.0 || scaleY == 0.0) return;
  float r = X / scaleY;
  r *= r;
  for (int i = 0; i < fold; i++) {
    priY[i * 2 * incpriY] /= r;
    priY[i * 2 * incpriY + 1] /= r;
    if (priY[i * incpriY] == 0.0) {
      binned_cmsupdate(fold - i, 0.0, priY + i * 2 * incpriY, incpriY, carY + i * 2 * inccarY, inccarY);
      return;
    }
  }
}

}  // extern "C"
