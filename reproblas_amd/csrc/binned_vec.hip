// binned_vec.hip -- elementwise kernels over device arrays of binned values:
// the vector add behind binned_dbdbaddv (reference src/binned/dbdbaddv.c:18-23),
// double -> binned (binned_dbdconv, src/binned/dmdconv.c:18-21, with the beta
// handling of src/reproBLAS/rdgemv.c:50-61 and rdgemm.c:80-95) and binned ->
// double (binned_ddbconv, src/binned/ddmconv.c:23-81).
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rb {

namespace {

template <int CAP, bool FIXED>
__global__ __launch_bounds__(256) void addv_kernel(int fold_rt, int64_t n, const double* X, int64_t incX, double* Y,
                                                   int64_t incY) {
  const int fold = FIXED ? CAP : fold_rt;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* x = X + i * incX * 2 * fold;
  double* y = Y + i * incY * 2 * fold;
  double px[CAP], cx[CAP], py[CAP], cy[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      px[j] = x[j];
      cx[j] = x[fold + j];
      py[j] = y[j];
      cy[j] = y[fold + j];
    }
  add(fold, py, cy, px, cx);
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      y[j] = py[j];
      y[fold + j] = cy[j];
    }
}

template <int CAP, bool FIXED>
__global__ __launch_bounds__(256) void to_binned_kernel(int fold_rt, int64_t n, const double* v, int64_t incv,
                                                        double beta, double* Yb) {
  const int fold = FIXED ? CAP : fold_rt;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[CAP], c[CAP];
  set_zero(fold, p, c);
  if (beta != 0.0) {
    const double x = (beta == 1.0) ? v[i * incv] : v[i * incv] * beta;
    // binned_dmdadd: update, deposit (with its NaN/Inf shortcut), renorm
    update(fold, x, p, c);
    if (nan_or_inf(x) || nan_or_inf(p[0])) {
      p[0] += x;
    } else {
      deposit(fold, dmindex0(p[0]), x, p);
    }
    renorm(fold, p, c);
  }
  double* y = Yb + i * 2 * fold;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      y[j] = p[j];
      y[fold + j] = c[j];
    }
}

template <int CAP, bool FIXED>
__global__ __launch_bounds__(256) void to_double_kernel(int fold_rt, int64_t n, const double* Yb, double* v,
                                                        int64_t incv) {
  const int fold = FIXED ? CAP : fold_rt;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* y = Yb + i * 2 * fold;
  double p[CAP], c[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      p[j] = y[j];
      c[j] = y[fold + j];
    }
  v[i * incv] = to_double(fold, p, c);
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

#define RB_FOLD_DISPATCH(KERNEL, ...)                                                         \
  switch (fold) {                                                                             \
    case 2: KERNEL<2, true><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__); break;            \
    case 3: KERNEL<3, true><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__); break;            \
    case 4: KERNEL<4, true><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__); break;            \
    default:                                                                                  \
      if (fold <= 8)                                                                          \
        KERNEL<8, false><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__);                      \
      else                                                                                    \
        KERNEL<kMaxFold, false><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__);               \
      break;                                                                                  \
  }

bool dbdbaddv(int fold, int64_t n, const double* X, int64_t incX, double* Y, int64_t incY, hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(addv_kernel, n, X, incX, Y, incY);
  return RB_HIP(hipGetLastError());
}

bool conv_to_binned(int fold, int64_t n, const double* y, int64_t incy, double beta, double* Yb, hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(to_binned_kernel, n, y, incy, beta, Yb);
  return RB_HIP(hipGetLastError());
}

bool conv_to_double(int fold, int64_t n, const double* Yb, double* y, int64_t incy, hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(to_double_kernel, n, Yb, y, incy);
  return RB_HIP(hipGetLastError());
}

}  // namespace rb
