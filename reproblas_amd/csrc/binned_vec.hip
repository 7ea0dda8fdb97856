// binned_vec.hip -- elementwise kernels over device arrays of binned values
// (double_binned and float_binned): the vector add behind binned_dbdbaddv /
// binned_sbsbaddv (reference src/binned/dbdbaddv.c:18-23, sbsbaddv.c),
// scalar -> binned (binned_dbdconv / sbsconv, src/binned/dmdconv.c:18-21,
// with the beta handling of src/reproBLAS/rdgemv.c:50-61 and rdgemm.c:80-95)
// and binned -> scalar (binned_ddbconv / ssbconv, src/binned/ddmconv.c:23-81,
// ssmconv.c).
#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rb {

namespace {

template <int CAP, bool FIXED, class F>
__global__ __launch_bounds__(256) void addv_kernel(int fold_rt, int64_t n, const F* X, int64_t incX, F* Y,
                                                   int64_t incY) {
  const int fold = FIXED ? CAP : fold_rt;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const F* x = X + i * incX * 2 * fold;
  F* y = Y + i * incY * 2 * fold;
  F px[CAP], cx[CAP], py[CAP], cy[CAP];
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

template <int CAP, bool FIXED, class F>
__global__ __launch_bounds__(256) void to_binned_kernel(int fold_rt, int64_t n, const F* v, int64_t incv, F beta,
                                                        F* Yb) {
  const int fold = FIXED ? CAP : fold_rt;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  F p[CAP], c[CAP];
  set_zero(fold, p, c);
  if (beta != F(0)) {
    const F x = (beta == F(1)) ? v[i * incv] : v[i * incv] * beta;
    // binned_dmdadd: update, deposit (with its NaN/Inf shortcut), renorm
    update(fold, x, p, c);
    if (nan_or_inf(x) || nan_or_inf(p[0])) {
      p[0] += x;
    } else {
      deposit(fold, dmindex0(p[0]), x, p);
    }
    renorm(fold, p, c);
  }
  F* y = Yb + i * 2 * fold;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      y[j] = p[j];
      y[fold + j] = c[j];
    }
}

template <int CAP, bool FIXED, class F>
__global__ __launch_bounds__(256) void from_binned_kernel(int fold_rt, int64_t n, const F* Yb, F* v, int64_t incv) {
  const int fold = FIXED ? CAP : fold_rt;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const F* y = Yb + i * 2 * fold;
  F p[CAP], c[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      p[j] = y[j];
      c[j] = y[fold + j];
    }
  v[i * incv] = convert(fold, p, c);
}

// y(i, o) = conv(binned(beta * y(i, o))) in place for an inner x outer array,
// element (i, o) at y[i*inc_i + o*inc_o]: the whole result of
// reproBLAS_rdgemv / rdgemm when no product is added (alpha == 0 or K == 0;
// src/reproBLAS/rdgemm.c:56-103, rdgemv.c:48-84). beta == 0 gives +0.
template <int CAP, bool FIXED, class F>
__global__ __launch_bounds__(256) void beta_only_kernel(int fold_rt, int64_t n, int64_t inner, F* y, int64_t inc_i,
                                                        int64_t inc_o, F beta) {
  const int fold = FIXED ? CAP : fold_rt;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    F* e = y + (t % inner) * inc_i + (t / inner) * inc_o;
    F p[CAP], c[CAP];
    set_zero(fold, p, c);
    if (beta != F(0)) {
      const F x = (beta == F(1)) ? *e : *e * beta;
      update(fold, x, p, c);
      if (nan_or_inf(x) || nan_or_inf(p[0])) {
        p[0] += x;
      } else {
        deposit(fold, dmindex0(p[0]), x, p);
      }
      renorm(fold, p, c);
    }
    *e = convert(fold, p, c);
  }
}

// Complex binned values: both components, interleaved (4*fold per value).
template <int CAP, bool FIXED, class F>
__global__ __launch_bounds__(256) void zaddv_kernel(int fold_rt, int64_t n, const F* X, int64_t incX, F* Y,
                                                    int64_t incY) {
  const int fold = FIXED ? CAP : fold_rt;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const F* x = X + i * incX * 4 * fold;
  F* y = Y + i * incY * 4 * fold;
  for (int ch = 0; ch < 2; ++ch) {
    F px[CAP], cx[CAP], py[CAP], cy[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) {
        px[j] = x[2 * j + ch];
        cx[j] = x[2 * fold + 2 * j + ch];
        py[j] = y[2 * j + ch];
        cy[j] = y[2 * fold + 2 * j + ch];
      }
    add(fold, py, cy, px, cx);
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) {
        y[2 * j + ch] = py[j];
        y[2 * fold + 2 * j + ch] = cy[j];
      }
  }
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

template <class F>
bool fold_ok(int fold) {
  if (fold >= 2 && fold <= Fmt<F>::max_fold) return true;
  set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
  return false;
}

}  // namespace

#define RB_FOLD_DISPATCH(KERNEL, F, ...)                                                      \
  switch (fold) {                                                                             \
    case 2: KERNEL<2, true, F><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__); break;         \
    case 3: KERNEL<3, true, F><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__); break;         \
    case 4: KERNEL<4, true, F><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__); break;         \
    default:                                                                                  \
      if (fold <= 8)                                                                          \
        KERNEL<8, false, F><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__);                   \
      else                                                                                    \
        KERNEL<Fmt<F>::max_fold, false, F><<<grid_for(n), 256, 0, s>>>(fold, __VA_ARGS__);    \
      break;                                                                                  \
  }

bool dbdbaddv(int fold, int64_t n, const double* X, int64_t incX, double* Y, int64_t incY, hipStream_t s) {
  if (!fold_ok<double>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(addv_kernel, double, n, X, incX, Y, incY);
  return RB_HIP(hipGetLastError());
}

bool sbsbaddv(int fold, int64_t n, const float* X, int64_t incX, float* Y, int64_t incY, hipStream_t s) {
  if (!fold_ok<float>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(addv_kernel, float, n, X, incX, Y, incY);
  return RB_HIP(hipGetLastError());
}

bool zaddv(int fold, int64_t n, const double* X, int64_t incX, double* Y, int64_t incY, hipStream_t s) {
  if (!fold_ok<double>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(zaddv_kernel, double, n, X, incX, Y, incY);
  return RB_HIP(hipGetLastError());
}

bool zaddv(int fold, int64_t n, const float* X, int64_t incX, float* Y, int64_t incY, hipStream_t s) {
  if (!fold_ok<float>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(zaddv_kernel, float, n, X, incX, Y, incY);
  return RB_HIP(hipGetLastError());
}

bool conv_to_binned(int fold, int64_t n, const double* y, int64_t incy, double beta, double* Yb, hipStream_t s) {
  if (!fold_ok<double>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(to_binned_kernel, double, n, y, incy, beta, Yb);
  return RB_HIP(hipGetLastError());
}

bool conv_to_binned(int fold, int64_t n, const float* y, int64_t incy, float beta, float* Yb, hipStream_t s) {
  if (!fold_ok<float>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(to_binned_kernel, float, n, y, incy, beta, Yb);
  return RB_HIP(hipGetLastError());
}

bool conv_from_binned(int fold, int64_t n, const double* Yb, double* y, int64_t incy, hipStream_t s) {
  if (!fold_ok<double>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(from_binned_kernel, double, n, Yb, y, incy);
  return RB_HIP(hipGetLastError());
}

bool conv_from_binned(int fold, int64_t n, const float* Yb, float* y, int64_t incy, hipStream_t s) {
  if (!fold_ok<float>(fold)) return false;
  if (n <= 0) return true;
  RB_FOLD_DISPATCH(from_binned_kernel, float, n, Yb, y, incy);
  return RB_HIP(hipGetLastError());
}

bool conv_to_double(int fold, int64_t n, const double* Yb, double* y, int64_t incy, hipStream_t s) {
  return conv_from_binned(fold, n, Yb, y, incy, s);
}

namespace {
template <class F>
bool beta_only_impl(int fold, int64_t inner, int64_t outer, F* y, int64_t inc_i, int64_t inc_o, F beta,
                    hipStream_t s) {
  if (!fold_ok<F>(fold)) return false;
  if (inner <= 0 || outer <= 0) return true;
  const int64_t n = inner * outer;
  const unsigned g = grid_for(n) < 65536u ? grid_for(n) : 65536u;  // grid-stride beyond
#define RB_BETA_LAUNCH(CAP, FIXED) \
  beta_only_kernel<CAP, FIXED, F><<<g, 256, 0, s>>>(fold, n, inner, y, inc_i, inc_o, beta)
  switch (fold) {
    case 2: RB_BETA_LAUNCH(2, true); break;
    case 3: RB_BETA_LAUNCH(3, true); break;
    case 4: RB_BETA_LAUNCH(4, true); break;
    default:
      if (fold <= 8)
        RB_BETA_LAUNCH(8, false);
      else
        RB_BETA_LAUNCH(Fmt<F>::max_fold, false);
      break;
  }
#undef RB_BETA_LAUNCH
  return RB_HIP(hipGetLastError());
}
}  // namespace

bool beta_only(int fold, int64_t inner, int64_t outer, double* y, int64_t inc_i, int64_t inc_o, double beta,
               hipStream_t s) {
  return beta_only_impl<double>(fold, inner, outer, y, inc_i, inc_o, beta, s);
}
bool beta_only(int fold, int64_t inner, int64_t outer, float* y, int64_t inc_i, int64_t inc_o, float beta,
               hipStream_t s) {
  return beta_only_impl<float>(fold, inner, outer, y, inc_i, inc_o, beta, s);
}

}  // namespace rb
