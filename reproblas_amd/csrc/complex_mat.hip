// complex_mat.hip -- complex binned gemv and gemm on device operands
// (binnedBLAS_zbzgemv / zbzgemm / cbcgemv / cbcgemm and reproBLAS_rzgemv /
// rzgemm / rcgemv / rcgemm semantics) through the real binned kernels.
//
// Reference path: src/binnedBLAS/zbzgemv.ccog and zbzgemm.ccog (one binned
// complex dotu / dotc per output over op(A) and alpha*x, resp. alpha*op(A)
// and op(B)), src/reproBLAS/rzgemv.c and rzgemm.c (beta*y / beta*C by the
// complex product, conversion per component).
//
// Both components of an output are real binned sums over real products:
//   re = sum a_r*b_r + a_i*(-b_i),   im = sum a_r*b_i + a_i*b_r
// (conjugation flips the sign of one factor's imaginary part). With the
// real image RA of op(A) (each complex row interleaved: a_r, a_i, ...) and
// the real image RB of op(B) ((b_r, -b_i) stacked over (b_i, b_r) per k),
// the complex product is one real binned gemm of M x 2K by 2K x 2N whose
// outputs are the real (columns 0..N-1) and imaginary (N..2N-1) parts, each
// a binned sum over exactly the reference's terms: the real kernels' bits
// equal the reference's. gemv reads A once: with contiguous rows of op(A) a
// complex row-dot kernel (complex_impl.hpp, zgemv_rows), with terms lda apart the
// complex mode of the column gemv (gemv.hip, zgemv_cols).
#include <algorithm>

#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rb {

namespace {

constexpr int kT = 256;
inline unsigned grid1(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

inline bool is_n(char t) { return t == 'n' || t == 'N'; }
inline bool is_t(char t) { return t == 't' || t == 'T'; }
inline bool row_major(char o) { return o == 'r' || o == 'R'; }

// How alpha (or beta) multiplies: 0 -- it is 1 (operand used as is);
// 1 -- the complex product (v_r*a_r - v_i*a_i, v_r*a_i + v_i*a_r);
// 2 -- the reference's column-major transposed-A expression of zbzgemm.ccog
//      (imaginary part v_r*a_i + v_i*a_i; for the conjugate branch
//      (v_r*a_r + v_i*a_i, v_r*a_i - v_i*a_i) on the stored element).
template <class F>
__device__ __forceinline__ void times_scalar(int mode, F vr, F vi, F ar, F ai, F& outr, F& outi) {
  if (mode == 0) {
    outr = vr;
    outi = vi;
  } else {
    outr = vr * ar - vi * ai;
    outi = vr * ai + vi * ar;
  }
}

// RA (M x 2K real, row-major): alpha * op(A)_ik, interleaved.
template <class F>
__global__ __launch_bounds__(kT) void prep_a_kernel(int64_t M, int64_t K, int rm, int trans, const F* A, int64_t lda,
                                                    F al_r, F al_i, int mode, F* RA) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= M * K) return;
  const int64_t i = e / K, k = e % K;
  // op(A)_ik is stored at (i, k) for 'N', at (k, i) otherwise
  const int64_t r = trans == 0 ? i : k, c = trans == 0 ? k : i;
  const int64_t off = rm ? r * lda + c : c * lda + r;
  const F vr = A[2 * off], vi = A[2 * off + 1];
  F outr, outi;
  if (mode == 2) {
    if (trans == 1) {  // 'T'
      outr = vr * al_r - vi * al_i;
      outi = vr * al_i + vi * al_i;
    } else {  // conjugate
      outr = vr * al_r + vi * al_i;
      outi = vr * al_i - vi * al_i;
    }
  } else {
    const F ci = trans == 2 ? -vi : vi;  // conj(A) for the conjugate transpose
    times_scalar(mode, vr, ci, al_r, al_i, outr, outi);
  }
  RA[i * 2 * K + 2 * k] = outr;
  RA[i * 2 * K + 2 * k + 1] = outi;
}

// RB (2K x 2N real, row-major): rows 2k, 2k+1 = (b_r, -b_i | b_i, b_r) of op(B)_k.
template <class F>
__global__ __launch_bounds__(kT) void prep_b_kernel(int64_t K, int64_t N, int rm, int trans, const F* B, int64_t ldb,
                                                    F* RB) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= K * N) return;
  const int64_t k = e / N, j = e % N;
  const int64_t r = trans == 0 ? k : j, c = trans == 0 ? j : k;
  const int64_t off = rm ? r * ldb + c : c * ldb + r;
  const F br = B[2 * off];
  const F bi = trans == 2 ? -B[2 * off + 1] : B[2 * off + 1];
  F* r0 = RB + (2 * k) * (2 * N);
  F* r1 = RB + (2 * k + 1) * (2 * N);
  r0[j] = br;
  r1[j] = -bi;
  r0[N + j] = bi;
  r1[N + j] = br;
}

// Complex binned element (i, j) of a rows x cols array with leading
// dimension ld (in complex binned elements, storage order rm) <-> its two
// components in real binned arrays R[(i*ldr + j)] and R[(i*ldr + off1 + j)].
template <class F>
__device__ __forceinline__ int64_t zoff(int64_t i, int64_t j, int64_t ld, int rm) {
  return rm ? i * ld + j : j * ld + i;
}

template <class F>
__global__ __launch_bounds__(kT) void split_kernel(int fold, int64_t rows, int64_t cols, const F* Z, int64_t ld,
                                                   int rm, F* R, int64_t ldr, int64_t off1) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t i = e / cols, j = e % cols;
  const F* z = Z + zoff<F>(i, j, ld, rm) * 4 * fold;
  F* r0 = R + (i * ldr + j) * 2 * fold;
  F* r1 = R + (i * ldr + off1 + j) * 2 * fold;
  for (int q = 0; q < fold; ++q) {
    r0[q] = z[2 * q];
    r1[q] = z[2 * q + 1];
    r0[fold + q] = z[2 * fold + 2 * q];
    r1[fold + q] = z[2 * fold + 2 * q + 1];
  }
}

template <class F>
__global__ __launch_bounds__(kT) void join_kernel(int fold, int64_t rows, int64_t cols, const F* R, int64_t ldr,
                                                  int64_t off1, F* Z, int64_t ld, int rm) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t i = e / cols, j = e % cols;
  F* z = Z + zoff<F>(i, j, ld, rm) * 4 * fold;
  const F* r0 = R + (i * ldr + j) * 2 * fold;
  const F* r1 = R + (i * ldr + off1 + j) * 2 * fold;
  for (int q = 0; q < fold; ++q) {
    z[2 * q] = r0[q];
    z[2 * q + 1] = r1[q];
    z[2 * fold + 2 * q] = r0[fold + q];
    z[2 * fold + 2 * q + 1] = r1[fold + q];
  }
}

// binned(beta*c) per component into R (beta mode: 0 -- zero, 1 -- c itself,
// 2 -- the complex product), reference rzgemv.c / rzgemm.c beta handling.
template <int CAP, class F>
__global__ __launch_bounds__(kT) void beta_split_kernel(int fold, int64_t rows, int64_t cols, const F* C, int64_t ld,
                                                        int rm, F b_r, F b_i, int mode, F* R, int64_t ldr,
                                                        int64_t off1) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t i = e / cols, j = e % cols;
  F v[2] = {F(0), F(0)};
  if (mode != 0) {
    const F* c = C + 2 * zoff<F>(i, j, ld, rm);
    if (mode == 1) {
      v[0] = c[0];
      v[1] = c[1];
    } else {
      v[0] = c[0] * b_r - c[1] * b_i;
      v[1] = c[0] * b_i + c[1] * b_r;
    }
  }
  for (int ch = 0; ch < 2; ++ch) {
    F p[CAP], q[CAP];
    set_zero(fold, p, q);
    if (mode != 0) {
      // binned_dmdconv: update, deposit (non-finite short cut), renorm
      update(fold, v[ch], p, q);
      if (nan_or_inf(v[ch]) || nan_or_inf(p[0]))
        p[0] += v[ch];
      else
        deposit(fold, dmindex0(p[0]), v[ch], p);
      renorm(fold, p, q);
    }
    F* r = R + (i * ldr + (ch ? off1 : 0) + j) * 2 * fold;
RB_UNROLL_CAP
    for (int t = 0; t < CAP; ++t)
      if (t < fold) {
        r[t] = p[t];
        r[fold + t] = q[t];
      }
  }
}

template <int CAP, class F>
__global__ __launch_bounds__(kT) void conv_join_kernel(int fold, int64_t rows, int64_t cols, const F* R, int64_t ldr,
                                                       int64_t off1, F* C, int64_t ld, int rm) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t i = e / cols, j = e % cols;
  F* c = C + 2 * zoff<F>(i, j, ld, rm);
  for (int ch = 0; ch < 2; ++ch) {
    const F* r = R + (i * ldr + (ch ? off1 : 0) + j) * 2 * fold;
    F p[CAP], q[CAP];
RB_UNROLL_CAP
    for (int t = 0; t < CAP; ++t)
      if (t < fold) {
        p[t] = r[t];
        q[t] = r[fold + t];
      }
    c[ch] = convert(fold, p, q);
  }
}

// alpha*x gathered contiguously (interleaved) for the gemv kernels.
template <class F>
__global__ __launch_bounds__(kT) void alpha_x_kernel(int64_t n, const F* x, int64_t incx, F al_r, F al_i, int mode,
                                                     F* xs) {
  const int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (j >= n) return;
  const F* p = x + 2 * j * incx;
  times_scalar(mode, p[0], p[1], al_r, al_i, xs[2 * j], xs[2 * j + 1]);
}

inline int trans_code(char t) { return is_n(t) ? 0 : (is_t(t) ? 1 : 2); }

template <class F>
bool real_gemm(int fold, int64_t M, int64_t N, int64_t K, const F* A, const F* B, F* Cb, hipStream_t s);
template <>
bool real_gemm<double>(int fold, int64_t M, int64_t N, int64_t K, const double* A, const double* B, double* Cb,
                       hipStream_t s) {
  return dbdgemm(fold, 'R', 'N', 'N', M, N, K, 1.0, A, K, B, N, Cb, N, s);
}
template <>
bool real_gemm<float>(int fold, int64_t M, int64_t N, int64_t K, const float* A, const float* B, float* Cb,
                      hipStream_t s) {
  return sbsgemm(fold, 'R', 'N', 'N', M, N, K, 1.0f, A, K, B, N, Cb, N, s);
}
template <class F>
bool fold_ok(int fold) {
  if (fold >= 2 && fold <= Fmt<F>::max_fold) return true;
  set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
  return false;
}

#define RB_ZFOLD_LAUNCH(KERNEL, F, GRID, ...)                                       \
  do {                                                                             \
    if (fold <= 2)                                                                 \
      KERNEL<2, F><<<GRID, kT, 0, s>>>(__VA_ARGS__);                               \
    else if (fold <= 3)                                                            \
      KERNEL<3, F><<<GRID, kT, 0, s>>>(__VA_ARGS__);                               \
    else if (fold <= 4)                                                            \
      KERNEL<4, F><<<GRID, kT, 0, s>>>(__VA_ARGS__);                               \
    else if (fold <= 8)                                                            \
      KERNEL<8, F><<<GRID, kT, 0, s>>>(__VA_ARGS__);                               \
    else                                                                           \
      KERNEL<Fmt<F>::max_fold, F><<<GRID, kT, 0, s>>>(__VA_ARGS__);                \
  } while (0)

// The common core: op(A), op(B) (or the gemv rows/vectors) into real images,
// then the real kernels accumulate into R (real and imaginary parts).
template <class F>
bool zgemm_into(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const F* alpha, const F* A,
                int64_t lda, const F* B, int64_t ldb, F* R, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0 || (alpha[0] == F(0) && alpha[1] == F(0))) return true;
  const bool rm = row_major(order);
  const int tca = trans_code(ta), tcb = trans_code(tb);
  int mode = (alpha[0] == F(1) && alpha[1] == F(0)) ? 0 : 1;
  if (mode == 1 && !rm && tca != 0) mode = 2;  // the reference's column-major transposed-A expression
  F* RA = (F*)workspace(s, kSlotZA, sizeof(F) * (size_t)(M * 2 * K));
  F* RB = (F*)workspace(s, kSlotZB, sizeof(F) * (size_t)(2 * K * 2 * N));
  if (!RA || !RB) return false;
  prep_a_kernel<F><<<grid1(M * K), kT, 0, s>>>(M, K, rm, tca, A, lda, alpha[0], alpha[1], mode, RA);
  prep_b_kernel<F><<<grid1(K * N), kT, 0, s>>>(K, N, rm, tcb, B, ldb, RB);
  if (!RB_HIP(hipGetLastError())) return false;
  return real_gemm<F>(fold, M, 2 * N, 2 * K, RA, RB, R, s);
}

// Accumulates the real and imaginary parts of output i into R[2i], R[2i+1].
template <class F>
bool zgemv_into(int fold, char order, char trans, int64_t M, int64_t N, const F* alpha, const F* A, int64_t lda,
                const F* x, int64_t incx, F* R, hipStream_t s) {
  if (M <= 0 || N <= 0 || (alpha[0] == F(0) && alpha[1] == F(0))) return true;
  const bool rm = row_major(order);
  const int tc = trans_code(trans);
  const int64_t Mo = tc == 0 ? M : N, No_all = tc == 0 ? N : M;
  const int mode = (alpha[0] == F(1) && alpha[1] == F(0)) ? 0 : 1;
  const int conj = (tc == 2) ? 1 : 0;
  // rows of op(A) are contiguous for row-major 'N' and column-major 'T'/'C':
  // two real row-dot passes over the interleaved rows; otherwise the terms of
  // an output are lda apart: one complex column pass
  const bool rows_contig = rm ? (tc == 0) : (tc != 0);
  // column-major conjugate transpose: the reference dots MIN(X_BLOCK, N - j)
  // rows of each M-row column per block j < M (zbzgemv.ccog:214-220,
  // cbcgemv.ccog the same), i.e. the first N rows when M > N; for M < N it
  // reads past the column (no defined answer) and the M rows are used
  const int64_t No = (!rm && tc == 2 && M > N) ? N : No_all;
  F* xs = (F*)workspace(s, kSlotZB, sizeof(F) * (size_t)(2 * No));
  if (!xs) return false;
  alpha_x_kernel<F><<<grid1(No), kT, 0, s>>>(No, x, incx, alpha[0], alpha[1], mode, xs);
  if (!RB_HIP(hipGetLastError())) return false;
  if (rows_contig) return zgemv_rows(fold, Mo, No, A, lda, conj, xs, R, s);
  return zgemv_cols(fold, Mo, No, A, lda, conj, xs, R, s);
}

template <class F>
bool zbzgemm_impl(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const F* alpha,
                  const F* A, int64_t lda, const F* B, int64_t ldb, F* Zb, int64_t ldz, hipStream_t s) {
  if (!fold_ok<F>(fold)) return false;
  if (M <= 0 || N <= 0 || K <= 0 || (alpha[0] == F(0) && alpha[1] == F(0))) return true;
  const int rm = row_major(order);
  F* R = (F*)workspace(s, kSlotZC, sizeof(F) * (size_t)(M * 2 * N * 2 * fold));
  if (!R) return false;
  split_kernel<F><<<grid1(M * N), kT, 0, s>>>(fold, M, N, Zb, ldz, rm, R, 2 * N, N);
  if (!zgemm_into<F>(fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, R, s)) return false;
  join_kernel<F><<<grid1(M * N), kT, 0, s>>>(fold, M, N, R, 2 * N, N, Zb, ldz, rm);
  return RB_HIP(hipGetLastError());
}

template <class F>
bool rzgemm_impl(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const F* alpha, const F* A,
                 int64_t lda, const F* B, int64_t ldb, const F* beta, F* C, int64_t ldc, hipStream_t s) {
  if (!fold_ok<F>(fold)) return false;
  if (M <= 0 || N <= 0) return true;
  const int rm = row_major(order);
  F* R = (F*)workspace(s, kSlotZC, sizeof(F) * (size_t)(M * 2 * N * 2 * fold));
  if (!R) return false;
  const int bmode = (beta[0] == F(0) && beta[1] == F(0)) ? 0 : ((beta[0] == F(1) && beta[1] == F(0)) ? 1 : 2);
  RB_ZFOLD_LAUNCH(beta_split_kernel, F, grid1(M * N), fold, M, N, C, ldc, rm, beta[0], beta[1], bmode, R, 2 * N, N);
  if (!zgemm_into<F>(fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, R, s)) return false;
  RB_ZFOLD_LAUNCH(conv_join_kernel, F, grid1(M * N), fold, M, N, R, 2 * N, N, C, ldc, rm);
  return RB_HIP(hipGetLastError());
}

template <class F>
bool zbzgemv_impl(int fold, char order, char trans, int64_t M, int64_t N, const F* alpha, const F* A, int64_t lda,
                  const F* x, int64_t incx, F* Yb, int64_t incy, hipStream_t s) {
  if (!fold_ok<F>(fold)) return false;
  if (M <= 0 || N <= 0 || (alpha[0] == F(0) && alpha[1] == F(0))) return true;
  const int64_t Mo = is_n(trans) ? M : N;
  F* R = (F*)workspace(s, kSlotZC, sizeof(F) * (size_t)(2 * Mo * 2 * fold));
  if (!R) return false;
  // Y as an Mo x 1 column with leading dimension incy (row-major view);
  // parts interleaved in R
  split_kernel<F><<<grid1(Mo), kT, 0, s>>>(fold, Mo, 1, Yb, incy, 1, R, 2, 1);
  if (!zgemv_into<F>(fold, order, trans, M, N, alpha, A, lda, x, incx, R, s)) return false;
  join_kernel<F><<<grid1(Mo), kT, 0, s>>>(fold, Mo, 1, R, 2, 1, Yb, incy, 1);
  return RB_HIP(hipGetLastError());
}

template <class F>
bool rzgemv_impl(int fold, char order, char trans, int64_t M, int64_t N, const F* alpha, const F* A, int64_t lda,
                 const F* x, int64_t incx, const F* beta, F* y, int64_t incy, hipStream_t s) {
  if (!fold_ok<F>(fold)) return false;
  if (M <= 0 || N <= 0) return true;
  const int64_t Mo = is_n(trans) ? M : N;
  F* R = (F*)workspace(s, kSlotZC, sizeof(F) * (size_t)(2 * Mo * 2 * fold));
  if (!R) return false;
  const int bmode = (beta[0] == F(0) && beta[1] == F(0)) ? 0 : ((beta[0] == F(1) && beta[1] == F(0)) ? 1 : 2);
  RB_ZFOLD_LAUNCH(beta_split_kernel, F, grid1(Mo), fold, Mo, 1, y, incy, 1, beta[0], beta[1], bmode, R, 2, 1);
  if (!zgemv_into<F>(fold, order, trans, M, N, alpha, A, lda, x, incx, R, s)) return false;
  RB_ZFOLD_LAUNCH(conv_join_kernel, F, grid1(Mo), fold, Mo, 1, R, 2, 1, y, incy, 1);
  return RB_HIP(hipGetLastError());
}

}  // namespace

bool zbzgemm(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const double* alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double* Zb, int64_t ldz, hipStream_t s) {
  return zbzgemm_impl<double>(fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, Zb, ldz, s);
}
bool zbzgemm(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const float* alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float* Zb, int64_t ldz, hipStream_t s) {
  return zbzgemm_impl<float>(fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, Zb, ldz, s);
}
bool rzgemm_device(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const double* alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, const double* beta, double* C,
                   int64_t ldc, hipStream_t s) {
  return rzgemm_impl<double>(fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
bool rzgemm_device(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const float* alpha,
                   const float* A, int64_t lda, const float* B, int64_t ldb, const float* beta, float* C, int64_t ldc,
                   hipStream_t s) {
  return rzgemm_impl<float>(fold, order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
bool zbzgemv(int fold, char order, char trans, int64_t M, int64_t N, const double* alpha, const double* A,
             int64_t lda, const double* x, int64_t incx, double* Yb, int64_t incy, hipStream_t s) {
  return zbzgemv_impl<double>(fold, order, trans, M, N, alpha, A, lda, x, incx, Yb, incy, s);
}
bool zbzgemv(int fold, char order, char trans, int64_t M, int64_t N, const float* alpha, const float* A, int64_t lda,
             const float* x, int64_t incx, float* Yb, int64_t incy, hipStream_t s) {
  return zbzgemv_impl<float>(fold, order, trans, M, N, alpha, A, lda, x, incx, Yb, incy, s);
}
bool rzgemv_device(int fold, char order, char trans, int64_t M, int64_t N, const double* alpha, const double* A,
                   int64_t lda, const double* x, int64_t incx, const double* beta, double* y, int64_t incy,
                   hipStream_t s) {
  return rzgemv_impl<double>(fold, order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}
bool rzgemv_device(int fold, char order, char trans, int64_t M, int64_t N, const float* alpha, const float* A,
                   int64_t lda, const float* x, int64_t incx, const float* beta, float* y, int64_t incy,
                   hipStream_t s) {
  return rzgemv_impl<float>(fold, order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}

}  // namespace rb
