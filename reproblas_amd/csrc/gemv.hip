// gemv.hip -- binned gemv on device operands.
//
// Reference: binnedBLAS_dbdgemv (src/binnedBLAS/dbdgemv.ccog:57-162) under
// reproBLAS_rdgemv (src/reproBLAS/rdgemv.c:35-87):
//   y_i = ddbconv( dbdconv(beta*y_i) (+) sum_j fl(op(A)_ij * fl(alpha*x_j)) )
// with alpha applied to x (dbdgemv.ccog:86-95, 110-119), beta == 0 ignoring y
// and beta == 1 converting y unscaled (rdgemv.c:50-61). Each output is one
// binned dot, so any split of its terms gives the same bits.
//
// Two memory patterns cover the four Order x Trans cases:
//   row-dot  -- the terms of an output are contiguous (row-major N,
//               column-major T): fl(alpha*x) is first gathered into a
//               contiguous vector; then a persistent grid of waves, one
//               output per wave at a time, streams the rows as BLAS1 dots
//               (wave_sum.hpp) and one lane folds in beta*y and converts.
//   col-axpy -- the terms of an output are lda apart (column-major N,
//               row-major T): each lane owns two adjacent outputs and walks
//               the columns, so a wave reads 1 KiB contiguous per column with
//               16-byte loads; every output keeps its own index. Column ranges
//               are split across workgroups; a second kernel merges the
//               per-range binned partials, folds in beta*y and converts.
#include <algorithm>

#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"
#include "wave_sum.hpp"

namespace rb {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

struct GemvArgs {
  int64_t nout, len;
  const double* A;
  int64_t lda;
  const double* x;
  int64_t incx;
  double alpha;
  double* Yb;  // binned outputs (accumulated into) or null
  int64_t incYb;
  double* y;  // double outputs when Yb is null: y = conv(binned(beta*y) (+) sum)
  int64_t incy;
  double beta;
  double* partials;  // col-axpy: [chunks][nout][2*fold]
  int64_t chunk;     // col-axpy: columns per chunk
  int64_t chunks;
  int fold;
};

__device__ __forceinline__ double scaled_x(const GemvArgs& a, int64_t k) {
  const double v = a.x[k * a.incx];
  return a.alpha == 1.0 ? v : v * a.alpha;
}

// binned(beta*y) as rdgemv.c:50-61 forms it (binned_dbdconv = dmdadd on zero)
template <int CAP>
__device__ __forceinline__ void beta_init(int fold, double beta, double yv, double (&p)[CAP], double (&c)[CAP]) {
  set_zero(fold, p, c);
  if (beta == 0.0) return;
  const double v = (beta == 1.0) ? yv : yv * beta;
  update(fold, v, p, c);
  if (nan_or_inf(v) || nan_or_inf(p[0])) {
    p[0] += v;
  } else {
    deposit(fold, dmindex0(p[0]), v, p);
  }
  renorm(fold, p, c);
}

// Final step for one output: fold the sum S into the caller's value.
template <int CAP>
__device__ __forceinline__ void finish_output(const GemvArgs& a, int fold, int64_t o, double (&p)[CAP],
                                              double (&c)[CAP]) {
  if (a.Yb) {
    double* yb = a.Yb + o * a.incYb * 2 * fold;
    double py[CAP], cy[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { py[j] = yb[j]; cy[j] = yb[fold + j]; }
    renorm(fold, py, cy);
    add(fold, py, cy, p, c);
    renorm(fold, py, cy);
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { yb[j] = py[j]; yb[fold + j] = cy[j]; }
  } else {
    double py[CAP], cy[CAP];
    beta_init(fold, a.beta, a.beta == 0.0 ? 0.0 : a.y[o * a.incy], py, cy);
    add(fold, py, cy, p, c);
    renorm(fold, py, cy);
    a.y[o * a.incy] = to_double(fold, py, cy);
  }
}

// ---------------------------------------------------------------- row-dot --
// xs[k] = fl(alpha * x[k*incx]), the reference's scaled x (dbdgemv.ccog:86-95)
__global__ __launch_bounds__(kThreads) void scale_x_kernel(GemvArgs a, double* xs) {
  const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (k < a.len) xs[k] = scaled_x(a, k);
}

// Terms per lane per tile of the row-dot (the BLAS1 dot's shape: 8 KiB per
// wave tile).
constexpr int kURow = 8;

__device__ __forceinline__ Span row_span(const GemvArgs& a, const double* xs, int64_t o) {
  return make_span<kURow>(a.A + o * a.lda, 1, xs, 1, a.len);
}

// The wave's value of row o is complete: merge the lanes, convert, store.
template <int CAP, int MODE>
__device__ __forceinline__ void finish_row(const GemvArgs& a, const double* xs, int fold, int64_t o, int lane,
                                           const WaveRun& w, State<CAP>& s) {
  wave_finish<CAP, kDot, kURow, MODE>(row_span(a, xs, o), fold, 0, 1, lane, w, s);
  if (lane == 0) finish_output(a, fold, o, s.p, s.c);
}

// Wave g of W owns rows g, g + W, ... and walks their tiles as one sequence,
// so the next tile (the first of its next row, at a row's end) is in flight
// while the current one is deposited; two register buffers take turns.
template <int CAP, bool FIXED, int MODE>
__global__ __launch_bounds__(kThreads) void rowdot_kernel(GemvArgs a, const double* xs) {
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * kWaves;
  int64_t o = (int64_t)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (o >= a.nout) return;  // whole wave leaves together
  const int64_t tiles = (a.len + 64 * kURow - 1) / (64 * kURow);
  State<CAP> s;
  WaveRun w;
  wave_start(fold, w, s);
  double ax[kURow], ay[kURow], bx[kURow], by[kURow];
  int64_t t = 0;
  load_tile<kDot, kURow, MODE>(row_span(a, xs, o), t, lane, ax, ay);
  while (true) {
    int64_t on = o, tn = t + 1;
    if (tn == tiles) { on = o + nw; tn = 0; }
    if (on < a.nout) load_tile<kDot, kURow, MODE>(row_span(a, xs, on), tn, lane, bx, by);
    deposit_tile<CAP, kDot, kURow>(fold, ax, ay, w, s);
    if (tn == 0) {
      finish_row<CAP, MODE>(a, xs, fold, o, lane, w, s);
      wave_start(fold, w, s);
    }
    if (on >= a.nout) break;
    o = on;
    t = tn;
    on = o;
    tn = t + 1;
    if (tn == tiles) { on = o + nw; tn = 0; }
    if (on < a.nout) load_tile<kDot, kURow, MODE>(row_span(a, xs, on), tn, lane, ax, ay);
    deposit_tile<CAP, kDot, kURow>(fold, bx, by, w, s);
    if (tn == 0) {
      finish_row<CAP, MODE>(a, xs, fold, o, lane, w, s);
      wave_start(fold, w, s);
    }
    if (on >= a.nout) break;
    o = on;
    t = tn;
  }
}

// --------------------------------------------------------------- col-axpy --
// Block: 128 outputs (2 per lane) x one column chunk; wave w takes columns
// chunk_start + w, +4, +8, ...
template <int CAP, bool FIXED, bool VEC>
__global__ __launch_bounds__(kThreads) void colaxpy_kernel(GemvArgs a) {
  constexpr int U = 4;  // columns per step
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (scalar)
  const int64_t o0 = (int64_t)blockIdx.x * 128 + 2 * lane;
  const int64_t kbeg = (int64_t)blockIdx.y * a.chunk;
  int64_t kend = kbeg + a.chunk;
  if (kend > a.len) kend = a.len;
  const bool in0 = o0 < a.nout, in1 = o0 + 1 < a.nout;

  double p0[CAP], c0[CAP], p1[CAP], c1[CAP];
  int idx0 = kMaxIndex, idx1 = kMaxIndex;
  set_bins(fold, kMaxIndex, p0, c0);
  set_bins(fold, kMaxIndex, p1, c1);
  double lim0 = admit_limit(kMaxIndex), lim1 = lim0;
  int dep = 0;
  for (int64_t k = kbeg + wave; k < kend; k += kWaves * U) {
    double t0[U], t1[U];
    double m0 = 0.0, m1 = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t kk = k + (int64_t)u * kWaves;
      const bool kin = kk < kend;
      double a0 = 0.0, a1 = 0.0, xs = 0.0;
      if (kin) {
        xs = scaled_x(a, kk);
        const double* col = a.A + kk * a.lda + o0;
        if (VEC && in1) {
          const double2 d = *reinterpret_cast<const double2*>(col);
          a0 = d.x;
          a1 = d.y;
        } else {
          a0 = in0 ? col[0] : 0.0;
          a1 = in1 ? col[1] : 0.0;
        }
      }
      t0[u] = a0 * xs;
      t1[u] = a1 * xs;
      m0 = fmax(m0, fabs(t0[u]));
      m1 = fmax(m1, fabs(t1[u]));
    }
    if (dep + U > kEndurance) {
      renorm(fold, p0, c0);
      renorm(fold, p1, c1);
      dep = 0;
    }
    if (__any(m0 >= lim0 || m1 >= lim1)) {
      // every output has its own index: per-lane raise
      const int n0 = dindex(m0), n1 = dindex(m1);
      if (n0 < idx0) {
        raise_index(fold, idx0 - n0, n0, p0, c0);
        idx0 = n0;
        lim0 = admit_limit(idx0);
      }
      if (n1 < idx1) {
        raise_index(fold, idx1 - n1, n1, p1, c1);
        idx1 = n1;
        lim1 = admit_limit(idx1);
      }
    }
    const bool top0 = idx0 == 0, top1 = idx1 == 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      deposit(fold, top0, t0[u], p0);
      deposit(fold, top1, t1[u], p1);
    }
    dep += U;
  }
  renorm(fold, p0, c0);
  renorm(fold, p1, c1);

  // merge the four waves' values per output: through LDS for the small
  // capacities; the general capacity writes one partial per wave instead
  double* part;
  if constexpr (CAP <= 8) {
    __shared__ double lds[kWaves - 1][128][2 * CAP];
    if (wave > 0) {
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) {
          lds[wave - 1][2 * lane][j] = p0[j];
          lds[wave - 1][2 * lane][CAP + j] = c0[j];
          lds[wave - 1][2 * lane + 1][j] = p1[j];
          lds[wave - 1][2 * lane + 1][CAP + j] = c1[j];
        }
    }
    __syncthreads();
    if (wave != 0) return;
    for (int w = 0; w < kWaves - 1; ++w) {
      double qp[CAP], qc[CAP];
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) { qp[j] = lds[w][2 * lane][j]; qc[j] = lds[w][2 * lane][CAP + j]; }
      add(fold, p0, c0, qp, qc);
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) { qp[j] = lds[w][2 * lane + 1][j]; qc[j] = lds[w][2 * lane + 1][CAP + j]; }
      add(fold, p1, c1, qp, qc);
    }
    part = a.partials + ((int64_t)blockIdx.y * a.nout) * 2 * fold;
  } else {
    part = a.partials + (((int64_t)blockIdx.y * kWaves + wave) * a.nout) * 2 * fold;
  }
  if (in0) {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { part[o0 * 2 * fold + j] = p0[j]; part[o0 * 2 * fold + fold + j] = c0[j]; }
  }
  if (in1) {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { part[(o0 + 1) * 2 * fold + j] = p1[j]; part[(o0 + 1) * 2 * fold + fold + j] = c1[j]; }
  }
}

// merge the per-chunk partials of each output, fold in beta*y (or Yb), convert
template <int CAP, bool FIXED>
__global__ __launch_bounds__(kThreads) void colaxpy_finish(GemvArgs a) {
  const int fold = FIXED ? CAP : a.fold;
  const int64_t o = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (o >= a.nout) return;
  double p[CAP], c[CAP];
  const double* src = a.partials + o * 2 * fold;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = src[j]; c[j] = src[fold + j]; }
  const int64_t records = (CAP <= 8) ? a.chunks : a.chunks * kWaves;
  for (int64_t s = 1; s < records; ++s) {
    const double* q = a.partials + (s * a.nout + o) * 2 * fold;
    double qp[CAP], qc[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { qp[j] = q[j]; qc[j] = q[fold + j]; }
    add(fold, p, c, qp, qc);
  }
  if (nan_or_inf(p[0])) {
    bool nan = false, pinf = false, ninf = false;
    for (int64_t k = 0; k < a.len; ++k) {
      const double t = a.A[o + k * a.lda] * scaled_x(a, k);
      nan |= (t != t);
      pinf |= (t == __builtin_inf());
      ninf |= (t == -__builtin_inf());
    }
    p[0] = (nan || (pinf && ninf) || !(pinf || ninf)) ? __builtin_nan("") : (pinf ? __builtin_inf() : -__builtin_inf());
  }
  finish_output(a, fold, o, p, c);
}

// Persistent row-dot workgroups per CU (compile-time overridable for tuning
// builds).
#ifndef RB_GEMV_ROW_BLOCKS_PER_CU
#define RB_GEMV_ROW_BLOCKS_PER_CU 4
#endif
constexpr int kRowBlocksPerCU = RB_GEMV_ROW_BLOCKS_PER_CU;

template <int CAP, bool FIXED>
bool launch_gemv(bool rowdot, GemvArgs& a, hipStream_t s) {
  const DeviceInfo* d = device_info();
  if (!d) return false;
  if (rowdot) {
    double* xs = (double*)workspace(s, kSlotGemvX, sizeof(double) * (size_t)a.len);
    if (!xs) return false;
    scale_x_kernel<<<(unsigned)((a.len + kThreads - 1) / kThreads), kThreads, 0, s>>>(a, xs);
    const int64_t blocks = std::min<int64_t>((a.nout + kWaves - 1) / kWaves, (int64_t)d->cus * kRowBlocksPerCU);
    const bool vec = ((uintptr_t)a.A % 16 == 0) && (a.lda % 2 == 0);
    if (vec)
      rowdot_kernel<CAP, FIXED, 0><<<(unsigned)blocks, kThreads, 0, s>>>(a, xs);
    else
      rowdot_kernel<CAP, FIXED, 1><<<(unsigned)blocks, kThreads, 0, s>>>(a, xs);
    return RB_HIP(hipGetLastError());
  }
  const int64_t rowblocks = (a.nout + 127) / 128;
  int64_t chunks = (4 * (int64_t)d->cus + rowblocks - 1) / rowblocks;
  const int64_t max_chunks = (a.len + 255) / 256;  // keep >= 256 columns per chunk
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  if (chunks > 65535) chunks = 65535;
  a.chunk = (a.len + chunks - 1) / chunks;
  a.chunks = (a.len + a.chunk - 1) / a.chunk;
  const int64_t records = (CAP <= 8) ? a.chunks : a.chunks * kWaves;
  a.partials = (double*)workspace(s, kSlotPartials, sizeof(double) * (size_t)records * a.nout * 2 * a.fold);
  if (!a.partials) return false;
  const bool vec = ((uintptr_t)a.A % 16 == 0) && (a.lda % 2 == 0);
  dim3 grid((unsigned)rowblocks, (unsigned)a.chunks);
  if (vec)
    colaxpy_kernel<CAP, FIXED, true><<<grid, kThreads, 0, s>>>(a);
  else
    colaxpy_kernel<CAP, FIXED, false><<<grid, kThreads, 0, s>>>(a);
  if (!RB_HIP(hipGetLastError())) return false;
  colaxpy_finish<CAP, FIXED><<<(unsigned)((a.nout + kThreads - 1) / kThreads), kThreads, 0, s>>>(a);
  return RB_HIP(hipGetLastError());
}

bool gemv_common(int fold, char order, char trans, int64_t M, int64_t N, double alpha, const double* A, int64_t lda,
                 const double* x, int64_t incx, double* Yb, int64_t incYb, double* y, int64_t incy, double beta,
                 hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  const bool rowmajor = (order == 'r' || order == 'R');
  const bool notrans = (trans == 'n' || trans == 'N');
  GemvArgs a;
  a.nout = notrans ? M : N;
  a.len = notrans ? N : M;
  a.A = A;
  a.lda = lda;
  a.x = x;
  a.incx = incx;
  a.alpha = alpha;
  a.Yb = Yb;
  a.incYb = incYb;
  a.y = y;
  a.incy = incy;
  a.beta = beta;
  a.partials = nullptr;
  a.chunk = a.chunks = 0;
  a.fold = fold;
  if (a.nout <= 0) return true;
  // the terms of an output are contiguous for row-major N and column-major T
  const bool rowdot = (rowmajor == notrans);
  switch (fold) {
    case 2: return launch_gemv<2, true>(rowdot, a, s);
    case 3: return launch_gemv<3, true>(rowdot, a, s);
    case 4: return launch_gemv<4, true>(rowdot, a, s);
    default:
      if (fold <= 8) return launch_gemv<8, false>(rowdot, a, s);
      return launch_gemv<kMaxFold, false>(rowdot, a, s);
  }
}

}  // namespace

bool dbdgemv(int fold, char order, char trans, int64_t M, int64_t N, double alpha, const double* A, int64_t lda,
             const double* x, int64_t incx, double* Yb, int64_t incYb, hipStream_t s) {
  if (M <= 0 || N <= 0 || alpha == 0.0) return true;  // dbdgemv.ccog:75-77
  return gemv_common(fold, order, trans, M, N, alpha, A, lda, x, incx, Yb, incYb, nullptr, 0, 0.0, s);
}

// y = conv(binned(beta*y) (+) sum) in place, for reproBLAS_rdgemv
bool rdgemv_device(int fold, char order, char trans, int64_t M, int64_t N, double alpha, const double* A,
                   int64_t lda, const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s) {
  if (M <= 0 || N <= 0) return true;  // rdgemv.c:43-45
  const bool notrans = (trans == 'n' || trans == 'N');
  const int64_t nout = notrans ? M : N;
  if (alpha == 0.0) {
    // dbdgemv adds nothing; y still goes through binned(beta*y) and back
    double* tmp = (double*)workspace(s, kSlotBinnedOut, sizeof(double) * (size_t)nout * 2 * fold);
    if (!tmp) return false;
    return conv_to_binned(fold, nout, y, incy, beta, tmp, s) && conv_to_double(fold, nout, tmp, y, incy, s);
  }
  return gemv_common(fold, order, trans, M, N, alpha, A, lda, x, incx, nullptr, 0, y, incy, beta, s);
}

}  // namespace rb
