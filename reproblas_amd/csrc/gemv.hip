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
//               output (double: a pair of outputs sharing the x tiles) per
//               wave at a time, streams the rows as BLAS1 dots
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

template <class F>
struct GemvArgs {
  int64_t nout, len;
  const F* A;
  int64_t lda;
  const F* x;
  int64_t incx;
  F alpha;
  F* Yb;  // binned outputs (accumulated into) or null
  int64_t incYb;
  F* y;  // plain outputs when Yb is null: y = conv(binned(beta*y) (+) sum)
  int64_t incy;
  F beta;
  F* partials;  // col-axpy: [chunks][nout][2*fold]
  int64_t chunk;     // col-axpy: columns per chunk
  int64_t chunks;
  int fold;
  int conj;  // complex col-axpy: conjugate A's elements
};

template <class F>
__device__ __forceinline__ F scaled_x(const GemvArgs<F>& a, int64_t k) {
  const F v = a.x[k * a.incx];
  return a.alpha == F(1) ? v : v * a.alpha;
}

// binned(beta*y) as rdgemv.c:50-61 forms it (binned_dbdconv = dmdadd on zero)
template <int CAP, class F>
__device__ __forceinline__ void beta_init(int fold, F beta, F yv, F (&p)[CAP], F (&c)[CAP]) {
  set_zero(fold, p, c);
  if (beta == F(0)) return;
  const F v = (beta == F(1)) ? yv : yv * beta;
  update(fold, v, p, c);
  if (nan_or_inf(v) || nan_or_inf(p[0])) {
    p[0] += v;
  } else {
    deposit(fold, dmindex0(p[0]), v, p);
  }
  renorm(fold, p, c);
}

// Final step for one output: fold the sum S into the caller's value.
template <int CAP, class F>
__device__ __forceinline__ void finish_output(const GemvArgs<F>& a, int fold, int64_t o, F (&p)[CAP], F (&c)[CAP]) {
  if (a.Yb) {
    F* yb = a.Yb + o * a.incYb * 2 * fold;
    F py[CAP], cy[CAP];
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
    F py[CAP], cy[CAP];
    beta_init(fold, a.beta, a.beta == F(0) ? F(0) : a.y[o * a.incy], py, cy);
    add(fold, py, cy, p, c);
    renorm(fold, py, cy);
    a.y[o * a.incy] = convert(fold, py, cy);
  }
}

// ---------------------------------------------------------------- row-dot --
// xs[k] = fl(alpha * x[k*incx]), the reference's scaled x (dbdgemv.ccog:86-95)
template <class F>
__global__ __launch_bounds__(kThreads) void scale_x_kernel(GemvArgs<F> a, F* xs) {
  const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (k < a.len) xs[k] = scaled_x(a, k);
}

// Terms per lane per tile of the row-dot (compile-time overridable for
// tuning builds; default the BLAS1 dot's shape: 8 KiB per wave tile). On
// MI355X, dgemv / sgemv 16384^2 column-major T at fold 3: 8 doubles per lane
// 5.8 / 5.9 TB/s, 4 doubles 5.5 / 4.9, 2 doubles 5.6 / 5.5.
#ifndef RB_GEMV_ROW_U
#define RB_GEMV_ROW_U 8
#endif
template <class F>
constexpr int u_row() { return RB_GEMV_ROW_U * (int)(sizeof(double) / sizeof(F)); }

template <class F>
__device__ __forceinline__ Span<F> row_span(const GemvArgs<F>& a, const F* xs, int64_t o) {
  return make_span<u_row<F>()>(a.A + o * a.lda, 1, xs, 1, a.len);
}

// The wave's value of row o is complete: merge the lanes, convert, store.
template <int CAP, int MODE, class F>
__device__ __forceinline__ void finish_row(const GemvArgs<F>& a, const F* xs, int fold, int64_t o, int lane,
                                           const WaveRun<F>& w, State<CAP, F>& s) {
  wave_finish<CAP, kDot, u_row<F>(), MODE>(row_span(a, xs, o), fold, 0, 1, lane, w, s, nullptr);
  if (lane == 0) finish_output(a, fold, o, s.p, s.c);
}

// Wave g of W owns rows g, g + W, ... and walks their tiles as one sequence,
// so the next tile (the first of its next row, at a row's end) is in flight
// while the current one is deposited; two register buffers take turns.
template <int CAP, bool FIXED, int MODE, class F>
__global__ __launch_bounds__(kThreads) void rowdot_kernel(GemvArgs<F> a, const F* xs) {
  constexpr int kURow = u_row<F>();
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * kWaves;
  int64_t o = (int64_t)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (o >= a.nout) return;  // whole wave leaves together
  const int64_t tiles = (a.len + 64 * kURow - 1) / (64 * kURow);
  State<CAP, F> s;
  WaveRun<F> w;
  wave_start(fold, w, s);
  F ax[kURow], ay[kURow], bx[kURow], by[kURow];
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

// Rows 2p and 2p + 1 of tile t: A tiles of both, the xs tile once.
template <int MODE, class F>
__device__ __forceinline__ void load_row_pair(const GemvArgs<F>& a, const F* xs, int64_t p, int64_t t, int lane,
                                              F (&v0)[u_row<F>()], F (&v1)[u_row<F>()], F (&vx)[u_row<F>()]) {
  constexpr int kURow = u_row<F>();
  load_tile<kDot, kURow, MODE>(row_span(a, xs, 2 * p), t, lane, v0, vx);
  if (2 * p + 1 < a.nout) {
    F unused[kURow];
    load_tile<kSum, kURow, MODE>(row_span(a, xs, 2 * p + 1), t, lane, v1, unused);
  } else {
#pragma unroll
    for (int k = 0; k < kURow; ++k) v1[k] = F(0);
  }
}

// Row pairs: wave g of W owns pairs g, g + W, ... (rows 2p and 2p + 1) and
// deposits both rows of a tile against one xs tile, so each xs element is
// read once per two rows. Each row keeps its own bins and wave index.
template <int CAP, bool FIXED, int MODE, class F>
__global__ __launch_bounds__(kThreads) void rowdot2_kernel(GemvArgs<F> a, const F* xs) {
  constexpr int kURow = u_row<F>();
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * kWaves;
  const int64_t pairs = (a.nout + 1) / 2;
  int64_t p = (int64_t)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (p >= pairs) return;  // whole wave leaves together
  const int64_t tiles = (a.len + 64 * kURow - 1) / (64 * kURow);
  State<CAP, F> s0, s1;
  WaveRun<F> w0, w1;
  wave_start(fold, w0, s0);
  wave_start(fold, w1, s1);
  F a0[kURow], a1[kURow], ax[kURow], b0[kURow], b1[kURow], bx[kURow];
  int64_t t = 0;
  load_row_pair<MODE>(a, xs, p, t, lane, a0, a1, ax);
  while (true) {
    int64_t pn = p, tn = t + 1;
    if (tn == tiles) { pn = p + nw; tn = 0; }
    if (pn < pairs) load_row_pair<MODE>(a, xs, pn, tn, lane, b0, b1, bx);
    deposit_tile<CAP, kDot, kURow>(fold, a0, ax, w0, s0);
    deposit_tile<CAP, kDot, kURow>(fold, a1, ax, w1, s1);
    if (tn == 0) {
      finish_row<CAP, MODE>(a, xs, fold, 2 * p, lane, w0, s0);
      if (2 * p + 1 < a.nout) finish_row<CAP, MODE>(a, xs, fold, 2 * p + 1, lane, w1, s1);
      wave_start(fold, w0, s0);
      wave_start(fold, w1, s1);
    }
    if (pn >= pairs) break;
    p = pn;
    t = tn;
    pn = p;
    tn = t + 1;
    if (tn == tiles) { pn = p + nw; tn = 0; }
    if (pn < pairs) load_row_pair<MODE>(a, xs, pn, tn, lane, a0, a1, ax);
    deposit_tile<CAP, kDot, kURow>(fold, b0, bx, w0, s0);
    deposit_tile<CAP, kDot, kURow>(fold, b1, bx, w1, s1);
    if (tn == 0) {
      finish_row<CAP, MODE>(a, xs, fold, 2 * p, lane, w0, s0);
      if (2 * p + 1 < a.nout) finish_row<CAP, MODE>(a, xs, fold, 2 * p + 1, lane, w1, s1);
      wave_start(fold, w0, s0);
      wave_start(fold, w1, s1);
    }
    if (pn >= pairs) break;
    p = pn;
    t = tn;
  }
}

// --------------------------------------------------------------- col-axpy --
// Block: 128 outputs (2 per lane) x one column chunk; wave w takes columns
// chunk_start + w, +4, +8, ...
//
// CPLX: the two outputs of a lane are the real and imaginary parts of one
// complex output i (o0 = 2i): A holds interleaved complex elements (lda in
// complex elements), x the contiguous interleaved alpha*x, and each column
// k gives every part two terms, a_r*x_r and -(a_i*x_i) (conj: a_i*x_i) for
// the real part, a_i*x_r (conj: -(a_i*x_r)) and a_r*x_i for the imaginary
// one -- the terms of the reference's zbzdotu / zbzdotc (zmzdotu.ccog,
// zmzdotc.ccog).
template <class F>
__device__ __forceinline__ void complex_terms(F ar, F ai, F xr, F xi, int conj, F& r0, F& r1, F& i0, F& i1) {
  r0 = ar * xr;
  r1 = conj ? ai * xi : (ai * xi) * F(-1);
  i0 = conj ? (ai * xr) * F(-1) : ai * xr;
  i1 = ar * xi;
}

// Columns per step of the col-axpy (real / complex) and its workgroups per
// CU (compile-time overridable for tuning builds). On MI355X, column-major N
// at fold 3, zgemv 8192^2 / dgemv 16384^2: the defaults 4.95 / 6.0 TB/s;
// complex 2 columns 5.0, 8 columns 4.6; real 8 columns 5.7; 2 workgroups
// per CU 4.2 / 5.15, 8 per CU 4.6 / 5.3.
#ifndef RB_GEMV_COL_U
#define RB_GEMV_COL_U 4
#endif
#ifndef RB_GEMV_COL_U_CPLX
#define RB_GEMV_COL_U_CPLX 4
#endif
#ifndef RB_GEMV_COL_BLOCKS_PER_CU
#define RB_GEMV_COL_BLOCKS_PER_CU 4
#endif

template <int CAP, bool FIXED, bool VEC, class F, bool CPLX = false>
__global__ __launch_bounds__(kThreads) void colaxpy_kernel(GemvArgs<F> a) {
  constexpr int U = CPLX ? RB_GEMV_COL_U_CPLX : RB_GEMV_COL_U;  // columns per step
  constexpr int TU = CPLX ? 2 * U : U;  // terms per output per step
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (scalar)
  const int64_t o0 = (int64_t)blockIdx.x * 128 + 2 * lane;
  const int64_t kbeg = (int64_t)blockIdx.y * a.chunk;
  int64_t kend = kbeg + a.chunk;
  if (kend > a.len) kend = a.len;
  const bool in0 = o0 < a.nout, in1 = o0 + 1 < a.nout;

  F p0[CAP], c0[CAP], p1[CAP], c1[CAP];
  int idx0 = Fmt<F>::max_index, idx1 = Fmt<F>::max_index;
  set_bins(fold, Fmt<F>::max_index, p0, c0);
  set_bins(fold, Fmt<F>::max_index, p1, c1);
  F lim0 = admit_limit<F>(Fmt<F>::max_index), lim1 = lim0;
  int dep = 0;
  for (int64_t k = kbeg + wave; k < kend; k += kWaves * U) {
    F t0[TU], t1[TU];
    F m0 = F(0), m1 = F(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t kk = k + (int64_t)u * kWaves;
      const bool kin = kk < kend;
      F a0 = F(0), a1 = F(0), xs = F(0);
      if constexpr (CPLX) {
        F xr = F(0), xi = F(0);
        if (kin && in0) {
          const F* el = a.A + 2 * (kk * a.lda + (o0 >> 1));
          if constexpr (!VEC) {
            a0 = el[0];
            a1 = el[1];
          } else if constexpr (sizeof(F) == 8) {
            const double2 d = *reinterpret_cast<const double2*>(el);
            a0 = d.x;
            a1 = d.y;
          } else {
            const float2 d = *reinterpret_cast<const float2*>(el);
            a0 = d.x;
            a1 = d.y;
          }
        }
        if (kin) {
          xr = a.x[2 * kk];
          xi = a.x[2 * kk + 1];
        }
        complex_terms(a0, a1, xr, xi, a.conj, t0[u], t0[U + u], t1[u], t1[U + u]);
        m0 = fmax(m0, fmax(fabs(t0[u]), fabs(t0[U + u])));
        m1 = fmax(m1, fmax(fabs(t1[u]), fabs(t1[U + u])));
        continue;
      }
      if (kin) {
        xs = scaled_x(a, kk);
        const F* col = a.A + kk * a.lda + o0;
        if (VEC && in1) {
          if constexpr (sizeof(F) == 8) {
            const double2 d = *reinterpret_cast<const double2*>(col);
            a0 = d.x;
            a1 = d.y;
          } else {
            const float2 d = *reinterpret_cast<const float2*>(col);
            a0 = d.x;
            a1 = d.y;
          }
        } else {
          a0 = in0 ? col[0] : F(0);
          a1 = in1 ? col[1] : F(0);
        }
      }
      t0[u] = a0 * xs;
      t1[u] = a1 * xs;
      m0 = fmax(m0, fabs(t0[u]));
      m1 = fmax(m1, fabs(t1[u]));
    }
    if (dep + TU > Fmt<F>::endurance) {
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
        lim0 = admit_limit<F>(idx0);
      }
      if (n1 < idx1) {
        raise_index(fold, idx1 - n1, n1, p1, c1);
        idx1 = n1;
        lim1 = admit_limit<F>(idx1);
      }
    }
    const bool top0 = idx0 == 0, top1 = idx1 == 0;
#pragma unroll
    for (int u = 0; u < TU; ++u) {
      deposit(fold, top0, t0[u], p0);
      deposit(fold, top1, t1[u], p1);
    }
    dep += TU;
  }
  renorm(fold, p0, c0);
  renorm(fold, p1, c1);

  // merge the four waves' values per output: through LDS for the small
  // capacities; the general capacity writes one partial per wave instead
  F* part;
  if constexpr (CAP <= 8) {
    __shared__ F lds[kWaves - 1][128][2 * CAP];
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
      F qp[CAP], qc[CAP];
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
template <int CAP, bool FIXED, class F, bool CPLX = false>
__global__ __launch_bounds__(kThreads) void colaxpy_finish(GemvArgs<F> a) {
  const int fold = FIXED ? CAP : a.fold;
  const int64_t o = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (o >= a.nout) return;
  F p[CAP], c[CAP];
  const F* src = a.partials + o * 2 * fold;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = src[j]; c[j] = src[fold + j]; }
  const int64_t records = (CAP <= 8) ? a.chunks : a.chunks * kWaves;
  for (int64_t s = 1; s < records; ++s) {
    const F* q = a.partials + (s * a.nout + o) * 2 * fold;
    F qp[CAP], qc[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { qp[j] = q[j]; qc[j] = q[fold + j]; }
    add(fold, p, c, qp, qc);
  }
  if (nan_or_inf(p[0])) {
    bool nan = false, pinf = false, ninf = false;
    for (int64_t k = 0; k < a.len; ++k) {
      F t[2];
      if constexpr (CPLX) {
        const F* el = a.A + 2 * (k * a.lda + (o >> 1));
        F r0, r1, i0, i1;
        complex_terms(el[0], el[1], a.x[2 * k], a.x[2 * k + 1], a.conj, r0, r1, i0, i1);
        t[0] = (o & 1) ? i0 : r0;
        t[1] = (o & 1) ? i1 : r1;
      } else {
        t[0] = t[1] = a.A[o + k * a.lda] * scaled_x(a, k);
      }
      for (int q = 0; q < 2; ++q) {
        nan |= (t[q] != t[q]);
        pinf |= (t[q] == (F)__builtin_inf());
        ninf |= (t[q] == -(F)__builtin_inf());
      }
    }
    p[0] = (nan || (pinf && ninf) || !(pinf || ninf)) ? (F)__builtin_nan("")
                                                      : (pinf ? (F)__builtin_inf() : -(F)__builtin_inf());
  }
  finish_output(a, fold, o, p, c);
}

// Persistent row-dot workgroups per CU (compile-time overridable for tuning
// builds).
#ifndef RB_GEMV_ROW_BLOCKS_PER_CU
#define RB_GEMV_ROW_BLOCKS_PER_CU 4
#endif
constexpr int kRowBlocksPerCU = RB_GEMV_ROW_BLOCKS_PER_CU;
// 1: the row-dot walks rows in pairs (rowdot2_kernel), 0: one row per wave.
// On MI355X, 16384^2 column-major T at fold 3: double 5.8 -> 6.4 TB/s with
// pairs; float 5.9 TB/s without, 5.6-5.75 with pairs at 4, 8 or 16 floats per
// lane (pairs off). 2 or 8 workgroups per CU: within noise of 4.
#ifndef RB_GEMV_ROW_PAIR_D
#define RB_GEMV_ROW_PAIR_D 1
#endif
#ifndef RB_GEMV_ROW_PAIR_F
#define RB_GEMV_ROW_PAIR_F 0
#endif
template <class F>
constexpr bool row_pair() { return sizeof(F) == 8 ? RB_GEMV_ROW_PAIR_D : RB_GEMV_ROW_PAIR_F; }

template <int CAP, bool FIXED, class F, bool CPLX>
bool launch_cols(GemvArgs<F>& a, hipStream_t s);

template <int CAP, bool FIXED, class F>
bool launch_gemv(bool rowdot, GemvArgs<F>& a, hipStream_t s) {
  const DeviceInfo* d = device_info();
  if (!d) return false;
  if (rowdot) {
    F* xs = (F*)workspace(s, kSlotGemvX, sizeof(F) * (size_t)a.len);
    if (!xs) return false;
    scale_x_kernel<<<(unsigned)((a.len + kThreads - 1) / kThreads), kThreads, 0, s>>>(a, xs);
    const bool vec = ((uintptr_t)a.A % 16 == 0) && (a.lda % 2 == 0);
    if constexpr (row_pair<F>()) {
      const int64_t pairs = (a.nout + 1) / 2;
      const int64_t blocks = std::min<int64_t>((pairs + kWaves - 1) / kWaves, (int64_t)d->cus * kRowBlocksPerCU);
      if (vec)
        rowdot2_kernel<CAP, FIXED, 0, F><<<(unsigned)blocks, kThreads, 0, s>>>(a, xs);
      else
        rowdot2_kernel<CAP, FIXED, 1, F><<<(unsigned)blocks, kThreads, 0, s>>>(a, xs);
    } else {
      const int64_t blocks = std::min<int64_t>((a.nout + kWaves - 1) / kWaves, (int64_t)d->cus * kRowBlocksPerCU);
      if (vec)
        rowdot_kernel<CAP, FIXED, 0, F><<<(unsigned)blocks, kThreads, 0, s>>>(a, xs);
      else
        rowdot_kernel<CAP, FIXED, 1, F><<<(unsigned)blocks, kThreads, 0, s>>>(a, xs);
    }
    return RB_HIP(hipGetLastError());
  }
  return launch_cols<CAP, FIXED, F, false>(a, s);
}

// col-axpy launch: column chunks so that about four workgroups per CU run,
// then the merge of the per-chunk partials
template <int CAP, bool FIXED, class F, bool CPLX>
bool launch_cols(GemvArgs<F>& a, hipStream_t s) {
  const DeviceInfo* d = device_info();
  if (!d) return false;
  const int64_t rowblocks = (a.nout + 127) / 128;
  int64_t chunks = (RB_GEMV_COL_BLOCKS_PER_CU * (int64_t)d->cus + rowblocks - 1) / rowblocks;
  const int64_t max_chunks = (a.len + 255) / 256;  // keep >= 256 columns per chunk
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  if (chunks > 65535) chunks = 65535;
  a.chunk = (a.len + chunks - 1) / chunks;
  a.chunks = (a.len + a.chunk - 1) / a.chunk;
  const int64_t records = (CAP <= 8) ? a.chunks : a.chunks * kWaves;
  a.partials = (F*)workspace(s, kSlotPartials, sizeof(F) * (size_t)records * a.nout * 2 * a.fold);
  if (!a.partials) return false;
  // vector loads: two adjacent real elements, or one complex element, aligned
  const bool vec = ((uintptr_t)a.A % (2 * sizeof(F)) == 0) && (CPLX || a.lda % 2 == 0);
  dim3 grid((unsigned)rowblocks, (unsigned)a.chunks);
  if (vec)
    colaxpy_kernel<CAP, FIXED, true, F, CPLX><<<grid, kThreads, 0, s>>>(a);
  else
    colaxpy_kernel<CAP, FIXED, false, F, CPLX><<<grid, kThreads, 0, s>>>(a);
  if (!RB_HIP(hipGetLastError())) return false;
  colaxpy_finish<CAP, FIXED, F, CPLX><<<(unsigned)((a.nout + kThreads - 1) / kThreads), kThreads, 0, s>>>(a);
  return RB_HIP(hipGetLastError());
}

template <class F>
bool gemv_common(int fold, char order, char trans, int64_t M, int64_t N, F alpha, const F* A, int64_t lda,
                 const F* x, int64_t incx, F* Yb, int64_t incYb, F* y, int64_t incy, F beta, hipStream_t s) {
  if (fold < 2 || fold > Fmt<F>::max_fold) {
    set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
    return false;
  }
  const bool rowmajor = (order == 'r' || order == 'R');
  const bool notrans = (trans == 'n' || trans == 'N');
  GemvArgs<F> a;
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
  a.conj = 0;
  if (a.nout <= 0) return true;
  // the terms of an output are contiguous for row-major N and column-major T
  const bool rowdot = (rowmajor == notrans);
  switch (fold) {
    case 2: return launch_gemv<2, true>(rowdot, a, s);
    case 3: return launch_gemv<3, true>(rowdot, a, s);
    case 4: return launch_gemv<4, true>(rowdot, a, s);
    default:
      if (fold <= 8) return launch_gemv<8, false>(rowdot, a, s);
      return launch_gemv<Fmt<F>::max_fold, false>(rowdot, a, s);
  }
}

// Complex outputs whose terms are lda apart (column-major N, row-major T/C):
// R[2i] and R[2i+1] (real binned) accumulate the real and imaginary parts of
// output i over sum_k A[k*lda + i] * xs[k] (complex, A conjugated if conj);
// xs: alpha*x, contiguous interleaved.
template <class F>
bool zcols_impl(int fold, int64_t Mo, int64_t No, const F* A, int64_t lda, int conj, const F* xs, F* R,
                hipStream_t s) {
  GemvArgs<F> a;
  a.nout = 2 * Mo;
  a.len = No;
  a.A = A;
  a.lda = lda;
  a.x = xs;
  a.incx = 1;
  a.alpha = F(1);
  a.Yb = R;
  a.incYb = 1;
  a.y = nullptr;
  a.incy = 0;
  a.beta = F(0);
  a.partials = nullptr;
  a.chunk = a.chunks = 0;
  a.fold = fold;
  a.conj = conj;
  if (Mo <= 0 || No <= 0) return true;
  switch (fold) {
    case 2: return launch_cols<2, true, F, true>(a, s);
    case 3: return launch_cols<3, true, F, true>(a, s);
    case 4: return launch_cols<4, true, F, true>(a, s);
    default:
      if (fold <= 8) return launch_cols<8, false, F, true>(a, s);
      return launch_cols<Fmt<F>::max_fold, false, F, true>(a, s);
  }
}

// y = conv(binned(beta*y) (+) sum) in place (rdgemv.c:35-87, rsgemv.c)
template <class F>
bool rgemv_impl(int fold, char order, char trans, int64_t M, int64_t N, F alpha, const F* A, int64_t lda, const F* x,
                int64_t incx, F beta, F* y, int64_t incy, hipStream_t s) {
  if (M <= 0 || N <= 0) return true;  // rdgemv.c:43-45
  const bool notrans = (trans == 'n' || trans == 'N');
  const int64_t nout = notrans ? M : N;
  if (alpha == F(0)) {
    // dbdgemv adds nothing; y still goes through binned(beta*y) and back
    return beta_only(fold, nout, 1, y, incy, 0, beta, s);
  }
  return gemv_common<F>(fold, order, trans, M, N, alpha, A, lda, x, incx, nullptr, 0, y, incy, beta, s);
}

}  // namespace

bool zgemv_cols(int fold, int64_t Mo, int64_t No, const double* A, int64_t lda, int conj, const double* xs,
                double* R, hipStream_t s) {
  return zcols_impl<double>(fold, Mo, No, A, lda, conj, xs, R, s);
}
bool zgemv_cols(int fold, int64_t Mo, int64_t No, const float* A, int64_t lda, int conj, const float* xs, float* R,
                hipStream_t s) {
  return zcols_impl<float>(fold, Mo, No, A, lda, conj, xs, R, s);
}

bool dbdgemv(int fold, char order, char trans, int64_t M, int64_t N, double alpha, const double* A, int64_t lda,
             const double* x, int64_t incx, double* Yb, int64_t incYb, hipStream_t s) {
  if (M <= 0 || N <= 0 || alpha == 0.0) return true;  // dbdgemv.ccog:75-77
  return gemv_common<double>(fold, order, trans, M, N, alpha, A, lda, x, incx, Yb, incYb, nullptr, 0, 0.0, s);
}

bool sbsgemv(int fold, char order, char trans, int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
             const float* x, int64_t incx, float* Yb, int64_t incYb, hipStream_t s) {
  if (M <= 0 || N <= 0 || alpha == 0.0f) return true;  // sbsgemv.ccog:72-74
  return gemv_common<float>(fold, order, trans, M, N, alpha, A, lda, x, incx, Yb, incYb, nullptr, 0, 0.0f, s);
}

bool rdgemv_device(int fold, char order, char trans, int64_t M, int64_t N, double alpha, const double* A,
                   int64_t lda, const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s) {
  return rgemv_impl<double>(fold, order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}

bool rsgemv_device(int fold, char order, char trans, int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                   const float* x, int64_t incx, float beta, float* y, int64_t incy, hipStream_t s) {
  return rgemv_impl<float>(fold, order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}

}  // namespace rb
