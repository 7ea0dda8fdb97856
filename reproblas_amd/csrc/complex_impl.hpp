// complex_impl.hpp -- single-pass binned reductions over complex device vectors:
// sum, dotu, dotc (complex binned results) and asum, the scaled sum of
// squares behind nrm2 (real binned results), for double complex and float
// complex data (interleaved real, imaginary pairs).
//
// Reference path: binnedBLAS_zmzsum (src/binnedBLAS/zmzsum.ccog:57-...),
// zmzdotu / zmzdotc (zmzdotu.ccog:59-..., zmzdotc.ccog:59-...), dmzasum and
// dmzssq (dmzasum.ccog, dmzssq.ccog) and their float-complex counterparts
// (cmcsum.ccog, cmcdotu.ccog, cmcdotc.ccog, smcasum.ccog, smcssq.ccog),
// converted by binned_zzbconv_sub / ddbconv.
//
// A complex binned value is two real binned values with their own indices:
// the real component takes x_r (sum), x_r*y_r and -(x_i*y_i) (dotu), x_r*y_r
// and x_i*y_i (dotc); the imaginary one x_i, x_i*y_r and x_r*y_i (dotu),
// -(x_i*y_r) and x_r*y_i (dotc) (zmzdotu.ccog:10288-10296 and the matching
// scalar section of zmzdotc.ccog). asum deposits |x_r| and |x_i|, nrm2 the
// squares of x_r/scl and x_i/scl, into the two components of a complex
// temporary that the reference adds into one real value at the end; here the
// two channels merge exactly the same way. Each lane keeps both channels in
// registers; each channel's index (and scale) is common to the wave, as in
// the real kernels (wave_sum.hpp), and all merges are exact, so the bits do
// not depend on the tiling.
#pragma once
#include <float.h>

#include <algorithm>

#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"
#include "wave_sum.hpp"

namespace rb {

template <class F>
struct ZArgs {
  int64_t n;  // complex elements
  const F* x;
  int64_t incx;  // in complex elements
  const F* y;
  int64_t incy;
  F* partials;  // gridDim.x records of 2*(2*fold+1) elements
  F* mins;      // nrm2: gridDim.x smallest nonzero |component| of each workgroup
  int replay;   // nrm2: zssq_replay_kernel follows this launch (see zssq_replay_kernel)
  unsigned* ticket;  // [0] the ticket, [1] nrm2: the replay flag
  const F* state_in;  // complex binned (sum/dot) or real binned (asum/ssq)
  F scale_in;
  const F* scale_in_dev;
  F* state_out;
  F* scale_out;
  F* result;  // 2 elements (sum/dot) or 1 (asum/nrm2)
  int fold;
};

// One launch of the complex BLAS1 kernel of a term on prepared arguments
// (mode: 0 unit stride 16-byte aligned, 2 otherwise), compiled in units of
// their own (complex_inst.hip built once per term and type).
#define RB_ZLAUNCH_DECL(NAME)                                                    \
  bool NAME(int fold, int mode, const ZArgs<double>& a, int blocks, hipStream_t s); \
  bool NAME(int fold, int mode, const ZArgs<float>& a, int blocks, hipStream_t s);
RB_ZLAUNCH_DECL(zblas1_launch_sum)
RB_ZLAUNCH_DECL(zblas1_launch_dotu)
RB_ZLAUNCH_DECL(zblas1_launch_dotc)
RB_ZLAUNCH_DECL(zblas1_launch_asum)
RB_ZLAUNCH_DECL(zblas1_launch_ssq)
#undef RB_ZLAUNCH_DECL

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Unit-stride workgroups per CU of the complex reductions and persistent
// workgroups per CU of the complex row-dot gemv (compile-time overridable for
// tuning builds). Measured on MI355X (2^27 double complex, fold 3): 1/CU zsum
// 6.0, dznrm2 5.5, zdotc 6.2 TB/s; 2/CU 6.2, 6.1, 6.5 TB/s. Row-dot gemv
// (zgemv 8192^2 column-major T, fold 3, 110 VGPRs: 4 waves per SIMD resident):
// 2/CU 248 us, 4/CU 216-231 us, 8/CU (one row per wave) 208-221 us.
#ifndef RB_ZBLAS1_BLOCKS_PER_CU
#define RB_ZBLAS1_BLOCKS_PER_CU 2
#endif
#ifndef RB_ZROW_BLOCKS_PER_CU
#define RB_ZROW_BLOCKS_PER_CU 8
#endif


template <class F>
struct ZSpan {
  const F* x;
  const F* y;
  int64_t incx, incy, n, tiles, full_tiles;
};

constexpr bool is_dot(int T) { return T == kZDotU || T == kZDotC; }
// merge rule of a channel
constexpr int chan_kind(int T) { return T == kZSSq ? kSSq : kSum; }
// deposit rule of a channel's raw terms
constexpr int dep_kind(int T) { return T == kZASum ? kASum : (T == kZSSq ? kSSq : kSum); }
// terms per lane per tile and channel
constexpr int nterms(int T, int U) { return is_dot(T) ? 2 * U : U; }

template <int U, class F>
__device__ __forceinline__ ZSpan<F> make_zspan(const F* x, int64_t incx, const F* y, int64_t incy, int64_t n) {
  ZSpan<F> s;
  s.x = x;
  s.y = y;
  s.incx = incx;
  s.incy = incy;
  s.n = n;
  s.tiles = (n + 64 * U - 1) / (64 * U);
  s.full_tiles = n / (64 * U);
  return s;
}

// What the caller knows of tile t: nothing (the load checks), that every
// element is in range (no checks, no branch: the registers a pipelined caller
// prefetches into are then written by the loads alone, so the loads stay in
// flight until the deposit that needs them), or that it is the partial last
// tile.
constexpr int kTileAny = 0, kTileFull = 1, kTilePartial = 2;

// The U complex elements of tile t for this lane. MODE 0: unit stride,
// 16-byte aligned: 16-byte loads (one double complex or two float complex
// elements each, lanes on consecutive addresses), non-temporal for x and for
// y unless Y_REUSED (a vector every wave re-reads, kept in L2); MODE 1: unit
// stride, element loads; MODE 2: any stride. Out-of-range elements read as 0.
template <int T, int U, int MODE, int KNOWN = kTileAny, bool Y_REUSED = false, class F>
__device__ __forceinline__ void load_ztile(const ZSpan<F>& a, int64_t t, int lane, F (&xr)[U], F (&xi)[U],
                                           F (&yr)[U], F (&yi)[U]) {
  const int64_t base = t * (64 * U);
  const bool full = KNOWN == kTileFull || (KNOWN == kTileAny && t < a.full_tiles);
  if (MODE == 0 && full) {
    if constexpr (sizeof(F) == 8) {
      const d2* px = reinterpret_cast<const d2*>(a.x + 2 * base);
      const d2* py = reinterpret_cast<const d2*>(a.y + 2 * base);
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const d2 v = __builtin_nontemporal_load(px + k * 64 + lane);
        xr[k] = v.x;
        xi[k] = v.y;
        if (is_dot(T)) {
          const d2 w = Y_REUSED ? py[k * 64 + lane] : __builtin_nontemporal_load(py + k * 64 + lane);
          yr[k] = w.x;
          yi[k] = w.y;
        }
      }
    } else {
      const f4* px = reinterpret_cast<const f4*>(a.x + 2 * base);
      const f4* py = reinterpret_cast<const f4*>(a.y + 2 * base);
#pragma unroll
      for (int k = 0; k < U / 2; ++k) {
        const f4 v = __builtin_nontemporal_load(px + k * 64 + lane);
        xr[2 * k] = v.x;
        xi[2 * k] = v.y;
        xr[2 * k + 1] = v.z;
        xi[2 * k + 1] = v.w;
        if (is_dot(T)) {
          const f4 w = Y_REUSED ? py[k * 64 + lane] : __builtin_nontemporal_load(py + k * 64 + lane);
          yr[2 * k] = w.x;
          yi[2 * k] = w.y;
          yr[2 * k + 1] = w.z;
          yi[2 * k + 1] = w.w;
        }
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int64_t e = base + k * 64 + lane;
    const bool in = full || e < a.n;
    const int64_t ix = (MODE == 2) ? e * a.incx : e;
    xr[k] = in ? a.x[2 * ix] : F(0);
    xi[k] = in ? a.x[2 * ix + 1] : F(0);
    if (is_dot(T)) {
      const int64_t iy = (MODE == 2) ? e * a.incy : e;
      yr[k] = in ? a.y[2 * iy] : F(0);
      yi[k] = in ? a.y[2 * iy + 1] : F(0);
    }
  }
}

// The terms of one channel (raw values for asum/ssq, whose deposit applies
// |.| or the scaled square).
template <int T, int U, class F>
__device__ __forceinline__ void channel_terms(int ch, const F (&xr)[U], const F (&xi)[U], const F (&yr)[U],
                                              const F (&yi)[U], F (&t)[nterms(T, U)]) {
#pragma unroll
  for (int k = 0; k < U; ++k) {
    if (!is_dot(T)) {
      t[k] = ch == 0 ? xr[k] : xi[k];
    } else if (T == kZDotU) {
      if (ch == 0) {
        t[k] = xr[k] * yr[k];
        t[U + k] = (xi[k] * yi[k]) * F(-1);
      } else {
        t[k] = xi[k] * yr[k];
        t[U + k] = xr[k] * yi[k];
      }
    } else {
      if (ch == 0) {
        t[k] = xr[k] * yr[k];
        t[U + k] = xi[k] * yi[k];
      } else {
        t[k] = (xi[k] * yr[k]) * F(-1);
        t[U + k] = xr[k] * yi[k];
      }
    }
  }
}

// Class of a channel's terms over the tiles this wave walked (NaN if any
// term is NaN or both infinities occur, else the infinity that occurs).
template <int T, int U, int MODE, class F>
__device__ F zwave_term_class(const ZSpan<F>& a, int64_t first, int64_t step, int lane, int ch) {
  constexpr int NT = nterms(T, U);
  bool nan = false, pinf = false, ninf = false;
  for (int64_t t = first; t < a.tiles; t += step) {
    F xr[U], xi[U], yr[U], yi[U], v[NT];
    load_ztile<T, U, MODE>(a, t, lane, xr, xi, yr, yi);
    channel_terms<T, U>(ch, xr, xi, yr, yi, v);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const F u = (T == kZASum || T == kZSSq) ? fabs(v[k]) : v[k];
      nan |= (u != u);
      pinf |= (u == (F)__builtin_inf());
      ninf |= (u == -(F)__builtin_inf());
    }
  }
  const bool any_nan = __any(nan), any_p = __any(pinf), any_n = __any(ninf);
  if (any_nan || (any_p && any_n) || !(any_p || any_n)) return (F)__builtin_nan("");
  return any_p ? (F)__builtin_inf() : -(F)__builtin_inf();
}

template <int CAP, int T, int U, class F>
__device__ __forceinline__ void deposit_ztile(int fold, const F (&xr)[U], const F (&xi)[U], const F (&yr)[U],
                                              const F (&yi)[U], WaveRun<F> (&w)[2], State<CAP, F> (&s)[2]) {
  constexpr int NT = nterms(T, U);
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    F t[NT];
    channel_terms<T, U>(ch, xr, xi, yr, yi, t);
    deposit_tile<CAP, dep_kind(T), NT>(fold, t, t, w[ch], s[ch]);
  }
}

// One wave reduces the tiles t0, t0 + tstep, ... into both channels.
template <int CAP, int T, int U, int MODE, class F>
__device__ __forceinline__ void zwave_accumulate(const ZSpan<F>& a, int fold, int64_t t0, int64_t tstep, int lane,
                                                 State<CAP, F> (&s)[2], F scale0 = F(0), F* min_nz = nullptr,
                                                 bool track_min = false) {
  WaveRun<F> w[2];
  // nrm2: both channels square at no less than the caller's scale (the
  // parallel schedule; zssq_replay_kernel replays the reference's where it
  // matters)
  wave_start(fold, w[0], s[0], scale0, min_nz != nullptr && track_min);
  wave_start(fold, w[1], s[1], scale0, min_nz != nullptr && track_min);
  F axr[U], axi[U], ayr[U], ayi[U], bxr[U], bxi[U], byr[U], byi[U];
  int64_t t = t0;
  if (t < a.tiles) load_ztile<T, U, MODE>(a, t, lane, axr, axi, ayr, ayi);
  while (t < a.tiles) {
    int64_t tn = t + tstep;
    if (tn < a.tiles) load_ztile<T, U, MODE>(a, tn, lane, bxr, bxi, byr, byi);
    deposit_ztile<CAP, T, U>(fold, axr, axi, ayr, ayi, w, s);
    if (tn >= a.tiles) break;
    t = tn;
    tn = t + tstep;
    if (tn < a.tiles) load_ztile<T, U, MODE>(a, tn, lane, axr, axi, ayr, ayi);
    deposit_ztile<CAP, T, U>(fold, bxr, bxi, byr, byi, w, s);
    t = tn;
  }
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    wave_butterfly(fold, w[ch], s[ch]);
    if (nan_or_inf(uniform_value(s[ch].p[0]))) s[ch].p[0] = zwave_term_class<T, U, MODE>(a, t0, tstep, lane, ch);
  }
  if (min_nz) {
    // nrm2: smallest nonzero |component| of the wave, every lane
    F m = w[0].mn < w[1].mn ? w[0].mn : w[1].mn;
#pragma unroll
    for (int st = 0; st < 6; ++st) {
      const F o = lane_partner(m, st);
      m = o < m ? o : m;
    }
    *min_nz = m;
  }
}

template <int CAP, int T, class F>
__device__ __forceinline__ void merge_wg(int fold, int wave, int lane, F (*lds)[2][2 * CAP + 1],
                                         State<CAP, F> (&s)[2], State<CAP, F> (&acc)[2]) {
  if (lane == 0) {
    store_state(lds[wave][0], fold, s[0]);
    store_state(lds[wave][1], fold, s[1]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      State<CAP, F> o;
      load_state(lds[0][ch], fold, acc[ch]);
      for (int w = 1; w < kWaves; ++w) {
        load_state(lds[w][ch], fold, o);
        merge<CAP, chan_kind(T)>(fold, acc[ch], o);
      }
    }
  }
}

template <int CAP, bool FIXED, int T, int U, int MODE, class F>
__global__ __launch_bounds__(kThreads) void zblas1_kernel(ZArgs<F> a) {
  constexpr int KM = chan_kind(T);
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t wave_gid = (int64_t)blockIdx.x * kWaves + wave;
  const int64_t nwaves = (int64_t)gridDim.x * kWaves;
  const ZSpan<F> sp = make_zspan<U>(a.x, a.incx, a.y, a.incy, a.n);
  State<CAP, F> s[2];
  const F scale0 = (T == kZSSq && a.state_in) ? (a.scale_in_dev ? *a.scale_in_dev : a.scale_in) : F(0);
  F wmin = (F)__builtin_inf();  // nrm2: smallest nonzero |component| of the wave
  zwave_accumulate<CAP, T, U, MODE>(sp, fold, wave_gid, nwaves, lane, s, scale0, T == kZSSq ? &wmin : nullptr,
                                    T == kZSSq && a.replay);

  __shared__ F lds[kWaves][2][2 * CAP + 1];
  __shared__ F lds_min[kWaves];
  __shared__ bool last;
  const int rec = 2 * (2 * fold + 1);
  {
    State<CAP, F> acc[2];
    if (T == kZSSq && lane == 0) lds_min[wave] = wmin;  // read after merge_wg's barrier
    merge_wg<CAP, T>(fold, wave, lane, lds, s, acc);
    if (threadIdx.x == 0) {
      if (T == kZSSq) {
        F mn = lds_min[0];
        for (int w = 1; w < kWaves; ++w) mn = lds_min[w] < mn ? lds_min[w] : mn;
        a.mins[blockIdx.x] = mn;
      }
      store_state(a.partials + (int64_t)blockIdx.x * rec, fold, acc[0]);
      store_state(a.partials + (int64_t)blockIdx.x * rec + (2 * fold + 1), fold, acc[1]);
      __threadfence();
      const unsigned ticket = atomicAdd(a.ticket, 1u);
      last = (ticket == gridDim.x - 1);
    }
  }
  __syncthreads();
  if (!last) return;

  // Last workgroup: merge every partial per channel, fold in the caller's
  // value, convert.
  __threadfence();
  State<CAP, F> acc[2];
  bool have = false;
  F mn = (F)__builtin_inf();  // nrm2: smallest nonzero |component| of the grid
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads) {
    if (T == kZSSq) {
      const F bm = __builtin_nontemporal_load(a.mins + b);
      mn = bm < mn ? bm : mn;
    }
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      State<CAP, F> o;
      const F* src = a.partials + (int64_t)b * rec + ch * (2 * fold + 1);
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) {
          o.p[j] = __builtin_nontemporal_load(src + j);
          o.c[j] = __builtin_nontemporal_load(src + fold + j);
        }
      o.scl = __builtin_nontemporal_load(src + 2 * fold);
      if (!have)
        acc[ch] = o;
      else
        merge<CAP, KM>(fold, acc[ch], o);
    }
    have = true;
  }
  if (!have) {
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      set_bins(fold, Fmt<F>::max_index, acc[ch].p, acc[ch].c);
      acc[ch].scl = F(0);
    }
  }
  wave_merge_general<CAP, KM>(fold, acc[0]);
  wave_merge_general<CAP, KM>(fold, acc[1]);
  F gmin = (F)__builtin_inf();  // nrm2, thread 0
  if (T == kZSSq) {
#pragma unroll
    for (int st = 0; st < 6; ++st) {
      const F o = lane_partner(mn, st);
      mn = o < mn ? o : mn;
    }
    if (lane == 0) lds_min[wave] = mn;
  }
  __syncthreads();
  if (T == kZSSq && threadIdx.x == 0) {
    gmin = lds_min[0];
    for (int w = 1; w < kWaves; ++w) gmin = lds_min[w] < gmin ? lds_min[w] : gmin;
  }
  State<CAP, F> tot[2];
  merge_wg<CAP, T>(fold, wave, lane, lds, acc, tot);
  if (threadIdx.x != 0) return;

  if (is_dot(T) || T == kZSum) {
    // complex result: each channel like the real kernel, interleaved layout
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      State<CAP, F> y;
      if (a.state_in) {
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) {
            y.p[j] = a.state_in[2 * j + ch];
            y.c[j] = a.state_in[2 * fold + 2 * j + ch];
          }
        renorm(fold, y.p, y.c);
      } else {
        set_zero(fold, y.p, y.c);
      }
      y.scl = F(0);
      if (a.n == 0) {
      } else if (y.p[0] == F(0)) {
        y = tot[ch];
      } else {
        merge<CAP, kSum>(fold, y, tot[ch]);
      }
      renorm(fold, y.p, y.c);
      if (a.state_out) {
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) {
            a.state_out[2 * j + ch] = y.p[j];
            a.state_out[2 * fold + 2 * j + ch] = y.c[j];
          }
      }
      if (a.result) a.result[ch] = convert(fold, y.p, y.c);
    }
  } else {
    // real result: the two channels are one sum (exact merge), then as the
    // real kernel folds in the caller's value
    State<CAP, F> v = tot[0];
    merge<CAP, KM>(fold, v, tot[1]);
    State<CAP, F> y;
    if (a.state_in) {
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) {
          y.p[j] = a.state_in[j];
          y.c[j] = a.state_in[fold + j];
        }
      renorm(fold, y.p, y.c);
    } else {
      set_zero(fold, y.p, y.c);
    }
    y.scl = a.scale_in_dev ? *a.scale_in_dev : a.scale_in;
    // an empty caller value keeps its scale (dmzssq squares at
    // max(scaleY, dscale(amax))): always the scaled merge
    if (a.n > 0) merge<CAP, KM>(fold, y, v);
    renorm(fold, y.p, y.c);
    if (T == kZSSq && a.replay && a.n > 0 && y.scl > F(0)) {
      // The reference's one-call schedule where it can differ: a finite
      // value and a component whose square at the final scale is subnormal
      // (its earlier blocks squared it at smaller scales, in normal range).
      // Unlike the real path, no lowest-bin test: the two channels share
      // one index and zmdrescale's zero test, so the last slot can move at
      // coarser bins too (fold 27, lowest bin 50). zssq_replay_kernel (next
      // on the stream) then writes the outputs from the caller's untouched
      // value.
      const bool fin = !nan_or_inf(y.p[0]) && y.p[0] != F(0);
      const F tiny = fp_from_bits<F>((typename Fmt<F>::U)(Lay<F>::ieee_bias - (Fmt<F>::max_exp / 2 - 1))
                                     << Lay<F>::mbits);
      if (fin && gmin / y.scl < tiny) {
        a.ticket[1] = 1u;
        *a.ticket = 0u;
        return;
      }
    }
    F scale = y.scl;
    if (T == kZSSq && nan_or_inf(y.p[0])) scale = F(1);
    if (a.state_out) {
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) {
          a.state_out[j] = y.p[j];
          a.state_out[fold + j] = y.c[j];
        }
    }
    if (a.scale_out) *a.scale_out = scale;
    if (a.result) {
      const F c = convert(fold, y.p, y.c);
      *a.result = (T == kZSSq) ? (F)((double)scale * sqrt((double)c)) : c;
    }
  }
  *a.ticket = 0u;
}

// The reference's complex sum of squares in one call (dmzssq.ccog:59-135,
// smcssq.ccog): blocks of 1024 (float: 256) elements in order; before each,
// a renorm when the deposits would pass the endurance (every second block),
// then the block's scale dscale(max |component|) -- the accumulator (both
// channels) rescaled by zmdrescale when it grows -- and an index update of
// both channels for the block's largest scaled square; then the squares of
// the real parts into channel 0 and of the imaginary parts into channel 1.
// The binned state therefore passes through the reference's exact states,
// unrenormalised ones included, at every rescale. One workgroup: block by
// block, every lane deposits its share of the block's squares into its own
// copy of the accumulator's primaries. A slot's slices depend only on the
// terms and on the slot's ulp, which is the same for the copies and the
// accumulator (within the endurance every primary stays in its binade; the
// tiny primaries of slots past the bin table, rescaled below 2^-(MAX_EXP-3),
// all have the subnormal ulp), so the sum over lanes of each copy's change,
// counted in ulps (sign times magnitude bits: linear there, across zero
// too), added to the accumulator gives the reference's primaries. At the
// end the two channels go into the caller's value by two scaled adds
// (dmdmaddsq), real part first. Runs only when the streaming kernel flagged
// the region (ticket[1]); writes the outputs and clears the flag.
// A value in ulps of the binades with the ulp of the smallest subnormal
// (and, by differences, of any one binade): sign times magnitude bits.
template <class F>
__device__ __forceinline__ int64_t ulp_count(F v) {
  const int64_t m = (int64_t)(bits(v) & ~((typename Fmt<F>::U)1 << (8 * sizeof(F) - 1)));
  return v < F(0) ? -m : m;
}
template <class F>
__device__ __forceinline__ F from_ulp_count(int64_t k) {
  return k < 0 ? -fp_from_bits<F>((typename Fmt<F>::U)(-k)) : fp_from_bits<F>((typename Fmt<F>::U)k);
}

template <int CAP, bool FIXED, class F>
__global__ __launch_bounds__(kThreads) void zssq_replay_kernel(ZArgs<F> a) {
  if (a.ticket[1] == 0u) return;
  constexpr int64_t NB = sizeof(F) == sizeof(double) ? 1024 : 256;  // dmzssq.ccog:36, smcssq.ccog:36
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  __shared__ F red[kWaves];
  __shared__ F bc_inv;
  __shared__ int bc_idx[2];
  __shared__ F bc_p[2][CAP];  // the accumulator's primaries before the block's deposits
  __shared__ bool bc_skip;
  __shared__ unsigned long long dsum[2][CAP];
  State<CAP, F> ch[2];  // thread 0: the reference's accumulator
  F scl = F(0);
  int deposits = 0;
  if (threadIdx.x == 0) {
    set_zero(fold, ch[0].p, ch[0].c);
    set_zero(fold, ch[1].p, ch[1].c);
  }
  const int64_t nb = (a.n + NB - 1) / NB;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t e0 = b * NB, e1 = (b + 1) * NB < a.n ? (b + 1) * NB : a.n;
    F m = F(0);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kThreads) {
      const int64_t ix = 2 * e * a.incx;
      m = fmax(m, fmax(fabs(a.x[ix]), fabs(a.x[ix + 1])));
    }
    m = wave_max(m);
    if (lane == 0) red[wave] = m;
    for (int i = threadIdx.x; i < 2 * CAP; i += kThreads) (&dsum[0][0])[i] = 0ull;
    __syncthreads();
    if (threadIdx.x == 0) {
      F amax = red[0];
      for (int w = 1; w < kWaves; ++w) amax = fmax(amax, red[w]);
      if (deposits + (int)(e1 - e0) > Fmt<F>::endurance) {
        renorm(fold, ch[0].p, ch[0].c);
        renorm(fold, ch[1].p, ch[1].c);
        deposits = 0;
      }
      const F ns = dscale(amax);
      if (ns > scl) {
        if (scl > F(0)) zrescale(fold, ns, scl, ch[0].p, ch[0].c, ch[1].p, ch[1].c);
        scl = ns;
      }
      F am = amax / scl;
      am = am * am;
      update(fold, am, ch[0].p, ch[0].c);
      update(fold, am, ch[1].p, ch[1].c);
      bc_inv = F(1) / scl;  // exact: scales are powers of two
      bc_skip = amax == F(0) || ch[0].p[0] == F(0) || ch[1].p[0] == F(0);
      bc_idx[0] = dmindex(ch[0].p[0]);
      bc_idx[1] = dmindex(ch[1].p[0]);
#pragma unroll
      for (int c = 0; c < 2; ++c)
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) bc_p[c][j] = ch[c].p[j];
    }
    __syncthreads();
    if (!bc_skip) {  // (an all-zero block deposits nothing but zeros)
      const F inv = bc_inv;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int I = bc_idx[c];
        F lp[CAP];
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) lp[j] = bc_p[c][j];
        for (int64_t e = e0 + threadIdx.x; e < e1; e += kThreads) {
          F v = a.x[2 * e * a.incx + c] * inv;
          v = v * v;
          deposit(fold, I == 0, v, lp);
        }
        int64_t d[CAP];
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) d[j] = ulp_count(lp[j]) - ulp_count(bc_p[c][j]);
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) {
#pragma unroll
            for (int st = 0; st < 6; ++st) d[j] += lane_partner_i64(d[j], st);
          }
        if (lane == 0) {
RB_UNROLL_CAP
          for (int j = 0; j < CAP; ++j)
            if (j < fold) atomicAdd(&dsum[c][j], (unsigned long long)d[j]);
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (!bc_skip) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
RB_UNROLL_CAP
          for (int j = 0; j < CAP; ++j)
            if (j < fold) ch[c].p[j] = from_ulp_count<F>(ulp_count(ch[c].p[j]) + (int64_t)dsum[c][j]);
        }
      }
      deposits += (int)(e1 - e0);
    }
    __syncthreads();  // red, bc_* and dsum are reused by the next block
  }
  if (threadIdx.x != 0) return;
  renorm(fold, ch[0].p, ch[0].c);
  renorm(fold, ch[1].p, ch[1].c);
  State<CAP, F> y;
  if (a.state_in) {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) {
        y.p[j] = a.state_in[j];
        y.c[j] = a.state_in[fold + j];
      }
  } else {
    set_zero(fold, y.p, y.c);
  }
  y.scl = a.scale_in_dev ? *a.scale_in_dev : a.scale_in;
  ch[0].scl = ch[1].scl = scl;
  merge<CAP, kSSq>(fold, y, ch[0]);
  merge<CAP, kSSq>(fold, y, ch[1]);
  if (a.state_out) {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) {
        a.state_out[j] = y.p[j];
        a.state_out[fold + j] = y.c[j];
      }
  }
  if (a.scale_out) *a.scale_out = y.scl;
  if (a.result) *a.result = (F)((double)y.scl * sqrt((double)convert(fold, y.p, y.c)));
  a.ticket[1] = 0u;
}

// Complex elements per lane per tile: 128 bytes of x per lane (64 of each
// operand for the dots).
template <int T, class F>
constexpr int z_units() {
  return (is_dot(T) ? 64 : 128) / (2 * (int)sizeof(F));
}

// ------------------------------------------------------ complex row gemv --
// Rows of op(A) contiguous (row-major N, column-major T/C): output o is the
// complex dotu (conj: dotc) of row o, A + 2*o*lda, with xs = alpha*x. A
// persistent grid of waves owns rows o, o + W, ... and walks their full tiles
// as one sequence (the next tile -- at a row's end, the first of the next row
// -- in flight while the current one is deposited), as the real row-dot gemv;
// a row's partial last tile is loaded and deposited on its own at its end.
// Output o accumulates into the real binned values R[2o] (real part) and
// R[2o+1] (imaginary part).
template <class F>
struct ZRowArgs {
  int64_t nout, len;
  const F* A;
  int64_t lda;  // complex elements
  const F* xs;  // alpha*x, contiguous interleaved
  F* R;
  int fold;
};

template <int U, class F>
__device__ __forceinline__ ZSpan<F> zrow_span(const ZRowArgs<F>& a, int64_t o) {
  return make_zspan<U>(a.A + 2 * o * a.lda, 1, a.xs, 1, a.len);
}

template <int CAP, int T, int U, int MODE, class F>
__device__ __forceinline__ void zfinish_row(const ZRowArgs<F>& a, int fold, int64_t o, int lane,
                                            WaveRun<F> (&w)[2], State<CAP, F> (&s)[2]) {
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    wave_butterfly(fold, w[ch], s[ch]);
    if (nan_or_inf(uniform_value(s[ch].p[0])))
      s[ch].p[0] = zwave_term_class<T, U, MODE>(zrow_span<U>(a, o), 0, 1, lane, ch);
    if (lane == 0) {
      F* r = a.R + (2 * o + ch) * 2 * fold;
      F py[CAP], cy[CAP];
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) {
          py[j] = r[j];
          cy[j] = r[fold + j];
        }
      renorm(fold, py, cy);
      add(fold, py, cy, s[ch].p, s[ch].c);
      renorm(fold, py, cy);
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) {
          r[j] = py[j];
          r[fold + j] = cy[j];
        }
    }
  }
}

// Complex elements per lane per tile of the row-dot gemv (compile-time
// overridable for tuning builds). 32 bytes of each operand per lane: on
// MI355X, zgemv 8192^2 column-major T/C at fold 3 runs 208-221 us with 2
// double complex per lane, 222-231 us with 4 and 241 us with 1.
#ifndef RB_ZROW_U_D
#define RB_ZROW_U_D 2
#endif
#ifndef RB_ZROW_U_F
#define RB_ZROW_U_F 4
#endif
template <class F>
constexpr int zrow_units() {
  return sizeof(F) == sizeof(double) ? RB_ZROW_U_D : RB_ZROW_U_F;
}

// The end of row o: its partial last tile (if any, loaded and deposited on
// its own), the merge into R, and fresh states for the next row.
template <int CAP, int T, int U, int MODE, class F>
__device__ __forceinline__ void zend_row(const ZRowArgs<F>& a, int fold, int64_t o, int64_t full, bool part,
                                         int lane, WaveRun<F> (&w)[2], State<CAP, F> (&s)[2]) {
  if (part) {
    F cxr[U], cxi[U], cyr[U], cyi[U];
    load_ztile<T, U, MODE, kTilePartial, true>(zrow_span<U>(a, o), full, lane, cxr, cxi, cyr, cyi);
    deposit_ztile<CAP, T, U>(fold, cxr, cxi, cyr, cyi, w, s);
  }
  zfinish_row<CAP, T, U, MODE>(a, fold, o, lane, w, s);
  wave_start(fold, w[0], s[0]);
  wave_start(fold, w[1], s[1]);
}

template <int CAP, bool FIXED, bool CONJ, int MODE, class F>
__global__ __launch_bounds__(kThreads) void zrowdot_kernel(ZRowArgs<F> a) {
  constexpr int T = CONJ ? kZDotC : kZDotU;
  constexpr int U = zrow_units<F>();
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * kWaves;
  int64_t o = (int64_t)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (o >= a.nout) return;  // whole wave leaves together
  const int64_t full = a.len / (64 * U);
  const bool part = full * (64 * U) < a.len;
  State<CAP, F> s[2];
  WaveRun<F> w[2];
  wave_start(fold, w[0], s[0]);
  wave_start(fold, w[1], s[1]);
  if (full == 0) {  // rows shorter than a tile: one partial tile each
    for (; o < a.nout; o += nw) zend_row<CAP, T, U, MODE>(a, fold, o, full, part, lane, w, s);
    return;
  }
  // The full tiles of rows o, o + W, ... as one sequence, the next one in
  // flight while the current one is deposited. The prefetch is unconditional
  // (past the last row it reloads tile 0 of the current row, unused).
  F axr[U], axi[U], ayr[U], ayi[U], bxr[U], bxi[U], byr[U], byi[U];
  int64_t t = 0;
  load_ztile<T, U, MODE, kTileFull, true>(zrow_span<U>(a, o), t, lane, axr, axi, ayr, ayi);
  while (true) {
    int64_t on = o, tn = t + 1;
    if (tn == full) {
      on = o + nw;
      tn = 0;
    }
    load_ztile<T, U, MODE, kTileFull, true>(zrow_span<U>(a, on < a.nout ? on : o), tn, lane, bxr, bxi, byr, byi);
    deposit_ztile<CAP, T, U>(fold, axr, axi, ayr, ayi, w, s);
    if (tn == 0) zend_row<CAP, T, U, MODE>(a, fold, o, full, part, lane, w, s);
    if (on >= a.nout) break;
    o = on;
    t = tn;
    on = o;
    tn = t + 1;
    if (tn == full) {
      on = o + nw;
      tn = 0;
    }
    load_ztile<T, U, MODE, kTileFull, true>(zrow_span<U>(a, on < a.nout ? on : o), tn, lane, axr, axi, ayr, ayi);
    deposit_ztile<CAP, T, U>(fold, bxr, bxi, byr, byi, w, s);
    if (tn == 0) zend_row<CAP, T, U, MODE>(a, fold, o, full, part, lane, w, s);
    if (on >= a.nout) break;
    o = on;
    t = tn;
  }
}

template <int CAP, bool FIXED, class F>
bool zlaunch_rows(const ZRowArgs<F>& a, bool conj, int mode, int blocks, hipStream_t s) {
  if (conj) {
    if (mode == 0)
      zrowdot_kernel<CAP, FIXED, true, 0, F><<<blocks, kThreads, 0, s>>>(a);
    else
      zrowdot_kernel<CAP, FIXED, true, 2, F><<<blocks, kThreads, 0, s>>>(a);
  } else {
    if (mode == 0)
      zrowdot_kernel<CAP, FIXED, false, 0, F><<<blocks, kThreads, 0, s>>>(a);
    else
      zrowdot_kernel<CAP, FIXED, false, 2, F><<<blocks, kThreads, 0, s>>>(a);
  }
  return RB_HIP(hipGetLastError());
}

template <class F>
bool zrows_impl(int fold, int64_t Mo, int64_t No, const F* A, int64_t lda, int conj, const F* xs, F* R,
                hipStream_t s) {
  if (Mo <= 0 || No <= 0) return true;
  const DeviceInfo* dev = device_info();
  if (!dev) return false;
  ZRowArgs<F> a;
  a.nout = Mo;
  a.len = No;
  a.A = A;
  a.lda = lda;
  a.xs = xs;
  a.R = R;
  a.fold = fold;
  // 16-byte loads need every row start 16-byte aligned
  const bool aligned = ((uintptr_t)A % 16 == 0) && ((uintptr_t)xs % 16 == 0) &&
                       ((2 * sizeof(F) * (size_t)lda) % 16 == 0);
  const int mode = aligned ? 0 : 1;
  const int64_t blocks = std::min<int64_t>((Mo + kWaves - 1) / kWaves, (int64_t)dev->cus * RB_ZROW_BLOCKS_PER_CU);
  // fold 3 (the default) has a dedicated instance; the others run the
  // general ones (register slots up to 8, then private memory)
  if (fold == 3) return zlaunch_rows<3, true>(a, conj, mode, (int)blocks, s);
  if (fold <= 8) return zlaunch_rows<8, false>(a, conj, mode, (int)blocks, s);
  return zlaunch_rows<Fmt<F>::max_fold, false>(a, conj, mode, (int)blocks, s);
}

template <int CAP, bool FIXED, int T, class F>
bool zlaunch_mode(int mode, const ZArgs<F>& a, int blocks, hipStream_t s) {
  constexpr int U = z_units<T, F>();
  // unaligned unit-stride operands take the element-load (strided) kernel
  if (mode == 0)
    zblas1_kernel<CAP, FIXED, T, U, 0, F><<<blocks, kThreads, 0, s>>>(a);
  else
    zblas1_kernel<CAP, FIXED, T, U, 2, F><<<blocks, kThreads, 0, s>>>(a);
  if constexpr (T == kZSSq) {
    if (a.replay) {
      if (!RB_HIP(hipGetLastError())) return false;
      zssq_replay_kernel<CAP, FIXED, F><<<1, kThreads, 0, s>>>(a);
    }
  }
  return RB_HIP(hipGetLastError());
}

// The kernel instance of term T for a fold (fold 3 has its own instance).
template <int T, class F>
bool zlaunch_fold_term(int fold, int mode, const ZArgs<F>& a, int blocks, hipStream_t s) {
  if (fold == 3) return zlaunch_mode<3, true, T>(mode, a, blocks, s);
  if (fold <= 8) return zlaunch_mode<8, false, T>(mode, a, blocks, s);
  return zlaunch_mode<Fmt<F>::max_fold, false, T>(mode, a, blocks, s);
}

int64_t zunits_of(ZTerm term, size_t esize) {
  const bool dot = (term == kZDotU || term == kZDotC);
  return (dot ? 64 : 128) / (2 * (int64_t)esize);
}

template <class F>
bool zblas1_impl(int fold, ZTerm term, int64_t n, const F* x, int64_t incx, const F* y, int64_t incy,
                 const F* state_in, F scale_in, const F* scale_in_dev, F* state_out, F* scale_out, F* result,
                 hipStream_t s) {
  if (fold < 2 || fold > Fmt<F>::max_fold) {
    set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
    return false;
  }
  const DeviceInfo* dev = device_info();
  if (!dev) return false;
  const bool dot = (term == kZDotU || term == kZDotC);
  ZArgs<F> a;
  a.n = n < 0 ? 0 : n;
  a.x = x;
  a.incx = incx;
  a.y = y;
  a.incy = incy;
  a.state_in = state_in;
  a.scale_in = scale_in;
  a.scale_in_dev = scale_in_dev;
  a.state_out = state_out;
  a.scale_out = scale_out;
  a.result = result;
  a.fold = fold;
  const int64_t per_tile = 64 * zunits_of(term, sizeof(F));
  const int64_t tiles = (a.n + per_tile - 1) / per_tile;
  int mode;
  const bool unit = (incx == 1) && (!dot || incy == 1);
  if (!unit) {
    mode = 2;
  } else {
    const bool aligned = ((uintptr_t)x % 16 == 0) && (!dot || (uintptr_t)y % 16 == 0);
    mode = aligned ? 0 : 1;
  }
  // grid: RB_ZBLAS1_BLOCKS_PER_CU workgroups per CU for aligned unit stride,
  // four otherwise; at least 2048 terms per wave for short vectors
  const int64_t min_tiles = std::max<int64_t>(1, 2048 / (per_tile * (dot ? 2 : 1)));
  int64_t blocks = (tiles + kWaves * min_tiles - 1) / (kWaves * min_tiles);
  const int64_t cap = (int64_t)dev->cus * (mode == 2 ? 4 : RB_ZBLAS1_BLOCKS_PER_CU);
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  // nrm2: the workgroups' smallest nonzero |component| after their records
  const size_t rec_bytes = sizeof(F) * (size_t)blocks * 2 * (2 * fold + 1);
  a.partials = (F*)workspace(s, kSlotPartials, rec_bytes + (term == kZSSq ? sizeof(F) * (size_t)blocks : 0));
  a.mins = a.partials ? (F*)((char*)a.partials + rec_bytes) : nullptr;
  // nrm2: the region where the reference's block schedule matters (see
  // zssq_replay_kernel): without a caller's value the scaled squares' top
  // index is 23..27 (double) or 7..11 (float), so only folds >= 25 (double)
  // or >= 9 (float) reach it; a caller's value of larger scale can put the
  // index anywhere.
  a.replay = term == kZSSq && a.n > 0 && (state_in != nullptr || fold >= (sizeof(F) == 8 ? 25 : 9));
  a.ticket = tickets(s, 2);
  if (!a.partials || !a.ticket) return false;
  switch (term) {
    case kZSum: return zblas1_launch_sum(fold, mode, a, (int)blocks, s);
    case kZDotU: return zblas1_launch_dotu(fold, mode, a, (int)blocks, s);
    case kZDotC: return zblas1_launch_dotc(fold, mode, a, (int)blocks, s);
    case kZASum: return zblas1_launch_asum(fold, mode, a, (int)blocks, s);
    default: return zblas1_launch_ssq(fold, mode, a, (int)blocks, s);
  }
}

}  // namespace
}  // namespace rb
