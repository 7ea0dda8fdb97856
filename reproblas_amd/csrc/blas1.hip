// blas1.hip -- single-pass binned reductions over device vectors:
// sum, asum, dot and the scaled sum of squares behind nrm2.
//
// Reference path: binnedBLAS_dmdsum (src/binnedBLAS/dmdsum.ccog:57-89),
// dmdasum (dmdasum.ccog:57-85), dmddot (dmddot.ccog:59-88) and dmdssq
// (dmdssq.ccog:60-101), merged the way binned_dbdbadd does
// (src/binned/dmdmadd.c:27-74), converted by binned_ddmconv.
//
// Layout of the work. The vector is cut into wave tiles of 64*U terms. Each
// wave walks tiles grid-stride. Every lane keeps its own fold-F primaries and
// carries in VGPRs; the index (and, for nrm2, the scale) is common to the
// wave. Per tile a lane takes the max |term| of its U terms; only when some
// lane's max is no longer admissible at the wave's index does the wave take
// the slow path (wave max, raise the index). Then every lane deposits its U
// terms (the reference's deposit, element by element) and the wave
// renormalises every 2048 deposits. At the end the 64 lanes merge by a
// butterfly of exact same-index adds, the waves of a workgroup merge in LDS,
// every workgroup writes one partial, and the last workgroup to finish (a
// ticket counter) merges all partials, folds in the caller's binned value and
// converts. HBM is read exactly once.
//
// Why this equals the reference bit for bit: the slice each bin receives from
// a term is independent of order and of the accumulator, the index ends at
// dindex(max |term|) either way, and the merges are exact.
//
// Non-finite terms poison a lane's top primary (the "|1" of an infinity is a
// NaN). A wave that ends poisoned re-reads its own tiles and replaces its top
// primary by the class of its terms (NaN if any NaN or both infinities, else
// the infinity); merges then combine classes by plain addition, as the
// reference's infinity path does.
#include <float.h>
#include <limits.h>

#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rb {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

struct Blas1Args {
  int64_t n;
  const double* x;
  int64_t incx;
  const double* y;
  int64_t incy;
  double* partials;    // gridDim.x records of (2*fold+1) doubles
  unsigned* ticket;
  const double* state_in;
  double scale_in;
  const double* scale_in_dev;  // when set, the scale of state_in is read here
  double* state_out;
  double* scale_out;
  double* result;
  int fold;
};

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double uniform_double(double v) {
  const uint64_t u = bits(v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return from_bits(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, shfl_xor_d(v, m));
  return v;
}

// The term as the reference's infinity path adds it (dmdsum.ccog:68-72,
// dmdasum and dmdssq use fabs, dmddot the product).
template <int T>
__device__ __forceinline__ double plain_term(double x, double y) {
  if (T == kSum) return x;
  if (T == kDot) return x * y;
  return fabs(x);
}

typedef double d2 __attribute__((ext_vector_type(2)));

// A strided vector range walked in wave tiles of 64*U terms.
struct Span {
  const double* x;
  int64_t incx;
  const double* y;  // second operand (dot) or null
  int64_t incy;
  int64_t n;
  int64_t tiles;       // tiles covering n
  int64_t full_tiles;  // tiles with no out-of-range element
};

template <int U>
__device__ __forceinline__ Span make_span(const double* x, int64_t incx, const double* y, int64_t incy, int64_t n) {
  Span sp;
  sp.x = x;
  sp.incx = incx;
  sp.y = y;
  sp.incy = incy;
  sp.n = n;
  sp.tiles = (n + 64 * U - 1) / (64 * U);
  sp.full_tiles = n / (64 * U);
  return sp;
}

// Loads the U raw terms of tile t for this lane. MODE 0: unit stride with
// 16-byte vector loads (operands 16-byte aligned); MODE 1: unit stride, 8-byte
// loads; MODE 2: arbitrary stride. Out-of-range terms read as 0, which every
// deposit leaves unchanged.
template <int T, int U, int MODE>
__device__ __forceinline__ void load_tile(const Span& a, int64_t t, int lane, double (&vx)[U], double (&vy)[U]) {
  const int64_t base = t * (64 * U);
  const bool full = t < a.full_tiles;
  if (MODE == 0 && full) {
    // streaming 16-byte loads (non-temporal: every byte is read once)
    const d2* xv = reinterpret_cast<const d2*>(a.x + base);
#pragma unroll
    for (int k = 0; k < U / 2; ++k) {
      const d2 d = __builtin_nontemporal_load(xv + k * 64 + lane);
      vx[2 * k] = d.x;
      vx[2 * k + 1] = d.y;
    }
    if (T == kDot) {
      const d2* yv = reinterpret_cast<const d2*>(a.y + base);
#pragma unroll
      for (int k = 0; k < U / 2; ++k) {
        const d2 d = __builtin_nontemporal_load(yv + k * 64 + lane);
        vy[2 * k] = d.x;
        vy[2 * k + 1] = d.y;
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int64_t e = base + k * 64 + lane;
    const bool in = full || e < a.n;
    const int64_t ix = (MODE == 2) ? e * a.incx : e;
    vx[k] = in ? a.x[ix] : 0.0;
    if (T == kDot) {
      const int64_t iy = (MODE == 2) ? e * a.incy : e;
      vy[k] = in ? a.y[iy] : 0.0;
    }
  }
}

template <int CAP>
struct State {
  double p[CAP];
  double c[CAP];
  double scl;  // nrm2 only
};

template <int CAP>
__device__ __forceinline__ void store_state(double* dst, int fold, const State<CAP>& s) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      dst[j] = s.p[j];
      dst[fold + j] = s.c[j];
    }
  dst[2 * fold] = s.scl;
}

template <int CAP>
__device__ __forceinline__ void load_state(const double* src, int fold, State<CAP>& s) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      s.p[j] = src[j];
      s.c[j] = src[fold + j];
    }
  s.scl = src[2 * fold];
}

// General merge of two renormalised values (with scales for nrm2).
template <int CAP, int T>
__device__ __forceinline__ void merge(int fold, State<CAP>& y, const State<CAP>& x) {
  if (T == kSSq) {
    if (nan_or_inf(x.p[0]) || nan_or_inf(y.p[0])) {
      // classes combine by addition; the scale no longer matters
      if (!nan_or_inf(y.p[0])) {
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j) { y.p[j] = x.p[j]; y.c[j] = x.c[j]; }
        y.scl = x.scl;
      } else if (nan_or_inf(x.p[0])) {
        y.p[0] += x.p[0];
      }
      return;
    }
    y.scl = add_scaled(fold, x.scl, x.p, x.c, y.scl, y.p, y.c);
  } else {
    add(fold, y.p, y.c, x.p, x.c);
  }
}

// Merge the 64 lanes of a wave (general add: lanes may differ in index).
template <int CAP, int T>
__device__ __forceinline__ void wave_merge_general(int fold, State<CAP>& s) {
#pragma unroll
  for (int m = 1; m <= 32; m <<= 1) {
    State<CAP> o;
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j) {
      o.p[j] = shfl_xor_d(s.p[j], m);
      o.c[j] = shfl_xor_d(s.c[j], m);
    }
    o.scl = shfl_xor_d(s.scl, m);
    // lower lane of each pair keeps (y = lower, x = upper) so both lanes agree
    merge<CAP, T>(fold, s, o);
  }
}

// Class of the terms of the tiles this wave walked: NaN if any term is NaN or
// both infinities occur, else the infinity that occurs. Terms as the
// reference's infinity path forms them.
template <int T, int U, int MODE>
__device__ double wave_term_class(const Span& a, int64_t first, int64_t step, int lane) {
  bool nan = false, pinf = false, ninf = false;
  for (int64_t t = first; t < a.tiles; t += step) {
    double vx[U], vy[U];
    load_tile<T, U, MODE>(a, t, lane, vx, vy);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const double v = plain_term<T>(vx[k], T == kDot ? vy[k] : 0.0);
      nan |= (v != v);
      pinf |= (v == __builtin_inf());
      ninf |= (v == -__builtin_inf());
    }
  }
  const bool any_nan = __any(nan), any_p = __any(pinf), any_n = __any(ninf);
  if (any_nan || (any_p && any_n) || !(any_p || any_n)) return __builtin_nan("");
  return any_p ? __builtin_inf() : -__builtin_inf();
}

// One wave reduces the tiles t0, t0 + tstep, ... of a span into `s` (every
// lane ends with the wave's value; a poisoned value is replaced by the class
// of the terms).
template <int CAP, int T, int U, int MODE>
__device__ __forceinline__ void wave_accumulate(const Span& a, int fold, int64_t t0, int64_t tstep, int lane,
                                                State<CAP>& s) {
  // Every wave starts at index 51, the reference's representation of an
  // empty sum (dmdupdate of 0.0 on a zero value), so every tile deposits.
  int idx = kMaxIndex;
  set_bins(fold, idx, s.p, s.c);
  double lim = admit_limit(idx);
  double scl = 0.0, inv = 0.0;  // nrm2: current scale of the wave and 1/scale
  int dep = 0;

  // the next tile's loads are issued before the current tile is deposited
  double cx[U], cy[U];
  if (t0 < a.tiles) load_tile<T, U, MODE>(a, t0, lane, cx, cy);
  for (int64_t t = t0; t < a.tiles; t += tstep) {
    double nx[U], ny[U];
    const int64_t tn = t + tstep;
    if (tn < a.tiles) load_tile<T, U, MODE>(a, tn, lane, nx, ny);
    double v[U];
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (T == kSum) v[k] = cx[k];
      if (T == kASum) v[k] = fabs(cx[k]);
      if (T == kDot) v[k] = cx[k] * cy[k];
      if (T == kSSq) v[k] = cx[k];
      m = fmax(m, fabs(v[k]));
    }
    bool need;
    if (T == kSSq) {
      const double ms = m * inv;
      need = (dscale(m) > scl) || (ms * ms >= lim);
    } else {
      need = (m >= lim);
    }
    // the endurance renorm comes before the update as in the reference drivers
    if (dep + U > kEndurance) {
      renorm(fold, s.p, s.c);
      dep = 0;
    }
    if (__any(need)) {
      double M = uniform_double(wave_max(m));
      if (T == kSSq) {
        const double ns = dscale(M);
        if (ns > scl) {
          if (scl > 0.0) {
            rescale(fold, ns, scl, s.p, s.c);
            idx = uniform_int(dmindex(s.p[0]));
          }
          scl = ns;
          inv = 1.0 / scl;
        }
        M = M / scl;  // reference dmdssq.ccog:98-99
        M = M * M;
      }
      const int nI = dindex(M);
      if (nI < idx) {
        raise_index(fold, idx - nI, nI, s.p, s.c);
        idx = nI;
      }
      lim = admit_limit(idx);
    }
    if (idx == 0) {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        double tv = v[k];
        if (T == kSSq) { tv = v[k] * inv; tv = tv * tv; }
        deposit(fold, true, tv, s.p);
      }
    } else {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        double tv = v[k];
        if (T == kSSq) { tv = v[k] * inv; tv = tv * tv; }
        deposit(fold, false, tv, s.p);
      }
    }
    dep += U;
#pragma unroll
    for (int k = 0; k < U; ++k) {
      cx[k] = nx[k];
      if (T == kDot) cy[k] = ny[k];
    }
  }
  renorm(fold, s.p, s.c);
  s.scl = scl;

  // Lanes share the index (and scale): exact same-index butterfly.
#pragma unroll
  for (int m = 1; m <= 32; m <<= 1) {
    double op[CAP], oc[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j) {
      op[j] = shfl_xor_d(s.p[j], m);
      oc[j] = shfl_xor_d(s.c[j], m);
    }
    if (!nan_or_inf(s.p[0]) && !nan_or_inf(op[0])) {
      add_same_index(fold, idx, s.p, s.c, op, oc);
    } else {
      s.p[0] = s.p[0] + op[0];  // keep the poison visible
    }
  }
  if (nan_or_inf(uniform_double(s.p[0]))) {
    s.p[0] = wave_term_class<T, U, MODE>(a, t0, tstep, lane);
  }
}

template <int CAP, bool FIXED, int T, int U, int MODE>
__global__ __launch_bounds__(kThreads) void blas1_kernel(Blas1Args a) {
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t wave_gid = (int64_t)blockIdx.x * kWaves + wave;
  const int64_t nwaves = (int64_t)gridDim.x * kWaves;
  const Span sp = make_span<U>(a.x, a.incx, a.y, a.incy, a.n);
  State<CAP> s;
  wave_accumulate<CAP, T, U, MODE>(sp, fold, wave_gid, nwaves, lane, s);

  // Workgroup merge through LDS, then one partial per workgroup.
  __shared__ double lds[kWaves][2 * CAP + 1];
  __shared__ bool last;
  if (lane == 0) store_state(lds[wave], fold, s);
  __syncthreads();
  if (threadIdx.x == 0) {
    State<CAP> acc, o;
    load_state(lds[0], fold, acc);
    for (int w = 1; w < kWaves; ++w) {
      load_state(lds[w], fold, o);
      merge<CAP, T>(fold, acc, o);
    }
    store_state(a.partials + (int64_t)blockIdx.x * (2 * fold + 1), fold, acc);
    __threadfence();
    const unsigned ticket = atomicAdd(a.ticket, 1u);
    last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;

  // Last workgroup: merge every partial (exact, so the order is free; it is
  // fixed anyway), fold in the caller's value, convert.
  __threadfence();
  State<CAP> acc;
  bool have = false;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads) {
    State<CAP> o;
    const double* src = a.partials + (int64_t)b * (2 * fold + 1);
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) {
        o.p[j] = __builtin_nontemporal_load(src + j);
        o.c[j] = __builtin_nontemporal_load(src + fold + j);
      }
    o.scl = __builtin_nontemporal_load(src + 2 * fold);
    if (!have) {
      acc = o;
      have = true;
    } else {
      merge<CAP, T>(fold, acc, o);
    }
  }
  if (!have) {
    set_bins(fold, kMaxIndex, acc.p, acc.c);
    acc.scl = 0.0;
  }
  wave_merge_general<CAP, T>(fold, acc);
  if (lane == 0) store_state(lds[wave], fold, acc);
  __syncthreads();
  if (threadIdx.x == 0) {
    State<CAP> o;
    load_state(lds[0], fold, acc);
    for (int w = 1; w < kWaves; ++w) {
      load_state(lds[w], fold, o);
      merge<CAP, T>(fold, acc, o);
    }
    // caller's value, renormalised first (the reference ends with a renorm)
    State<CAP> y;
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
    if (a.n == 0) {
      // nothing deposited: the caller's value only (the reference's final renorm)
    } else if (y.p[0] == 0.0) {
      y = acc;
    } else {
      merge<CAP, T>(fold, y, acc);
    }
    renorm(fold, y.p, y.c);
    double scale = y.scl;
    if (T == kSSq && nan_or_inf(y.p[0])) scale = 1.0;  // dscale(1.0), dmdssq.ccog tail
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
      const double v = to_double(fold, y.p, y.c);
      *a.result = (T == kSSq) ? scale * sqrt(v) : v;
    }
    *a.ticket = 0u;  // ready for the next launch on this stream
  }
}

template <int CAP, bool FIXED, int T, int U>
bool launch_mode(int mode, const Blas1Args& a, int blocks, hipStream_t s) {
  switch (mode) {
    case 0: blas1_kernel<CAP, FIXED, T, U, 0><<<blocks, kThreads, 0, s>>>(a); break;
    case 1: blas1_kernel<CAP, FIXED, T, U, 1><<<blocks, kThreads, 0, s>>>(a); break;
    default: blas1_kernel<CAP, FIXED, T, U, 2><<<blocks, kThreads, 0, s>>>(a); break;
  }
  return RB_HIP(hipGetLastError());
}

// Terms per lane per tile: 16 for the one-operand reductions, 8 for dot, so
// every tile moves 8 KiB per wave.
constexpr int kUOne = 16;
constexpr int kUDot = 8;

template <int CAP, bool FIXED>
bool launch_term(Term term, int mode, const Blas1Args& a, int blocks, hipStream_t s) {
  switch (term) {
    case kSum: return launch_mode<CAP, FIXED, kSum, kUOne>(mode, a, blocks, s);
    case kASum: return launch_mode<CAP, FIXED, kASum, kUOne>(mode, a, blocks, s);
    case kDot: return launch_mode<CAP, FIXED, kDot, kUDot>(mode, a, blocks, s);
    default: return launch_mode<CAP, FIXED, kSSq, kUOne>(mode, a, blocks, s);
  }
}

// ------------------------------------------------------- segment batches --
// A ragged batch of sums/dots over CSR segments of one vector pair. The host
// cuts every segment into chunks of at most `chunk` terms; each wave reduces
// whole chunks (grid-stride over the chunk table) into one binned partial per
// chunk; a second kernel merges each segment's chunk partials in order and
// converts. Chunk boundaries do not change the bits.
struct SegArgs {
  const double* x;
  const double* y;
  const SegChunk* chunks;
  int64_t nchunks;
  double* partials;  // nchunks records of (2*fold+1) doubles
  const int64_t* first_chunk;  // segment s owns chunks [first_chunk[s], first_chunk[s+1])
  int64_t nseg;
  double* results;
  int fold;
};

template <int CAP, bool FIXED, int T, int U>
__global__ __launch_bounds__(kThreads) void seg_chunk_kernel(SegArgs a) {
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int64_t wave_gid = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * kWaves;
  for (int64_t c = wave_gid; c < a.nchunks; c += nwaves) {
    const SegChunk ch = a.chunks[c];
    const double* xb = a.x + ch.begin;
    const double* yb = (T == kDot) ? a.y + ch.begin : nullptr;
    const Span sp = make_span<U>(xb, 1, yb, 1, ch.end - ch.begin);
    const bool aligned = ((uintptr_t)xb % 16 == 0) && (T != kDot || (uintptr_t)yb % 16 == 0);
    State<CAP> st;
    if (aligned)
      wave_accumulate<CAP, T, U, 0>(sp, fold, 0, 1, lane, st);
    else
      wave_accumulate<CAP, T, U, 1>(sp, fold, 0, 1, lane, st);
    if (lane == 0) store_state(a.partials + c * (2 * fold + 1), fold, st);
  }
}

template <int CAP, bool FIXED, int T>
__global__ __launch_bounds__(kThreads) void seg_finish_kernel(SegArgs a) {
  const int fold = FIXED ? CAP : a.fold;
  const int64_t sg = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sg >= a.nseg) return;
  const int64_t c0 = a.first_chunk[sg], c1 = a.first_chunk[sg + 1];
  if (c0 == c1) {
    a.results[sg] = 0.0;  // empty segment: the reference's N = 0 result
    return;
  }
  State<CAP> acc;
  load_state(a.partials + c0 * (2 * fold + 1), fold, acc);
  for (int64_t c = c0 + 1; c < c1; ++c) {
    State<CAP> o;
    load_state(a.partials + c * (2 * fold + 1), fold, o);
    merge<CAP, T>(fold, acc, o);
  }
  renorm(fold, acc.p, acc.c);
  const double v = to_double(fold, acc.p, acc.c);
  if (T == kSSq) {
    const double scale = nan_or_inf(acc.p[0]) ? 1.0 : acc.scl;
    a.results[sg] = scale * sqrt(v);
  } else {
    a.results[sg] = v;
  }
}

template <int CAP, bool FIXED, int T, int U>
bool launch_seg(const SegArgs& a, int grid, hipStream_t s) {
  if (a.nchunks > 0) seg_chunk_kernel<CAP, FIXED, T, U><<<grid, kThreads, 0, s>>>(a);
  if (a.nseg > 0) seg_finish_kernel<CAP, FIXED, T><<<(unsigned)((a.nseg + 255) / 256), 256, 0, s>>>(a);
  return RB_HIP(hipGetLastError());
}

template <int CAP, bool FIXED>
bool launch_seg_term(Term term, const SegArgs& a, int grid, hipStream_t s) {
  switch (term) {
    case kSum: return launch_seg<CAP, FIXED, kSum, kUOne>(a, grid, s);
    case kASum: return launch_seg<CAP, FIXED, kASum, kUOne>(a, grid, s);
    case kDot: return launch_seg<CAP, FIXED, kDot, kUDot>(a, grid, s);
    default: return launch_seg<CAP, FIXED, kSSq, kUOne>(a, grid, s);
  }
}

}  // namespace

bool fold_supported(int fold) { return fold >= 2 && fold <= kMaxFold; }

bool blas1(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
           const double* state_in, double scale_in, const double* scale_in_dev, double* state_out, double* scale_out,
           double* result, hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  const DeviceInfo* dev = device_info();
  if (!dev) return false;
  Blas1Args a;
  a.n = n < 0 ? 0 : n;
  a.x = x;
  a.incx = incx;
  a.y = y;
  a.incy = incy;
  const int64_t per_tile = 64 * (term == kDot ? kUDot : kUOne);
  const int64_t tiles = (a.n + per_tile - 1) / per_tile;
  a.state_in = state_in;
  a.scale_in = scale_in;
  a.scale_in_dev = scale_in_dev;
  a.state_out = state_out;
  a.scale_out = scale_out;
  a.result = result;
  a.fold = fold;

  // grid: enough waves to keep every CU streaming, at most one wave per tile
  int64_t blocks = (tiles + kWaves - 1) / kWaves;
  const int64_t cap = (int64_t)dev->cus * 8;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  a.partials = (double*)workspace(s, kSlotPartials, sizeof(double) * (size_t)blocks * (2 * fold + 1));
  a.ticket = tickets(s, 1);
  if (!a.partials || !a.ticket) return false;

  int mode;
  const bool unit = (incx == 1) && (term != kDot || incy == 1);
  if (!unit) {
    mode = 2;
  } else {
    const bool aligned = ((uintptr_t)x % 16 == 0) && (term != kDot || (uintptr_t)y % 16 == 0);
    mode = aligned ? 0 : 1;
  }
  switch (fold) {
    case 2: return launch_term<2, true>(term, mode, a, (int)blocks, s);
    case 3: return launch_term<3, true>(term, mode, a, (int)blocks, s);
    case 4: return launch_term<4, true>(term, mode, a, (int)blocks, s);
    default:
      if (fold <= 8) return launch_term<8, false>(term, mode, a, (int)blocks, s);
      return launch_term<kMaxFold, false>(term, mode, a, (int)blocks, s);
  }
}

bool segments(int fold, Term term, const double* x, const double* y, const SegChunk* chunks, int64_t nchunks,
              double* partials, const int64_t* first_chunk, int64_t nseg, double* results, int grid, hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  SegArgs a;
  a.x = x;
  a.y = y;
  a.chunks = chunks;
  a.nchunks = nchunks;
  a.partials = partials;
  a.first_chunk = first_chunk;
  a.nseg = nseg;
  a.results = results;
  a.fold = fold;
  if (grid < 1) grid = 1;
  switch (fold) {
    case 2: return launch_seg_term<2, true>(term, a, grid, s);
    case 3: return launch_seg_term<3, true>(term, a, grid, s);
    case 4: return launch_seg_term<4, true>(term, a, grid, s);
    default:
      if (fold <= 8) return launch_seg_term<8, false>(term, a, grid, s);
      return launch_seg_term<kMaxFold, false>(term, a, grid, s);
  }
}

}  // namespace rb
