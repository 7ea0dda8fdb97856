// blas1_impl.hpp -- single-pass binned reductions over device vectors:
// sum, asum, dot and the scaled sum of squares behind nrm2.
//
// Reference path: binnedBLAS_dmdsum (src/binnedBLAS/dmdsum.ccog:57-89),
// dmdasum (dmdasum.ccog:57-85), dmddot (dmddot.ccog:59-88) and dmdssq
// (dmdssq.ccog:60-101), merged the way binned_dbdbadd does
// (src/binned/dmdmadd.c:27-74), converted by binned_ddmconv.
//
// Layout of the work. The vector is cut into wave tiles of 64*U terms that
// the waves walk grid-stride, each reducing its tiles into one binned value
// (wave_sum.hpp). The waves of a workgroup merge in LDS, every workgroup
// writes one partial, and the last workgroup to finish (a ticket counter)
// merges all partials, folds in the caller's binned value and converts. HBM
// is read exactly once.
//
// Why this equals the reference bit for bit: the slice each bin receives from
// a term is independent of order and of the accumulator, the index ends at
// dindex(max |term|) either way, and the merges are exact.
#pragma once
#include <float.h>
#include <limits.h>

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
struct Blas1Args {
  int64_t n;
  const F* x;
  int64_t incx;
  const F* y;
  int64_t incy;
  F* partials;    // gridDim.x records of (2*fold+1) elements
  unsigned* ticket;
  const F* state_in;
  F scale_in;
  const F* scale_in_dev;  // when set, the scale of state_in is read here
  F* state_out;
  F* scale_out;
  F* result;
  int fold;
};

template <int CAP, bool FIXED, int T, int U, int MODE, class F>
__global__ __launch_bounds__(kThreads) void blas1_kernel(Blas1Args<F> a) {
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (scalar)
  const int64_t wave_gid = (int64_t)blockIdx.x * kWaves + wave;
  const int64_t nwaves = (int64_t)gridDim.x * kWaves;
  const Span sp = make_span<U>(a.x, a.incx, a.y, a.incy, a.n);
  State<CAP, F> s;
  wave_accumulate<CAP, T, U, MODE>(sp, fold, wave_gid, nwaves, lane, s);

  // Workgroup merge through LDS, then one partial per workgroup.
  __shared__ F lds[kWaves][2 * CAP + 1];
  __shared__ bool last;
  if (lane == 0) store_state(lds[wave], fold, s);
  __syncthreads();
  if (threadIdx.x == 0) {
    State<CAP, F> acc, o;
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
  State<CAP, F> acc;
  bool have = false;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads) {
    State<CAP, F> o;
    const F* src = a.partials + (int64_t)b * (2 * fold + 1);
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
    set_bins(fold, Fmt<F>::max_index, acc.p, acc.c);
    acc.scl = 0.0;
  }
  wave_merge_general<CAP, T>(fold, acc);
  if (lane == 0) store_state(lds[wave], fold, acc);
  __syncthreads();
  if (threadIdx.x == 0) {
    State<CAP, F> o;
    load_state(lds[0], fold, acc);
    for (int w = 1; w < kWaves; ++w) {
      load_state(lds[w], fold, o);
      merge<CAP, T>(fold, acc, o);
    }
    // caller's value, renormalised first (the reference ends with a renorm)
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
    if (a.n == 0) {
      // nothing deposited: the caller's value only (the reference's final renorm)
    } else if (y.p[0] == 0.0) {
      y = acc;
    } else {
      merge<CAP, T>(fold, y, acc);
    }
    renorm(fold, y.p, y.c);
    F scale = y.scl;
    if (T == kSSq && nan_or_inf(y.p[0])) scale = F(1);  // dscale(1.0), dmdssq.ccog tail
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
      const F v = convert(fold, y.p, y.c);
      // nrm2: scl * sqrt(conv) in double, as rdnrm2.c:31 and rsnrm2.c compute it
      *a.result = (T == kSSq) ? (F)((double)scale * sqrt((double)v)) : v;
    }
    *a.ticket = 0u;  // ready for the next launch on this stream
  }
}

template <int CAP, bool FIXED, int T, int U, class F>
bool launch_mode(int mode, const Blas1Args<F>& a, int blocks, hipStream_t s) {
  switch (mode) {
    case 0: blas1_kernel<CAP, FIXED, T, U, 0, F><<<blocks, kThreads, 0, s>>>(a); break;
    case 1: blas1_kernel<CAP, FIXED, T, U, 1, F><<<blocks, kThreads, 0, s>>>(a); break;
    default: blas1_kernel<CAP, FIXED, T, U, 2, F><<<blocks, kThreads, 0, s>>>(a); break;
  }
  return RB_HIP(hipGetLastError());
}

// Terms per lane per tile (double): 16 for the one-operand reductions, 8 for
// dot, so every tile moves 8 KiB per wave; float tiles hold twice as many.
//
// Grid: at most RB_BLAS1_BLOCKS_PER_CU workgroups (4 waves) per CU for unit stride. On
// MI355X a 2^28-term sum streams at ~6.7 TB/s with one workgroup per CU and
// drops to ~5.3 TB/s with eight: every wave keeps its next tile in flight, and
// fewer concurrent streams keep a narrower window of HBM pages open.
// (Compile-time overridable for tuning builds.)
#ifndef RB_BLAS1_U_ONE
#define RB_BLAS1_U_ONE 16
#endif
#ifndef RB_BLAS1_U_DOT
#define RB_BLAS1_U_DOT 8
#endif
#ifndef RB_BLAS1_BLOCKS_PER_CU
#define RB_BLAS1_BLOCKS_PER_CU 1
#endif
#ifndef RB_BLAS1_BLOCKS_PER_CU_STRIDED
#define RB_BLAS1_BLOCKS_PER_CU_STRIDED 4
#endif
// Small vectors: every wave takes at least this many terms. Each workgroup
// costs ~25 ns of serialised completion work (partial store, fence, ticket)
// and each extra tile per wave ~0.7 us of latency; 2048 terms per wave
// measured best at N = 2^18..2^20 (dsum 2^20: 16.4 -> 13.7 us).
#ifndef RB_BLAS1_MIN_TERMS_PER_WAVE
#define RB_BLAS1_MIN_TERMS_PER_WAVE 2048
#endif
// Single precision: terms per lane per tile and workgroups per CU of the
// float kernels (defaults: the same 8 KiB wave tiles as double).
#ifndef RB_BLAS1_U_ONE_S
#define RB_BLAS1_U_ONE_S 32
#endif
#ifndef RB_BLAS1_U_DOT_S
#define RB_BLAS1_U_DOT_S 16
#endif
#ifndef RB_BLAS1_BLOCKS_PER_CU_S
#define RB_BLAS1_BLOCKS_PER_CU_S 1
#endif
constexpr int kMinTermsPerWave = RB_BLAS1_MIN_TERMS_PER_WAVE;
constexpr int kUOne = RB_BLAS1_U_ONE;
constexpr int kUDot = RB_BLAS1_U_DOT;
constexpr int kBlocksPerCUStrided = RB_BLAS1_BLOCKS_PER_CU_STRIDED;

// Terms per lane per tile and unit-stride workgroups per CU for element type F.
template <class F>
constexpr int u_one() { return sizeof(F) == sizeof(double) ? kUOne : RB_BLAS1_U_ONE_S; }
template <class F>
constexpr int u_dot() { return sizeof(F) == sizeof(double) ? kUDot : RB_BLAS1_U_DOT_S; }
template <class F>
constexpr int blocks_per_cu() { return sizeof(F) == sizeof(double) ? RB_BLAS1_BLOCKS_PER_CU : RB_BLAS1_BLOCKS_PER_CU_S; }

template <int CAP, bool FIXED, class F>
bool launch_term(Term term, int mode, const Blas1Args<F>& a, int blocks, hipStream_t s) {
  switch (term) {
    case kSum: return launch_mode<CAP, FIXED, kSum, u_one<F>()>(mode, a, blocks, s);
    case kASum: return launch_mode<CAP, FIXED, kASum, u_one<F>()>(mode, a, blocks, s);
    case kDot: return launch_mode<CAP, FIXED, kDot, u_dot<F>()>(mode, a, blocks, s);
    default: return launch_mode<CAP, FIXED, kSSq, u_one<F>()>(mode, a, blocks, s);
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
  const int64_t wave_gid = (int64_t)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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

namespace {
template <class F>
bool blas1_impl(int fold, Term term, int64_t n, const F* x, int64_t incx, const F* y, int64_t incy,
                const F* state_in, F scale_in, const F* scale_in_dev, F* state_out, F* scale_out, F* result,
                hipStream_t s) {
  if (fold < 2 || fold > Fmt<F>::max_fold) {
    set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
    return false;
  }
  const DeviceInfo* dev = device_info();
  if (!dev) return false;
  Blas1Args<F> a;
  a.n = n < 0 ? 0 : n;
  a.x = x;
  a.incx = incx;
  a.y = y;
  a.incy = incy;
  const int64_t per_tile = 64 * (term == kDot ? u_dot<F>() : u_one<F>());
  const int64_t tiles = (a.n + per_tile - 1) / per_tile;
  a.state_in = state_in;
  a.scale_in = scale_in;
  a.scale_in_dev = scale_in_dev;
  a.state_out = state_out;
  a.scale_out = scale_out;
  a.result = result;
  a.fold = fold;

  int mode;
  const bool unit = (incx == 1) && (term != kDot || incy == 1);
  if (!unit) {
    mode = 2;
  } else {
    const bool aligned = ((uintptr_t)x % 16 == 0) && (term != kDot || (uintptr_t)y % 16 == 0);
    mode = aligned ? 0 : 1;
  }

  // grid: a few streaming waves per CU, at most one wave per tile
  const int64_t min_tiles = std::max<int64_t>(1, kMinTermsPerWave / per_tile);
  int64_t blocks = (tiles + kWaves * min_tiles - 1) / (kWaves * min_tiles);
  const int64_t cap = (int64_t)dev->cus * (mode == 2 ? kBlocksPerCUStrided : blocks_per_cu<F>());
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  a.partials = (F*)workspace(s, kSlotPartials, sizeof(F) * (size_t)blocks * (2 * fold + 1));
  a.ticket = tickets(s, 1);
  if (!a.partials || !a.ticket) return false;
  switch (fold) {
    case 2: return launch_term<2, true>(term, mode, a, (int)blocks, s);
    case 3: return launch_term<3, true>(term, mode, a, (int)blocks, s);
    case 4: return launch_term<4, true>(term, mode, a, (int)blocks, s);
    default:
      if (fold <= 8) return launch_term<8, false>(term, mode, a, (int)blocks, s);
      return launch_term<Fmt<F>::max_fold, false>(term, mode, a, (int)blocks, s);
  }
}
}  // namespace

}  // namespace rb
