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

// Cross-workgroup accumulators of the sum, asum and dot kernels, in the
// zeroed block that also holds the completion ticket (at byte kGridAccOffset;
// every launch leaves it zero again). Each workgroup adds its value's slot
// offsets and carries (integer form, wave_sum.hpp) at their absolute bin
// indexes -- index + slot, the same bin for every value -- so the last
// workgroup reads the sums aligned to the lowest index directly: a higher
// value's slots below that index's window are the ones its raise would
// drop. Integer atomics are associative, so the arrival order is free.
// Workgroup b adds into copy b % kCopies: atomics on one address serialise
// (~20 ns each), so the copies keep a 128-workgroup grid at 8 per address.
struct GridAcc {
  static constexpr int kCopies = 16;
  static constexpr int kBins = 104;  // >= max_index + max_fold (double: 103; float: 41)
  int64_t d[kCopies][kBins];         // slot offsets by absolute bin
  int64_t c[kCopies][kBins];         // carries by absolute bin
  int top[kCopies];                  // max of max_index - index (finite values)
  int flags;                         // class bits of poisoned values | 8: one saw a finite term of index 0
  int pad;
  int64_t fi;                        // max of INT64_MAX - first infinite term (poisoned values)
};
constexpr size_t kReplayFlagOffset = 128;  // nrm2: int, set when the replay kernel must run
constexpr size_t kGridAccOffset = 256;
constexpr size_t kSyncBytes = kGridAccOffset + sizeof(GridAcc);  // ticket block of one launch
constexpr int kGridAccWords = (int)(sizeof(GridAcc) / sizeof(int64_t));
static_assert(kSyncBytes <= kBlas1SyncBytes, "the caller-provided sync block must hold the grid accumulators");

template <class F>
struct Blas1Args {
  int64_t n;
  const F* x;
  int64_t incx;
  const F* y;
  int64_t incy;
  F* partials;    // nrm2: gridDim.x records of (2*fold+1) elements
  F* mins;        // nrm2: gridDim.x smallest nonzero |x| of each workgroup
  int replay;     // nrm2: ssq_replay_kernel follows this launch (see replay_ssq)
  unsigned* ticket;  // zeroed block of kSyncBytes: the ticket, then a GridAcc
  const F* state_in;
  F scale_in;
  const F* scale_in_dev;  // when set, the scale of state_in is read here
  F* state_out;
  F* scale_out;
  F* result;
  int fold;
};

// One launch of the BLAS1 kernel of a term on prepared arguments (mode: 0
// unit stride 16-byte aligned, 1 unit stride, 2 strided). Each is compiled
// in a unit of its own (blas1_inst.hip built once per term and type), so the
// kernel instances compile in parallel.
bool blas1_launch_sum(int mode, const Blas1Args<double>& a, int blocks, hipStream_t s);
bool blas1_launch_asum(int mode, const Blas1Args<double>& a, int blocks, hipStream_t s);
bool blas1_launch_dot(int mode, const Blas1Args<double>& a, int blocks, hipStream_t s);
bool blas1_launch_ssq(int mode, const Blas1Args<double>& a, int blocks, hipStream_t s);
bool blas1_launch_sum(int mode, const Blas1Args<float>& a, int blocks, hipStream_t s);
bool blas1_launch_asum(int mode, const Blas1Args<float>& a, int blocks, hipStream_t s);
bool blas1_launch_dot(int mode, const Blas1Args<float>& a, int blocks, hipStream_t s);
bool blas1_launch_ssq(int mode, const Blas1Args<float>& a, int blocks, hipStream_t s);

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Slot sums of the kWaves renormalised values vals[0..kWaves) (records of
// 2*fold+1 in LDS): the finite values join their lowest index I (integer
// form, wave_sum.hpp) and lane j of the calling wave sums the offsets (d) and
// carries (c) of slot j over them; cls gathers the class bits of the
// non-finite values. Every lane of the wave calls it; I and cls are uniform.
// Not for scaled (nrm2) values.
template <int CAP, class F>
__device__ __forceinline__ void slot_sums(int fold, F (*vals)[2 * CAP + 1], int lane, int& I, int& cls,
                                          int64_t& d, F& c) {
  I = INT_MAX;
  cls = 0;
  int idx[kWaves];
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const F p0 = vals[w][0];
    idx[w] = nan_or_inf(p0) ? INT_MAX : dmindex(p0);
    if (nan_or_inf(p0)) cls |= class_bits(p0);
    I = idx[w] < I ? idx[w] : I;
  }
  if (I == INT_MAX) I = Fmt<F>::max_index;
  const int j = lane < fold ? lane : 0;
  const int64_t b = (int64_t)bits(bin<F>(I + j));
  d = 0;
  c = F(0);
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const int src = j - (idx[w] - I);  // the value's own slot that lands on slot j
    if (idx[w] != INT_MAX && src >= 0) {
      d += (int64_t)bits(vals[w][src]) - b;
      c += vals[w][fold + src];
    }
  }
}

__device__ __forceinline__ int64_t lane_value_i64(int64_t v, int k) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(uint64_t)v, k);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)v >> 32), k);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Block length of the reference's driver for a term kind: binnedBLAS_dmdsum
// and dmddot walk blocks of 2048 and 1024 doubles (dmdsum.ccog:36,
// dmddot.ccog:36), smssum and smsdot blocks of 512 floats (smssum.ccog:37,
// smsdot.ccog:36). Each renormalises before a block once `endurance`
// deposits would be exceeded (dmdsum.ccog:81-84): before every block after
// the first when the block is the endurance, before every second one when
// it is half of it.
template <int T, class F>
constexpr int ref_block() {
  return sizeof(F) == sizeof(double) ? (T == kDot ? 1024 : 2048) : 512;
}

// The binned sum of the terms [lo, hi) by the 4 waves of one workgroup;
// thread 0 ends with the merged value (renormalised). Every thread calls it.
// nrm2: every term is squared at scale0, which must be at least dscale of
// every |x| in the range (no wave changes scale).
template <int CAP, int T, int U, int MODE, class F>
__device__ __forceinline__ void wg_range_sum(const Blas1Args<F>& a, int fold, int64_t lo, int64_t hi, int wave, int lane,
                             F (*lds)[2 * CAP + 1], State<CAP, F>& out, F scale0 = F(0)) {
  const Span<F> sp = make_span<U>(a.x + lo * a.incx, a.incx, T == kDot ? a.y + lo * a.incy : nullptr, a.incy,
                                  hi - lo);
  State<CAP, F> st;
  wave_accumulate<CAP, T, U, MODE>(sp, fold, wave, kWaves, lane, st, nullptr, scale0);
  __syncthreads();
  if (lane == 0) store_state(lds[wave], fold, st);
  __syncthreads();
  if (threadIdx.x == 0) {
    State<CAP, F> o;
    load_state(lds[0], fold, out);
    for (int w = 1; w < kWaves; ++w) {
      load_state(lds[w], fold, o);
      add(fold, out.p, out.c, o.p, o.c);
    }
  }
}

// The reference's infinity path (dmdsum.ccog:66-78, dmddot.ccog:68-81):
// at the first block holding an infinite term it adds that block's terms
// one by one, in order, to its top primary priY[0] -- the state of all
// earlier blocks, renormalised at the last renorm the cadence made and not
// since -- and afterwards only combines classes. The result is NaN when that
// sequential sum overflows to the infinity opposite to the block's first
// infinite term before reaching it, else that infinity. Called by every
// thread of the last workgroup when the merged class is one infinity and a
// finite term (or the caller's value) of index 0 occurs; pstar is the first
// infinite term. Thread 0 returns the reference's priY[0] after that block
// (which is then the result: later blocks add only same-sign infinities or
// finite terms to it).
template <int CAP, int T, int U, int MODE, class F>
__device__ __forceinline__ F replay_infinity_path(const Blas1Args<F>& a, int fold, int64_t pstar, int wave, int lane,
                                  F (*lds)[2 * CAP + 1], F* seq) {
  constexpr int64_t NB = ref_block<T, F>();
  constexpr int64_t kPer = Fmt<F>::endurance / NB;  // blocks per renorm
  const int64_t bs = pstar / NB;                    // the first block with an infinite term
  F p0 = F(0);
  if (bs > 0) {
    // blocks [r, bs) were deposited since the last renorm (before block r)
    const int64_t r = ((bs - 1) / kPer) * kPer;
    State<CAP, F> R, S;
    if (r >= 1) wg_range_sum<CAP, T, U, MODE>(a, fold, 0, r * NB, wave, lane, lds, R);
    wg_range_sum<CAP, T, U, MODE>(a, fold, r * NB, bs * NB, wave, lane, lds, S);
    if (threadIdx.x == 0) {
      if (r >= 1) {
        if (a.state_in) {
          State<CAP, F> y;
RB_UNROLL_CAP
          for (int j = 0; j < CAP; ++j)
            if (j < fold) { y.p[j] = a.state_in[j]; y.c[j] = a.state_in[fold + j]; }
          renorm(fold, y.p, y.c);
          add(fold, R.p, R.c, y.p, y.c);
        }
      } else if (a.state_in) {
        // no renorm yet: the caller's value as given (dmdsum.ccog:81-84 skips block 0)
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) { R.p[j] = a.state_in[j]; R.c[j] = a.state_in[fold + j]; }
      } else {
        set_zero(fold, R.p, R.c);
      }
      // the updates of blocks [r, bs) leave the index at min(I_R, I_S)
      // (dmdupdate.c:24-56); the top slot then holds R's content plus the
      // top-bin slices of those blocks, exact (dmddeposit.c:35-70 within the
      // endurance)
      const int is = dmindex(S.p[0]);
      int ip = is;
      if (R.p[0] == F(0)) {
        set_bins(fold, ip, R.p, R.c);
      } else {
        const int ir = dmindex(R.p[0]);
        ip = ir < is ? ir : is;
        raise_index(fold, ir - ip, ip, R.p, R.c);
      }
      raise_index(fold, is - ip, ip, S.p, S.c);
      p0 = (R.p[0] + (S.p[0] - bin<F>(ip))) + S.c[0] * carry_unit<F>(ip);
    }
  } else if (threadIdx.x == 0 && a.state_in) {
    p0 = a.state_in[0];
  }
  // the terms of block bs up to the first infinity, staged in order
  const int64_t b0 = bs * NB;
  for (int64_t e = b0 + threadIdx.x; e <= pstar; e += kThreads) {
    const F xv = a.x[e * a.incx];
    seq[e - b0] = plain_term<T>(xv, T == kDot ? a.y[e * a.incy] : F(0));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int64_t j = 0; j <= pstar - b0; ++j) p0 = p0 + seq[j];
  }
  return p0;
}

// The reference's sum of squares in one call (dmdssq.ccog:60-101): each
// block of ref_block() terms squares at the running scale of the block
// maxima so far (from the caller's scaleY), and the accumulator is rescaled
// (dmdrescale) whenever that scale grows. Its value is therefore the ordered
// merge of "epochs" -- maximal runs of blocks with one scale -- each summed
// exactly at its scale, with the accumulator rescaled at every epoch start.
// The result equals the parallel one unless a scaled square is subnormal and
// the value's lowest bin is fine enough to hold part of it (bin 51 for
// double, 19-20 for float); the last workgroup then replays this schedule.
// Called by every thread of one workgroup (ssq_replay_kernel, its own launch
// so the streaming kernel keeps its registers); thread 0 ends with the value
// (y) and its scale (y.scl).
template <int CAP, int U, int MODE, class F>
__device__ __forceinline__ void replay_ssq(const Blas1Args<F>& a, int fold, int wave, int lane, F (*lds)[2 * CAP + 1],
                                           State<CAP, F>& y) {
  constexpr int64_t NB = ref_block<kSSq, F>();
  constexpr int kMaxEpochs = 96;  // scales are distinct powers 2^(WIDTH*k+c): far fewer
  const int64_t nb = (a.n + NB - 1) / NB;
  const F scale_in = a.scale_in_dev ? *a.scale_in_dev : a.scale_in;
  __shared__ F blk_scale[kWaves];
  __shared__ int64_t ep_start[kMaxEpochs];
  __shared__ F ep_scale[kMaxEpochs];
  __shared__ int nep;
  __shared__ F run;
  if (threadIdx.x == 0) {
    nep = 0;
    run = scale_in;
  }
  __syncthreads();
  // pass 1: dscale(max |x|) of every block in order, epoch starts where the
  // running scale grows (block 0 always starts one)
  for (int64_t g = 0; g < nb; g += kWaves) {
    const int64_t b = g + wave;
    if (b < nb) {
      const int64_t e1 = (b + 1) * NB < a.n ? (b + 1) * NB : a.n;
      F m = F(0);
      for (int64_t e = b * NB + lane; e < e1; e += 64) m = fmax(m, fabs(a.x[e * a.incx]));
      m = wave_max(m);
      if (lane == 0) blk_scale[wave] = dscale(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 0; w < kWaves && g + w < nb; ++w) {
        const F sc = blk_scale[w];
        if (g + w == 0 || sc > run) {
          run = sc > run ? sc : run;
          if (nep < kMaxEpochs) {
            ep_start[nep] = g + w;
            ep_scale[nep] = run;
          }
          ++nep;
        }
      }
    }
    __syncthreads();
  }
  // pass 2: each epoch summed at its scale, merged in order into the
  // caller's value as given
  if (threadIdx.x == 0) {
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
    y.scl = scale_in;
  }
  const int ne = nep < kMaxEpochs ? nep : kMaxEpochs;
  F scl = scale_in;
  for (int e = 0; e < ne; ++e) {
    const int64_t lo = ep_start[e] * NB;
    const int64_t hi = (e + 1 < ne) ? ep_start[e + 1] * NB : a.n;
    const F sc = ep_scale[e];
    State<CAP, F> P;
    wg_range_sum<CAP, kSSq, U, MODE>(a, fold, lo, hi, wave, lane, lds, P, sc);
    if (threadIdx.x == 0) {
      // the caller's value is rescaled as given; later the accumulator was
      // renormalised before each block, so before each rescale
      if (sc > scl && scl > F(0)) rescale(fold, sc, scl, y.p, y.c);
      scl = sc;
      renorm(fold, y.p, y.c);
      add(fold, y.p, y.c, P.p, P.c);
      y.scl = scl;
    }
  }
}

// The outputs of a reduction from its final (renormalised) value y.
template <int T, int CAP, class F>
__device__ __forceinline__ void write_outputs(const Blas1Args<F>& a, int fold, const State<CAP, F>& y) {
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
}

template <int CAP, bool FIXED, int T, int U, int MODE, class F>
__global__ __launch_bounds__(kThreads) void blas1_kernel(Blas1Args<F> a) {
  constexpr bool kInfPath = (T == kSum || T == kDot);
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (scalar)
  const int64_t wave_gid = (int64_t)blockIdx.x * kWaves + wave;
  const int64_t nwaves = (int64_t)gridDim.x * kWaves;
  const Span<F> sp = make_span<U>(a.x, a.incx, a.y, a.incy, a.n);
  State<CAP, F> s;
  PoisonInfo pi;
  // nrm2: every wave squares at no less than the caller's scale
  const F scale0 = (T == kSSq && a.state_in) ? (a.scale_in_dev ? *a.scale_in_dev : a.scale_in) : F(0);
  F wmin = (F)__builtin_inf();  // nrm2: smallest nonzero |x| of the wave
  wave_accumulate<CAP, T, U, MODE>(sp, fold, wave_gid, nwaves, lane, s, kInfPath ? &pi : nullptr, scale0,
                                   T == kSSq ? &wmin : nullptr, T == kSSq && a.replay);

  // Workgroup merge through LDS, then one partial per workgroup.
  __shared__ F lds[kWaves][2 * CAP + 1];
  __shared__ int64_t lds_fi[kWaves];
  __shared__ int lds_big[kWaves];
  __shared__ F lds_min[kWaves];
  __shared__ bool last;
  if (lane == 0) {
    store_state(lds[wave], fold, s);
    if (T == kSSq) lds_min[wave] = wmin;
    if (kInfPath) {
      lds_fi[wave] = pi.first_inf;
      lds_big[wave] = pi.big;
    }
  }
  __syncthreads();
  if constexpr (T == kSSq) {
    // scaled values: the general scaled merge, one after another
    if (threadIdx.x == 0) {
      State<CAP, F> acc, o;
      load_state(lds[0], fold, acc);
      F mn = lds_min[0];
      for (int w = 1; w < kWaves; ++w) {
        load_state(lds[w], fold, o);
        merge<CAP, T>(fold, acc, o);
        mn = lds_min[w] < mn ? lds_min[w] : mn;
      }
      store_state(a.partials + (int64_t)blockIdx.x * (2 * fold + 1), fold, acc);
      a.mins[blockIdx.x] = mn;
    }
  } else if (wave == 0) {
    // the 4 wave values, slot-parallel in wave 0, then added into the grid
    // accumulators by thread 0 (a poisoned value contributes its class only)
    GridAcc* g = reinterpret_cast<GridAcc*>(reinterpret_cast<char*>(a.ticket) + kGridAccOffset);
    int I, cls;
    int64_t d;
    F c;
    slot_sums<CAP>(fold, lds, lane, I, cls, d, c);
    if (cls == 0) {
      const int cp = blockIdx.x % GridAcc::kCopies;
      for (int k = 0; k < fold; ++k) {
        const int64_t dk = lane_value_i64(d, k);
        const int64_t ck = (int64_t)lane_value(c, k);
        if (lane == 0) {
          __hip_atomic_fetch_add(&g->d[cp][I + k], dk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(&g->c[cp][I + k], ck, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      if (lane == 0)
        __hip_atomic_fetch_max(&g->top[cp], Fmt<F>::max_index - I, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (lane == 0) {
      int flags = cls;
      if (kInfPath) {
        int64_t fi = lds_fi[0];
        int big = lds_big[0];
        for (int w = 1; w < kWaves; ++w) {
          fi = lds_fi[w] < fi ? lds_fi[w] : fi;
          big |= lds_big[w];
        }
        flags |= big ? 8 : 0;
        __hip_atomic_fetch_max(&g->fi, INT64_MAX - fi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __hip_atomic_fetch_or(&g->flags, flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (threadIdx.x == 0) {
    // release the partial (nrm2) or the accumulator adds to the last
    // workgroup; the ticket itself is relaxed (fence-to-fence
    // synchronisation through it)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned ticket = __hip_atomic_fetch_add(a.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;

  // Last workgroup: the merged value of every workgroup (nrm2: merge the
  // partial records; the others: read the grid accumulators), then the
  // caller's value, then convert.
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  State<CAP, F> acc;  // thread 0: the grid's value
  int64_t pstar = -1;
  F gmin = (F)__builtin_inf();  // nrm2, thread 0: smallest nonzero |x| of the grid
  if constexpr (T == kSSq) {
    bool have = false;
    F mn = (F)__builtin_inf();
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads) {
      const F bm = __builtin_nontemporal_load(a.mins + b);
      mn = bm < mn ? bm : mn;
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
#pragma unroll
    for (int st = 0; st < 6; ++st) {
      const F o = lane_partner(mn, st);
      mn = o < mn ? o : mn;
    }
    if (lane == 0) {
      store_state(lds[wave], fold, acc);
      lds_min[wave] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      State<CAP, F> o;
      load_state(lds[0], fold, acc);
      gmin = lds_min[0];
      for (int w = 1; w < kWaves; ++w) {
        load_state(lds[w], fold, o);
        merge<CAP, T>(fold, acc, o);
        gmin = lds_min[w] < gmin ? lds_min[w] : gmin;
      }
    }
  } else {
    GridAcc* g = reinterpret_cast<GridAcc*>(reinterpret_cast<char*>(a.ticket) + kGridAccOffset);
    __shared__ int64_t replay_at;
    if (wave == 0) {
      // lanes 0..15 hold the copies; the sums over them take the four DPP
      // stages that stay inside a row of 16 lanes
      constexpr int kC = GridAcc::kCopies;
      static_assert(kC == 16, "the copy sums below are the four row-internal stages");
      const int cl = lane < kC ? lane : 0;
      const bool own = lane < kC;
      int top = own ? g->top[cl] : 0;
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int o = lane_partner32(top, st);
        top = o > top ? o : top;
      }
      top = __builtin_amdgcn_readlane(top, 0);
      const int I = Fmt<F>::max_index - top;
      int64_t d[CAP], c[CAP];
RB_UNROLL_CAP
      for (int j = 0; j < CAP; ++j)
        if (j < fold) {
          d[j] = own ? g->d[cl][I + j] : 0;
          c[j] = own ? g->c[cl][I + j] : 0;
        }
#pragma unroll
      for (int st = 0; st < 4; ++st) {
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) {
            d[j] += lane_partner_i64(d[j], st);
            c[j] += lane_partner_i64(c[j], st);
          }
      }
      const int flags = g->flags;
      const int cls = flags & 7;
      if (lane == 0) {
RB_UNROLL_CAP
        for (int j = 0; j < CAP; ++j)
          if (j < fold) acc.c[j] = (F)c[j];
        from_offsets(fold, I, d, acc.p, acc.c);
        if (cls) acc.p[0] = class_value<F>(cls);
        acc.scl = F(0);
      }
      // Replay decision: one infinity overall, a finite caller value, and a
      // magnitude that can overflow a sequential sum (a finite term of
      // index 0 somewhere, or in the caller's value).
      if constexpr (kInfPath) {
        const int64_t fim = g->fi;
        const bool one_inf = cls == 2 || cls == 4;
        const int64_t f = fim ? INT64_MAX - fim : INT64_MAX;
        const bool bg = (flags & 8) || (I == 0);
        const F y0 = a.state_in ? a.state_in[0] : F(0);
        // (f < n always holds for an infinite class; checked so no replay
        // can ever index past the vector)
        if (a.n > 0 && one_inf && f >= 0 && f < a.n && !nan_or_inf(y0) && (bg || dmindex0(y0))) pstar = f;
      }
      if (lane == 0) replay_at = pstar;
    }
    // wave 0 has read the accumulators: clear them for the next launch
    __syncthreads();
    pstar = replay_at;
    for (int i = threadIdx.x; i < kGridAccWords; i += kThreads) reinterpret_cast<int64_t*>(g)[i] = 0;
  }
  State<CAP, F> y;
  bool ssq_replay = false;  // thread 0, nrm2: leave the outputs to ssq_replay_kernel
  if (threadIdx.x == 0) {
    // caller's value, renormalised first (the reference ends with a renorm)
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
  }
  bool replayed = false;
  if constexpr (kInfPath) {
    if (pstar >= 0) {
      __shared__ F seq[ref_block<T, F>()];
      // the reference's top primary after its first infinite block: it
      // already holds the caller's value, and later blocks leave it as is
      const F top = replay_infinity_path<CAP, T, U, MODE>(a, fold, pstar, wave, lane, lds, seq);
      if (threadIdx.x == 0) y.p[0] = top;
      replayed = true;
    }
  }
  if (threadIdx.x == 0) {
    if (a.n > 0 && !replayed) {
      if (a.state_in) {
        // An empty caller value still carries its scale for nrm2: the
        // reference squares every term at max(scaleY, dscale(amax))
        // (dmdssq.ccog:62-101), which the scaled merge reproduces by
        // rescaling acc up to scaleY.
        merge<CAP, T>(fold, y, acc);
      } else {
        y = acc;  // what merging into the zero value (scale 0) gives
      }
    }
    // n == 0: nothing deposited, the caller's value only (the reference's final renorm)
    renorm(fold, y.p, y.c);
  }
  if constexpr (T == kSSq) {
    // The reference's one-call schedule where it can differ (replay_ssq):
    // a finite value whose lowest bin is fine enough to see a subnormal
    // square (its ulp below twice the smallest normal: bin 51 for double,
    // bins 19-20 for float) and a term whose square at the final scale is
    // subnormal (|x| / scale < 2^-(MAX_EXP/2 - 1); the reference's earlier
    // block scales are no larger, so its squares are no smaller).
    if (threadIdx.x == 0) {
      const bool fin = !nan_or_inf(y.p[0]) && y.p[0] != F(0);
      const F tiny = fp_from_bits<F>((typename Fmt<F>::U)(Lay<F>::ieee_bias - (Fmt<F>::max_exp / 2 - 1))
                                     << Lay<F>::mbits);
      const bool fine = fin && bin_exp<F>(dmindex(y.p[0]) + fold - 1) < Lay<F>::mbits + 2;
      ssq_replay = a.replay && a.n > 0 && fine && y.scl > F(0) && gmin / y.scl < tiny;
    }
  }
  if (threadIdx.x == 0) {
    // nrm2 in the replay region: ssq_replay_kernel (next on the stream)
    // writes the outputs from the caller's untouched value
    if (!(T == kSSq && ssq_replay)) write_outputs<T>(a, fold, y);
    if (T == kSSq && ssq_replay) *reinterpret_cast<int*>(reinterpret_cast<char*>(a.ticket) + kReplayFlagOffset) = 1;
    *a.ticket = 0u;  // ready for the next launch on this stream
  }
}

// One workgroup after an nrm2 launch whose replay switch is on: replays the
// reference's one-call schedule when that launch flagged the region
// (replay_ssq), writes the outputs and clears the flag; else returns.
template <int CAP, bool FIXED, int U, int MODE, class F>
__global__ __launch_bounds__(kThreads) void ssq_replay_kernel(Blas1Args<F> a) {
  int* flag = reinterpret_cast<int*>(reinterpret_cast<char*>(a.ticket) + kReplayFlagOffset);
  if (*flag == 0) return;
  const int fold = FIXED ? CAP : a.fold;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  __shared__ F lds[kWaves][2 * CAP + 1];
  State<CAP, F> y;
  replay_ssq<CAP, U, MODE>(a, fold, wave, lane, lds, y);
  __syncthreads();  // every thread has read the flag
  if (threadIdx.x == 0) {
    renorm(fold, y.p, y.c);
    write_outputs<kSSq>(a, fold, y);
    *flag = 0;
  }
}

template <int CAP, bool FIXED, int T, int U, class F>
bool launch_mode(int mode, const Blas1Args<F>& a, int blocks, hipStream_t s) {
  switch (mode) {
    case 0: blas1_kernel<CAP, FIXED, T, U, 0, F><<<blocks, kThreads, 0, s>>>(a); break;
    case 1: blas1_kernel<CAP, FIXED, T, U, 1, F><<<blocks, kThreads, 0, s>>>(a); break;
    default: blas1_kernel<CAP, FIXED, T, U, 2, F><<<blocks, kThreads, 0, s>>>(a); break;
  }
  if constexpr (T == kSSq) {
    if (a.replay) {
      switch (mode) {
        case 0: ssq_replay_kernel<CAP, FIXED, U, 0, F><<<1, kThreads, 0, s>>>(a); break;
        case 1: ssq_replay_kernel<CAP, FIXED, U, 1, F><<<1, kThreads, 0, s>>>(a); break;
        default: ssq_replay_kernel<CAP, FIXED, U, 2, F><<<1, kThreads, 0, s>>>(a); break;
      }
    }
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

// The kernel instance of term T for the arguments' fold.
template <int T, class F>
bool launch_fold_term(int mode, const Blas1Args<F>& a, int blocks, hipStream_t s) {
  constexpr int U = (T == kDot) ? u_dot<F>() : u_one<F>();
  switch (a.fold) {
    case 2: return launch_mode<2, true, T, U>(mode, a, blocks, s);
    case 3: return launch_mode<3, true, T, U>(mode, a, blocks, s);
    case 4: return launch_mode<4, true, T, U>(mode, a, blocks, s);
    default:
      if (a.fold <= 8) return launch_mode<8, false, T, U>(mode, a, blocks, s);
      return launch_mode<Fmt<F>::max_fold, false, T, U>(mode, a, blocks, s);
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
// ext_scratch (or null for the library workspace): device memory of
// ext_bytes for the partial records (nrm2), and ext_ticket: a zeroed block of
// kSyncBytes (the ticket and the grid accumulators) that the launch leaves
// zeroed.
template <class F>
bool blas1_impl(int fold, Term term, int64_t n, const F* x, int64_t incx, const F* y, int64_t incy,
                const F* state_in, F scale_in, const F* scale_in_dev, F* state_out, F* scale_out, F* result,
                hipStream_t s, void* ext_scratch = nullptr, size_t ext_bytes = 0, unsigned* ext_ticket = nullptr) {
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
  // nrm2: one partial record and one minimum per workgroup; the others
  // merge through the grid accumulators in the ticket block and need none
  const size_t rec_bytes = sizeof(F) * (size_t)blocks * (2 * fold + 1);
  const size_t need = term == kSSq ? rec_bytes + sizeof(F) * (size_t)blocks : 0;
  char* ws;
  if (ext_scratch) {
    if (need > ext_bytes) {
      set_error(kBadArgument,
                "this reduction needs %zu bytes of scratch (%lld workgroup records at fold %d); the fixed scratch "
                "holds %zu",
                need, (long long)blocks, fold, ext_bytes);
      return false;
    }
    ws = (char*)ext_scratch;
    a.ticket = ext_ticket;
  } else {
    ws = (char*)workspace(s, kSlotPartials, need > 0 ? need : 1);
    a.ticket = tickets(s, (int)((kSyncBytes + sizeof(unsigned) - 1) / sizeof(unsigned)));
  }
  a.partials = (F*)ws;
  a.mins = ws ? (F*)(ws + (term == kSSq ? rec_bytes : 0)) : nullptr;
  // nrm2: the region where the reference's block schedule matters (see
  // replay_ssq) needs the value's lowest bin among the finest ones. Without a
  // caller's value the scaled squares' top index is 23..27 (double) or 7..11
  // (float), so only folds >= 25 (double) or >= 9 (float) reach them; a
  // caller's value of larger scale can put the index anywhere.
  a.replay = term == kSSq && a.n > 0 && (state_in != nullptr || fold >= (sizeof(F) == 8 ? 25 : 9));
  if (!a.partials || !a.ticket) return false;
  switch (term) {
    case kSum: return blas1_launch_sum(mode, a, (int)blocks, s);
    case kASum: return blas1_launch_asum(mode, a, (int)blocks, s);
    case kDot: return blas1_launch_dot(mode, a, (int)blocks, s);
    default: return blas1_launch_ssq(mode, a, (int)blocks, s);
  }
}
}  // namespace

}  // namespace rb
