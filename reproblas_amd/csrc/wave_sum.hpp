// wave_sum.hpp -- one wave's binned reduction of a strided vector range of
// doubles or floats, shared by the BLAS1 kernels, the ragged segment kernels
// and the row-dot gemv kernel.
//
// Reference path: the per-block loops of binnedBLAS_dmdsum
// (src/binnedBLAS/dmdsum.ccog:57-89), dmdasum (dmdasum.ccog:57-85), dmddot
// (dmddot.ccog:59-88) and dmdssq (dmdssq.ccog:60-101), and the merge of
// binned_dbdbadd (src/binned/dmdmadd.c:27-74).
//
// A wave walks tiles of 64*U terms (t0, t0 + tstep, ...). Every lane keeps its
// own fold-F primaries and carries in VGPRs; the index (and, for nrm2, the
// scale) is common to the wave. Per tile a lane takes the max |term| of its U
// terms; only when some lane's max is no longer admissible at the wave's
// index does the wave take the slow path (wave max, raise the index). Then
// every lane deposits its U terms (the reference's deposit, element by
// element) and the wave renormalises every 2048 deposits. The 64 lanes merge
// by a butterfly of exact same-index adds.
//
// Non-finite terms poison a lane's top primary (the "|1" of an infinity is a
// NaN). A wave that ends poisoned re-reads its own tiles and replaces its top
// primary by the class of its terms (NaN if any NaN or both infinities, else
// the infinity); merges then combine classes by plain addition, as the
// reference's infinity path does.
//
// Included into an anonymous namespace of each kernel source.
#pragma once
#include <limits.h>

#include "binned_core.hpp"
#include "kernels.hpp"

namespace rb {
namespace {

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// Partner value for stage `st` (0..5) of a wave reduction tree. Stage st pairs
// every lane with one lane of the other half of its aligned group of
// 2^(st+1) lanes; when every lane combines its value with its partner's by an
// exact, commutative operation, after stage st all lanes of such a group hold
// the same value, and after stage 5 every lane holds the whole wave's. Stages
// 0-3 are DPP lane moves inside rows of 16 (quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror: these pair lane i with the other
// half of its group because the groups below already agree), stage 4 a
// ds_swizzle xor 16 inside 32 lanes, stage 5 reads lanes 0 and 32.
__device__ __forceinline__ int lane_partner32(int v, int st) {
  switch (st) {
    case 0: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true);
    case 1: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true);
    case 2: return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true);
    case 3: return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, true);
    case 4: return __builtin_amdgcn_ds_swizzle(v, 0x401F);
    default: {
      const int lo_half = __builtin_amdgcn_readlane(v, 0);
      const int hi_half = __builtin_amdgcn_readlane(v, 32);
      return (__lane_id() < 32) ? hi_half : lo_half;
    }
  }
}

__device__ __forceinline__ double lane_partner(double v, int st) {
  const uint64_t u = bits(v);
  const uint32_t lo = (uint32_t)lane_partner32((int)(uint32_t)u, st);
  const uint32_t hi = (uint32_t)lane_partner32((int)(uint32_t)(u >> 32), st);
  return from_bits(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float lane_partner(float v, int st) {
  return fp_from_bits<float>((uint32_t)lane_partner32((int)bits(v), st));
}

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double uniform_value(double v) {
  const uint64_t u = bits(v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return from_bits(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float uniform_value(float v) {
  return fp_from_bits<float>((uint32_t)__builtin_amdgcn_readfirstlane((int)bits(v)));
}
__device__ __forceinline__ double uniform_double(double v) { return uniform_value(v); }

// The value of lane k (k wave-uniform).
__device__ __forceinline__ double lane_value(double v, int k) {
  const uint64_t u = bits(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, k);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), k);
  return from_bits(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float lane_value(float v, int k) {
  return fp_from_bits<float>((uint32_t)__builtin_amdgcn_readlane((int)bits(v), k));
}

template <class F>
__device__ __forceinline__ F wave_max(F v) {
#pragma unroll
  for (int st = 0; st < 6; ++st) v = fmax(v, lane_partner(v, st));
  return v;
}

// The term as the reference's infinity path adds it (dmdsum.ccog:68-72,
// dmdasum and dmdssq use fabs, dmddot the product).
template <int K, class F>
__device__ __forceinline__ F plain_term(F x, F y) {
  if (K == kSum) return x;
  if (K == kDot) return x * y;
  return fabs(x);
}

typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// A strided vector range of element type F walked in wave tiles of 64*U terms.
template <class F>
struct Span {
  const F* x;
  int64_t incx;
  const F* y;  // second operand (dot) or null
  int64_t incy;
  int64_t n;
  int64_t tiles;       // tiles covering n
  int64_t full_tiles;  // tiles with no out-of-range element
};

template <int U, class F>
__device__ __forceinline__ Span<F> make_span(const F* x, int64_t incx, const F* y, int64_t incy, int64_t n) {
  Span<F> sp;
  sp.x = x;
  sp.incx = incx;
  sp.y = y;
  sp.incy = incy;
  sp.n = n;
  sp.tiles = (n + 64 * U - 1) / (64 * U);
  sp.full_tiles = n / (64 * U);
  return sp;
}

// U consecutive 16-byte vectors of a lane (64 lanes interleaved per vector
// row), streamed with non-temporal loads: every byte is read once.
template <int U, class F>
__device__ __forceinline__ void load_rows16(const F* base, int lane, F (&v)[U]) {
  if constexpr (sizeof(F) == 8) {
    const d2* p = reinterpret_cast<const d2*>(base);
#pragma unroll
    for (int k = 0; k < U / 2; ++k) {
      const d2 d = __builtin_nontemporal_load(p + k * 64 + lane);
      v[2 * k] = d.x;
      v[2 * k + 1] = d.y;
    }
  } else {
    const f4* p = reinterpret_cast<const f4*>(base);
#pragma unroll
    for (int k = 0; k < U / 4; ++k) {
      const f4 d = __builtin_nontemporal_load(p + k * 64 + lane);
      v[4 * k] = d.x;
      v[4 * k + 1] = d.y;
      v[4 * k + 2] = d.z;
      v[4 * k + 3] = d.w;
    }
  }
}

// Element index (within the span) of term k of this lane in tile t, in the
// layout load_tile uses: MODE 0 full tiles hold 16-byte vectors of V
// consecutive elements, 64 lanes interleaved per vector row; every other
// tile holds element k*64 + lane.
template <int U, int MODE, class F>
__device__ __forceinline__ int64_t tile_element(const Span<F>& a, int64_t t, int k, int lane) {
  constexpr int V = 16 / (int)sizeof(F);
  const int64_t base = t * (64 * U);
  if (MODE == 0 && t < a.full_tiles) return base + (int64_t)V * ((k / V) * 64 + lane) + (k % V);
  return base + (int64_t)k * 64 + lane;
}

// Loads the U raw terms of tile t for this lane. MODE 0: unit stride with
// 16-byte vector loads (operands 16-byte aligned); MODE 1: unit stride,
// element loads; MODE 2: arbitrary stride. Out-of-range terms read as 0,
// which every deposit leaves unchanged.
template <int K, int U, int MODE, class F>
__device__ __forceinline__ void load_tile(const Span<F>& a, int64_t t, int lane, F (&vx)[U], F (&vy)[U]) {
  const int64_t base = t * (64 * U);
  const bool full = t < a.full_tiles;
  if (MODE == 0 && full) {
    load_rows16<U>(a.x + base, lane, vx);
    if (K == kDot) load_rows16<U>(a.y + base, lane, vy);
    return;
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int64_t e = base + k * 64 + lane;
    const bool in = full || e < a.n;
    const int64_t ix = (MODE == 2) ? e * a.incx : e;
    vx[k] = in ? a.x[ix] : F(0);
    if (K == kDot) {
      const int64_t iy = (MODE == 2) ? e * a.incy : e;
      vy[k] = in ? a.y[iy] : F(0);
    }
  }
}

template <int CAP, class F = double>
struct State {
  F p[CAP];
  F c[CAP];
  F scl;  // nrm2 only
};

template <int CAP, class F>
__device__ __forceinline__ void store_state(F* dst, int fold, const State<CAP, F>& s) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      dst[j] = s.p[j];
      dst[fold + j] = s.c[j];
    }
  dst[2 * fold] = s.scl;
}

template <int CAP, class F>
__device__ __forceinline__ void load_state(const F* src, int fold, State<CAP, F>& s) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      s.p[j] = src[j];
      s.c[j] = src[fold + j];
    }
  s.scl = src[2 * fold];
}

// General merge of two renormalised values (with scales for nrm2).
template <int CAP, int K, class F>
__device__ __forceinline__ void merge(int fold, State<CAP, F>& y, const State<CAP, F>& x) {
  if (K == kSSq) {
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

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
  for (int st = 0; st < 6; ++st) {
    const int o = lane_partner32(v, st);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ int wave_or_int(int v) {
#pragma unroll
  for (int st = 0; st < 6; ++st) v |= lane_partner32(v, st);
  return v;
}

// Minimum and OR over each aligned group of 2^STAGES lanes.
template <int STAGES>
__device__ __forceinline__ int group_min_int(int v) {
#pragma unroll
  for (int st = 0; st < STAGES; ++st) {
    const int o = lane_partner32(v, st);
    v = o < v ? o : v;
  }
  return v;
}
template <int STAGES>
__device__ __forceinline__ int group_or_int(int v) {
#pragma unroll
  for (int st = 0; st < STAGES; ++st) v |= lane_partner32(v, st);
  return v;
}

__device__ __forceinline__ int64_t lane_partner_i64(int64_t v, int st) {
  const uint64_t u = (uint64_t)v;
  const uint32_t lo = (uint32_t)lane_partner32((int)(uint32_t)u, st);
  const uint32_t hi = (uint32_t)lane_partner32((int)(uint32_t)(u >> 32), st);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// ---- exact merges in integer form ----
// A value's slots at index I as integers: the offset of each primary from
// its bin, in units of that bin's ulp (exact: between renormalisations a
// primary stays in its bin's binade, so the offset is a difference of bit
// patterns), next to the carries. Renormalisation keeps the low
// `mbits - 2` bits of the offset in the primary and moves the rest, as a
// floor division, into the carry (src/binned/dmrenorm.c:70-85). Values at
// one index therefore merge exactly by summing offsets and carries, and the
// renormalised sum is a function of those sums alone: equal to any order of
// exact adds (src/binned/dmdmadd.c:27-74) followed by a renorm. A value at a
// higher index joins a lower index I by moving its slots down (slot j of the
// raised value is its slot j - shift, fresh top slots are offset 0), which
// is what the index raise of dmdupdate.c:42-55 does.
template <int CAP, class F>
__device__ __forceinline__ void to_offsets(int fold, int I, const F (&p)[CAP], int64_t (&d)[CAP]) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) d[j] = (int64_t)bits(p[j]) - (int64_t)bits(bin<F>(I + j));
}

// The renormalised value at index I whose slot offsets sum to d; the
// carries c already hold the carry sums and receive the overflow of d.
template <int CAP, class F>
__device__ __forceinline__ void from_offsets(int fold, int I, const int64_t (&d)[CAP], F (&p)[CAP], F (&c)[CAP]) {
  using U = typename Fmt<F>::U;
  constexpr int lo = Lay<F>::mbits - 2;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      p[j] = fp_from_bits<F>((U)(bits(bin<F>(I + j)) + (U)(d[j] & (((int64_t)1 << lo) - 1))));
      c[j] += (F)(d[j] >> lo);
    }
}

// Class of a non-finite value as bits (1 NaN, 2 +inf, 4 -inf), and back: the
// class that adding the values would give (NaN for any NaN or both
// infinities, else the one infinity).
template <class F>
__device__ __forceinline__ int class_bits(F v) {
  return (v != v) ? 1 : (v > F(0) ? 2 : 4);
}
template <class F>
__device__ __forceinline__ F class_value(int b) {
  if ((b & 1) || (b & 6) == 6) return (F)__builtin_nan("");
  return (b & 2) ? (F)__builtin_inf() : -(F)__builtin_inf();
}

// Exact merge of renormalised values held by the lanes of a wave (lanes with
// `valid` false hold nothing); afterwards every lane of each aligned group of
// 2^STAGES lanes holds its group's sum (STAGES = 6: the whole wave's). The
// values may differ in index: each joins the group's lowest index -- the
// alignment the general add (dmdmadd.c:27-74) makes pairwise -- and the
// group sums its slot offsets and carries (integer form above). Non-finite
// values combine by class, as adding their top primaries would. Not for
// scaled (nrm2) values.
template <int STAGES, int CAP, class F>
__device__ __forceinline__ void wave_merge_aligned(int fold, bool valid, State<CAP, F>& s) {
  const bool fin = valid && !nan_or_inf(s.p[0]);
  const int mine = fin ? dmindex(s.p[0]) : Fmt<F>::max_index;
  int I = group_min_int<STAGES>(fin ? mine : INT_MAX);
  const int cls = group_or_int<STAGES>((valid && !fin) ? class_bits(s.p[0]) : 0);
  if (I == INT_MAX) I = Fmt<F>::max_index;
  int64_t own[CAP], d[CAP];
  F c[CAP];
  to_offsets(fold, mine, s.p, own);
  const int sh = fin ? mine - I : CAP;  // slot j takes own slot j - sh
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) {
      d[j] = (j >= sh) ? pick(own, j - sh) : 0;
      c[j] = (j >= sh) ? pick(s.c, j - sh) : F(0);
    }
#pragma unroll
  for (int st = 0; st < STAGES; ++st) {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) {
        d[j] += lane_partner_i64(d[j], st);
        c[j] += lane_partner(c[j], st);
      }
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) s.c[j] = c[j];
  from_offsets(fold, I, d, s.p, s.c);
  if (cls) s.p[0] = class_value<F>(cls);
}

// Merge the 64 lanes of a wave (general add: lanes may differ in index).
template <int CAP, int K, class F>
__device__ __forceinline__ void wave_merge_general(int fold, State<CAP, F>& s) {
#pragma unroll
  for (int st = 0; st < 6; ++st) {
    State<CAP, F> o;
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j) {
      o.p[j] = lane_partner(s.p[j], st);
      o.c[j] = lane_partner(s.c[j], st);
    }
    o.scl = lane_partner(s.scl, st);
    merge<CAP, K>(fold, s, o);
  }
}

// What a wave reports about infinities besides its binned value, for the
// replay of the reference's infinity path (sum and dot only):
//   first_inf -- element index of the wave's first infinite term (INT64_MAX
//                when it has none),
//   big       -- the wave saw a finite term of index 0 (|t| >= 2^984 for
//                double), the only terms whose sequential sum can overflow.
struct PoisonInfo {
  int64_t first_inf;
  bool big;
};

// Minimum over the wave (every lane ends with it), by the lane moves of the
// reduction trees below on the two halves.
__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
#pragma unroll
  for (int st = 0; st < 6; ++st) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)lane_partner32((int)(uint32_t)u, st);
    const uint32_t hi = (uint32_t)lane_partner32((int)(uint32_t)(u >> 32), st);
    const int64_t o = (int64_t)(((uint64_t)hi << 32) | lo);
    v = o < v ? o : v;
  }
  return v;
}

// Class of the terms of the tiles this wave walked: NaN if any term is NaN or
// both infinities occur, else the infinity that occurs. Terms as the
// reference's infinity path forms them. With `info`, also the wave's first
// infinite term and whether a finite term of index 0 occurs.
template <int K, int U, int MODE, class F>
__device__ F wave_term_class(const Span<F>& a, int64_t first, int64_t step, int lane, PoisonInfo* info) {
  bool nan = false, pinf = false, ninf = false, big = false;
  int64_t fi = INT64_MAX;
  for (int64_t t = first; t < a.tiles; t += step) {
    F vx[U], vy[U];
    load_tile<K, U, MODE>(a, t, lane, vx, vy);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const F v = plain_term<K>(vx[k], K == kDot ? vy[k] : F(0));
      const bool inf = fabs(v) == (F)__builtin_inf();
      nan |= (v != v);
      pinf |= (v == (F)__builtin_inf());
      ninf |= (v == -(F)__builtin_inf());
      if (info) {
        const int64_t e = tile_element<U, MODE>(a, t, k, lane);
        if (inf && e < fi) fi = e;
        big |= !inf && fabs(v) >= admit_limit<F>(1);
      }
    }
  }
  if (info) {
    info->first_inf = wave_min_i64(fi);
    info->big = __any(big);
  }
  const bool any_nan = __any(nan), any_p = __any(pinf), any_n = __any(ninf);
  if (any_nan || (any_p && any_n) || !(any_p || any_n)) return (F)__builtin_nan("");
  return any_p ? (F)__builtin_inf() : -(F)__builtin_inf();
}

// Per-wave running state of wave_accumulate.
template <class F>
struct WaveRun {
  int idx;  // common index of the wave's bins
  F lim;    // admit_limit(idx)
  F scl;    // nrm2: current scale of the wave
  F inv;    // nrm2: 1 / scl
  F mn;     // nrm2: smallest nonzero |x| the lane has seen (+inf: none), when tracked
  bool track;
  int dep;  // deposits since the last renormalisation
};

// Deposits one tile of raw terms (cx, cy) into the lane's bins. In tiles of
// 8 or more terms the even terms go into the lane's primaries and the odd
// ones into a second set that starts the tile at the same bins, so the two
// dependency chains interleave; the second set's offsets then join the first
// (exact: together they hold exactly the slices one set would, within the
// endurance).
template <int CAP, int K, int U, class F>
__device__ __forceinline__ void deposit_tile(int fold, const F (&cx)[U], const F (&cy)[U], WaveRun<F>& w,
                                             State<CAP, F>& s) {
  F v[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    if (K == kSum) v[k] = cx[k];
    if (K == kASum) v[k] = fabs(cx[k]);
    if (K == kDot) v[k] = cx[k] * cy[k];
    if (K == kSSq) v[k] = cx[k];
  }
  // max |term| by a balanced tree (NaN ignored, as fmax does)
  F t[U];
#pragma unroll
  for (int k = 0; k < U; ++k) t[k] = fabs(v[k]);
#pragma unroll
  for (int n = U; n > 1; n = (n + 1) / 2)
#pragma unroll
    for (int k = 0; k + (n + 1) / 2 < n; ++k) t[k] = fmax(t[k], t[k + (n + 1) / 2]);
  const F m = fmax(F(0), t[0]);
  bool need;
  if (K == kSSq) {
    const F ms = m * w.inv;
    need = (dscale(m) > w.scl) || (ms * ms >= w.lim);
  } else {
    need = (m >= w.lim);
  }
  if (K == kSSq && w.track) {
    F mk = (F)__builtin_inf();
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const F ak = fabs(v[k]);  // the raw term (squared further down)
      mk = (ak > F(0) && ak < mk) ? ak : mk;
    }
    w.mn = mk < w.mn ? mk : w.mn;
  }
  // the endurance renorm comes before the update as in the reference drivers
  if (w.dep + U > Fmt<F>::endurance) {
    renorm(fold, s.p, s.c);
    w.dep = 0;
  }
  if (__any(need)) {
    F M = uniform_value(wave_max(m));
    if (K == kSSq) {
      const F ns = dscale(M);
      if (ns > w.scl) {
        if (w.scl > F(0)) {
          rescale(fold, ns, w.scl, s.p, s.c);
          w.idx = uniform_int(dmindex(s.p[0]));
        }
        w.scl = ns;
        w.inv = F(1) / w.scl;  // exact: scales are powers of two
      }
      M = M / w.scl;  // reference dmdssq.ccog:98-99
      M = M * M;
    }
    const int nI = dindex(M);
    if (nI < w.idx) {
      raise_index(fold, w.idx - nI, nI, s.p, s.c);
      w.idx = nI;
    }
    w.lim = admit_limit<F>(w.idx);
  }
  if (K == kSSq) {
#pragma unroll
    for (int k = 0; k < U; ++k) {
      v[k] = v[k] * w.inv;
      v[k] = v[k] * v[k];
    }
  }
  const bool top = (w.idx == 0);
  if constexpr (U >= 8) {
    F q[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) q[j] = bin<F>(w.idx + j);
    if (top) {
#pragma unroll
      for (int k = 0; k + 1 < U; k += 2) deposit2(fold, true, v[k], s.p, v[k + 1], q);
      if (U % 2) deposit(fold, true, v[U - 1], s.p);
    } else {
#pragma unroll
      for (int k = 0; k + 1 < U; k += 2) deposit2(fold, false, v[k], s.p, v[k + 1], q);
      if (U % 2) deposit(fold, false, v[U - 1], s.p);
    }
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) s.p[j] = s.p[j] + (q[j] - bin<F>(w.idx + j));
  } else {
    // short tiles: one chain (the second set costs more than it overlaps)
    if (top) {
#pragma unroll
      for (int k = 0; k < U; ++k) deposit(fold, true, v[k], s.p);
    } else {
#pragma unroll
      for (int k = 0; k < U; ++k) deposit(fold, false, v[k], s.p);
    }
  }
  w.dep += U;
}

// Fresh wave state: every wave starts at the last index, the reference's
// representation of an empty sum (dmdupdate of 0.0 on a zero value), so every
// tile deposits. nrm2 starts at the caller's scale (0 for none): the
// reference squares every term at max(scaleY, dscale(amax so far))
// (dmdssq.ccog:88-99), so no term is ever squared at a smaller scale.
template <int CAP, class F>
__device__ __forceinline__ void wave_start(int fold, WaveRun<F>& w, State<CAP, F>& s, F scale0 = F(0),
                                           bool track_min = false) {
  w.idx = Fmt<F>::max_index;
  set_bins(fold, w.idx, s.p, s.c);
  w.lim = admit_limit<F>(w.idx);
  w.scl = scale0;
  w.inv = scale0 > F(0) ? F(1) / scale0 : F(0);  // exact: scales are powers of two
  w.mn = (F)__builtin_inf();
  w.track = track_min;
  w.dep = 0;
}

// Exact same-index merge of the 64 lanes (they share the wave's index and
// scale) in integer form: every lane ends with the wave's renormalised value.
// A poisoned lane poisons the wave (the caller then replaces the top primary
// by the class of the terms).
template <int CAP, class F>
__device__ __forceinline__ void wave_butterfly(int fold, const WaveRun<F>& w, State<CAP, F>& s) {
  s.scl = w.scl;
  if (__any(nan_or_inf(s.p[0]))) {
    s.p[0] = (F)__builtin_nan("");
    return;
  }
  int64_t d[CAP];
  to_offsets(fold, w.idx, s.p, d);
#pragma unroll
  for (int st = 0; st < 6; ++st) {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) {
        d[j] += lane_partner_i64(d[j], st);
        s.c[j] += lane_partner(s.c[j], st);
      }
  }
  from_offsets(fold, w.idx, d, s.p, s.c);
}

template <int CAP, int K, int U, int MODE, class F>
__device__ __forceinline__ void wave_finish(const Span<F>& a, int fold, int64_t t0, int64_t tstep, int lane,
                                            const WaveRun<F>& w, State<CAP, F>& s, PoisonInfo* info) {
  wave_butterfly(fold, w, s);
  if (nan_or_inf(uniform_value(s.p[0]))) {
    s.p[0] = wave_term_class<K, U, MODE>(a, t0, tstep, lane, info);
  } else if (info) {
    info->first_inf = INT64_MAX;
    info->big = (w.idx == 0);
  }
}

// One wave reduces the tiles t0, t0 + tstep, ... of a span into `s` (every
// lane ends with the wave's value; a poisoned value is replaced by the class
// of the terms). `info` (or null) receives what the infinity-path replay needs;
// `min_nz` (or null; nrm2) the smallest nonzero |x| of the wave when
// track_min is set (+inf otherwise).
template <int CAP, int K, int U, int MODE, class F>
__device__ __forceinline__ void wave_accumulate(const Span<F>& a, int fold, int64_t t0, int64_t tstep, int lane,
                                                State<CAP, F>& s, PoisonInfo* info = nullptr, F scale0 = F(0),
                                                F* min_nz = nullptr, bool track_min = false) {
  WaveRun<F> w;
  wave_start(fold, w, s, scale0, min_nz != nullptr && track_min);

  // Two tile buffers in turn: the loads of the next tile are in flight while
  // the current one is deposited, and no buffer is ever copied.
  F ax[U], ay[U], bx[U], by[U];
  int64_t t = t0;
  if (t < a.tiles) load_tile<K, U, MODE>(a, t, lane, ax, ay);
  while (t < a.tiles) {
    int64_t tn = t + tstep;
    if (tn < a.tiles) load_tile<K, U, MODE>(a, tn, lane, bx, by);
    deposit_tile<CAP, K, U>(fold, ax, ay, w, s);
    if (tn >= a.tiles) break;
    t = tn;
    tn = t + tstep;
    if (tn < a.tiles) load_tile<K, U, MODE>(a, tn, lane, ax, ay);
    deposit_tile<CAP, K, U>(fold, bx, by, w, s);
    t = tn;
  }
  wave_finish<CAP, K, U, MODE>(a, fold, t0, tstep, lane, w, s, info);
  if (min_nz) {
    // smallest nonzero |x| of the wave (nrm2), every lane
    F m = w.mn;
#pragma unroll
    for (int st = 0; st < 6; ++st) {
      const F o = lane_partner(m, st);
      m = o < m ? o : m;
    }
    *min_nz = m;
  }
}

}  // namespace
}  // namespace rb
