// binned_core.hpp -- the binned arithmetic (FP64 double_binned and FP32
// float_binned), written once for host and device (gfx950) code.
//
// A binned value of fold F holds F primaries p[0..F) and F carries c[0..F).
// p[j] accumulates slices of every term that fall in bin (I + j), where I is
// the index of the value; bins are 40 bits wide for double and 13 for float
// (reference include/binned.h:84,95). The slice that each bin receives from a
// term is the round-to-nearest of the remaining term at that bin's ulp, with
// ties made impossible by forcing the term's last mantissa bit to one. That
// slice does not depend on the accumulator's contents or on the order of
// terms, so any lane layout, tiling or merge tree gives the same bits,
// provided:
//   * every accumulator is renormalised at least every `endurance` deposits
//     (2048 for double, 512 for float),
//   * merges use the exact add below on renormalised operands,
//   * conversion follows the reference's fixed operation order.
// Every function is a template on the element type T (double or float) whose
// constants come from Fmt<T>; the formulas are the reference's with
// DBL_/FLT_ constants substituted. All code here must be compiled with
// -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RB_HD __host__ __device__ __forceinline__
// Loops over a binned value's slots: fully unrolled (registers) for the
// small capacities, plain loops (private memory) for the general one.
#define RB_UNROLL_CAP _Pragma("unroll (CAP <= 8 ? CAP : 1)")

namespace rb {

template <class T>
struct Fmt;

// reference include/binned.h:84,105,123,142,184,206 (DBWIDTH, binned_DBMAXINDEX,
// binned_DBMAXFOLD, binned_DBENDURANCE, binned_DMCOMPRESSION/EXPANSION)
template <>
struct Fmt<double> {
  using U = uint64_t;
  static constexpr int mant = 53;       // DBL_MANT_DIG
  static constexpr int max_exp = 1024;  // DBL_MAX_EXP
  static constexpr int width = 40;
  static constexpr int max_index = 51;
  static constexpr int max_fold = 52;
  static constexpr int endurance = 2048;
  static constexpr double compression = 1.0 / (1 << 14);
  static constexpr double expansion = 1.0 * (1 << 14);
};

// reference include/binned.h:95,115,131,153,195,217 (SBWIDTH, binned_SBMAXINDEX,
// binned_SBMAXFOLD, binned_SBENDURANCE, binned_SMCOMPRESSION/EXPANSION)
template <>
struct Fmt<float> {
  using U = uint32_t;
  static constexpr int mant = 24;      // FLT_MANT_DIG
  static constexpr int max_exp = 128;  // FLT_MAX_EXP
  static constexpr int width = 13;
  static constexpr int max_index = 20;
  static constexpr int max_fold = 21;
  static constexpr int endurance = 512;
  static constexpr float compression = 1.0f / (1 << 12);
  static constexpr float expansion = 1.0f * (1 << 12);
};

// Derived layout constants (reference src/common/common.h:43-69).
template <class T>
struct Lay {
  static constexpr int mbits = Fmt<T>::mant - 1;                  // stored mantissa bits
  static constexpr int exp_all = 2 * Fmt<T>::max_exp - 1;         // exponent field of NaN/Inf
  static constexpr int exp_bias = Fmt<T>::max_exp - 2;            // the reference's EXP_BIAS
  static constexpr int ieee_bias = Fmt<T>::max_exp - 1;
  static constexpr int top_exp = 2 * Fmt<T>::max_exp - 2;         // exponent of bin 0
  static constexpr int bin_base = 2 * Fmt<T>::max_exp + Fmt<T>::mant - Fmt<T>::width - 1;  // bins 1..
};

// FP64 constants under their historical names (the double-only code).
constexpr int kWidth = Fmt<double>::width;
constexpr int kMaxIndex = Fmt<double>::max_index;
constexpr int kMaxFold = Fmt<double>::max_fold;
constexpr int kEndurance = Fmt<double>::endurance;
constexpr double kCompression = Fmt<double>::compression;
constexpr double kExpansion = Fmt<double>::expansion;

RB_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
RB_HD uint32_t bits(float x) { return __builtin_bit_cast(uint32_t, x); }
RB_HD double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
template <class T>
RB_HD T fp_from_bits(typename Fmt<T>::U u) {
  return __builtin_bit_cast(T, u);
}

// biased exponent field (reference src/common/common.h:43-55)
template <class T>
RB_HD int biased_exp(T x) {
  return (int)((bits(x) >> Lay<T>::mbits) & (typename Fmt<T>::U)Lay<T>::exp_all);
}
// exponent all ones (reference src/common/common.h:59-69)
template <class T>
RB_HD bool nan_or_inf(T x) {
  return biased_exp(x) == Lay<T>::exp_all;
}
// the deposit's "|1": last mantissa bit forced to one (one v_or_b32 on gfx950)
template <class T>
RB_HD T set_lsb(T x) {
  return fp_from_bits<T>(bits(x) | (typename Fmt<T>::U)1);
}

// Bin exponent of index i: bin 0 is 1.5*2^(MAX_EXP-1), bin i >= 1 is
// 0.75*2^(MAX_EXP+MANT-WIDTH+1-i*WIDTH), the last bin repeated beyond
// (reference src/binned/dmbins.c:23-39, smbins.c).
template <class T>
RB_HD int bin_exp(int i) {
  i = i > Fmt<T>::max_index ? Fmt<T>::max_index : i;
  return (i == 0) ? Lay<T>::top_exp : Lay<T>::bin_base - Fmt<T>::width * i;
}
// Bin value of index i, built from bits: no table, no memory traffic.
template <class T = double>
RB_HD T bin(int i) {
  using U = typename Fmt<T>::U;
  return fp_from_bits<T>(((U)bin_exp<T>(i) << Lay<T>::mbits) | ((U)1 << (Lay<T>::mbits - 1)));
}
// bin(i)/6 = 2^(exponent(bin(i)) - 2), the weight of one carry unit; exact,
// so equal to the reference's division (src/binned/ddmconv.c:47-75).
template <class T = double>
RB_HD T carry_unit(int i) {
  using U = typename Fmt<T>::U;
  return fp_from_bits<T>((U)(bin_exp<T>(i) - 2) << Lay<T>::mbits);
}
// Smallest magnitude whose index is below I: |t| < admit_limit(I) <=> dindex(t) >= I.
// 2^(MAX_EXP-WIDTH*I) for I >= 1 and +inf for I = 0 (from src/binned/dindex.c:40).
template <class T = double>
RB_HD T admit_limit(int I) {
  using U = typename Fmt<T>::U;
  return fp_from_bits<T>((U)(Lay<T>::exp_all - Fmt<T>::width * I) << Lay<T>::mbits);
}

// reference src/binned/dindex.c:19-41 and sindex.c. Zero maps to the last
// index; a subnormal through frexp's exponent (its bit length), capped; NaN
// and Inf map to 0 through C's truncating division. (For double every
// subnormal lands on 51; for float the largest ones land on 19.)
template <class T>
RB_HD int dindex(T x) {
  const int e = biased_exp(x);
  if (e == 0) {
    const auto m = bits(x) & (((typename Fmt<T>::U)1 << Lay<T>::mbits) - 1);
    if (m == 0) return Fmt<T>::max_index;
    int bl;  // bit length of the mantissa field
    if constexpr (sizeof(T) == 8) {
      bl = 64 - __builtin_clzll((unsigned long long)m);
    } else {
      bl = 32 - __builtin_clz((unsigned)m);
    }
    const int fe = bl + 1 - Lay<T>::ieee_bias - Lay<T>::mbits;  // frexp exponent
    const int idx = (Fmt<T>::max_exp - fe) / Fmt<T>::width;
    return idx < Fmt<T>::max_index ? idx : Fmt<T>::max_index;
  }
  return (Lay<T>::top_exp - e) / Fmt<T>::width;
}
// reference src/binned/dmindex.c:22-38, smindex.c
template <class T>
RB_HD int dmindex(T p0) {
  return (Lay<T>::bin_base - biased_exp(p0)) / Fmt<T>::width;
}
// reference src/binned/dmindex0.c:19-31, smindex0.c
template <class T>
RB_HD bool dmindex0(T p0) {
  return biased_exp(p0) == Lay<T>::top_exp;
}

// reference src/binned/dscale.c:25-31, sscale.c: a power of two 2^(e-1); e
// keeps C's truncating '%' on a possibly negative operand.
template <class T>
RB_HD T dscale(T x) {
  using U = typename Fmt<T>::U;
  int e = biased_exp(x);
  e = e < Fmt<T>::width ? Fmt<T>::width : e;
  e -= (e - Lay<T>::exp_bias - 1) % Fmt<T>::width;
  e -= Lay<T>::exp_bias;
  return fp_from_bits<T>((U)(e - 1 + Lay<T>::ieee_bias) << Lay<T>::mbits);
}

// Register-array helpers. CAP is the compile-time capacity; `fold` <= CAP is
// the fold in use (a constant after inlining for the fixed-fold kernels).
template <int CAP, class T>
RB_HD T pick(const T (&a)[CAP], int k) {
  if constexpr (CAP > 8) {
    return (k >= 0 && k < CAP) ? a[k] : T(0);
  } else {
    T v = T(0);
#pragma unroll
    for (int t = 0; t < CAP; ++t) v = (t == k) ? a[t] : v;
    return v;
  }
}

template <int CAP, class T>
RB_HD void set_zero(int fold, T (&p)[CAP], T (&c)[CAP]) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = T(0); c[j] = T(0); }
}

// Fresh primaries of index I (reference src/binned/dmdupdate.c:36-41).
template <int CAP, class T>
RB_HD void set_bins(int fold, int I, T (&p)[CAP], T (&c)[CAP]) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = bin<T>(I + j); c[j] = T(0); }
}

// Raise a non-empty value to index I < current: slots move down by
// shift = current - I, the lowest fall off, fresh top bins come in
// (reference src/binned/dmdupdate.c:42-55).
template <int CAP, class T>
RB_HD void raise_index(int fold, int shift, int I, T (&p)[CAP], T (&c)[CAP]) {
  if (shift <= 0) return;
  T np[CAP], nc[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      if (j >= shift) {
        np[j] = pick(p, j - shift);
        nc[j] = pick(c, j - shift);
      } else {
        np[j] = bin<T>(I + j);
        nc[j] = T(0);
      }
    }
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = np[j]; c[j] = nc[j]; }
}

// reference src/binned/dmdupdate.c:24-56 (update for a term of magnitude x)
template <int CAP, class T>
RB_HD void update(int fold, T x, T (&p)[CAP], T (&c)[CAP]) {
  if (nan_or_inf(p[0])) return;
  const int xi = dindex(x);
  if (p[0] == T(0)) {
    set_bins(fold, xi, p, c);
    return;
  }
  raise_index(fold, dmindex(p[0]) - xi, xi, p, c);
}

// One term into primaries that admit it. `top` selects the index-0 path that
// deposits the first slice of x*compression (reference
// src/binned/dmddeposit.c:35-70, smsdeposit.c, and the scalar loops of
// src/binnedBLAS/dmdsum.ccog:5969-6192).
template <int CAP, class T>
RB_HD void deposit(int fold, bool top, T x, T (&p)[CAP]) {
  if (top) {
    const T m = p[0];
    const T q = m + set_lsb(x * Fmt<T>::compression);
    p[0] = q;
    const T r = (m - q) * (Fmt<T>::expansion * T(0.5));
    x = (x + r) + r;
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold - 1 && !(top && j == 0)) {
      const T m = p[j];
      const T q = m + set_lsb(x);
      p[j] = q;
      x = x + (m - q);
    }
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j == fold - 1) p[j] = p[j] + set_lsb(x);
}

// Two terms into two independent sets of primaries (each exactly as
// deposit() above), written step by step in lockstep so that the two
// dependency chains interleave in the instruction stream.
template <int CAP, class T>
RB_HD void deposit2(int fold, bool top, T xa, T (&pa)[CAP], T xb, T (&pb)[CAP]) {
  if (top) {
    const T ma = pa[0], mb = pb[0];
    const T qa = ma + set_lsb(xa * Fmt<T>::compression);
    const T qb = mb + set_lsb(xb * Fmt<T>::compression);
    pa[0] = qa;
    pb[0] = qb;
    const T ra = (ma - qa) * (Fmt<T>::expansion * T(0.5));
    const T rb = (mb - qb) * (Fmt<T>::expansion * T(0.5));
    xa = (xa + ra) + ra;
    xb = (xb + rb) + rb;
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold - 1 && !(top && j == 0)) {
      const T ma = pa[j], mb = pb[j];
      const T qa = ma + set_lsb(xa);
      const T qb = mb + set_lsb(xb);
      pa[j] = qa;
      pb[j] = qb;
      xa = xa + (ma - qa);
      xb = xb + (mb - qb);
    }
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j == fold - 1) {
      pa[j] = pa[j] + set_lsb(xa);
      pb[j] = pb[j] + set_lsb(xb);
    }
}

// reference src/binned/dmrenorm.c:70-85, smrenorm.c: move the two bits below
// the leading one into the carry so p returns to [1.5, 1.75) * ufp.
template <int CAP, class T>
RB_HD void renorm(int fold, T (&p)[CAP], T (&c)[CAP]) {
  using U = typename Fmt<T>::U;
  constexpr int lo = Lay<T>::mbits - 2;
  if (p[0] == T(0) || nan_or_inf(p[0])) return;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      U u = bits(p[j]);
      c[j] += (T)((int)((u >> lo) & 3) - 2);
      u &= ~((U)1 << lo);
      u |= (U)1 << (lo + 1);
      p[j] = fp_from_bits<T>(u);
    }
  }
}

// Add of two renormalised values that share index I (the common case inside
// a wave): exact, then renormalised. Equal to the general add below when the
// indices agree.
template <int CAP, class T>
RB_HD void add_same_index(int fold, int I, T (&py)[CAP], T (&cy)[CAP], const T (&px)[CAP], const T (&cx)[CAP]) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      py[j] = py[j] + (px[j] - bin<T>(I + j));
      cy[j] = cy[j] + cx[j];
    }
  }
  renorm(fold, py, cy);
}

// reference src/binned/dmdmadd.c:27-74, smsmadd.c: Y += X for renormalised X, Y.
template <int CAP, class T>
RB_HD void add(int fold, T (&py)[CAP], T (&cy)[CAP], const T (&px)[CAP], const T (&cx)[CAP]) {
  if (px[0] == T(0)) return;
  if (py[0] == T(0)) {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { py[j] = px[j]; cy[j] = cx[j]; }
    return;
  }
  if (nan_or_inf(px[0]) || nan_or_inf(py[0])) {
    py[0] += px[0];
    return;
  }
  const int xi = dmindex(px[0]);
  const int yi = dmindex(py[0]);
  const int shift = yi - xi;
  if (shift > 0) {
    T np[CAP], nc[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j) {
      if (j < fold) {
        if (j >= shift) {
          np[j] = px[j] + (pick(py, j - shift) - bin<T>(yi + j - shift));
          nc[j] = cx[j] + pick(cy, j - shift);
        } else {
          np[j] = px[j];
          nc[j] = cx[j];
        }
      }
    }
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j)
      if (j < fold) { py[j] = np[j]; cy[j] = nc[j]; }
  } else {
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j) {
      if (j < fold && j >= -shift) {
        py[j] += pick(px, j + shift) - bin<T>(xi + j + shift);
        cy[j] += pick(cx, j + shift);
      }
    }
  }
  renorm(fold, py, cy);
}

// reference src/binned/ddmconv.c:23-81: binned -> double in a fixed order;
// scaled by 2^-66 while the index is at most 3 so top bins cannot overflow.
template <int CAP>
RB_HD double convert(int fold, const double (&p)[CAP], const double (&c)[CAP]) {
  if (nan_or_inf(p[0])) return p[0];
  if (p[0] == 0.0) return 0.0;
  const int I = dmindex(p[0]);
  double Y = 0.0;
  if (I <= (3 * 53) / kWidth) {
    const double sdown = from_bits((uint64_t)(1023 - 66) << 52);  // ldexp(0.5, 1-(2*53-40))
    const double sup = from_bits((uint64_t)(1023 + 66) << 52);    // ldexp(0.5, 1+(2*53-40))
    int nscaled = (3 * 53) / kWidth - I;
    nscaled = nscaled < fold ? nscaled : fold;
    nscaled = nscaled > 0 ? nscaled : 0;
    int i;
    if (I == 0) {
      Y += c[0] * (carry_unit<double>(0) * sdown * kExpansion);
      Y += pick(c, 1) * (carry_unit<double>(1) * sdown);
      Y += (p[0] - bin<double>(0)) * sdown * kExpansion;
      i = 2;
    } else {
      Y += c[0] * (carry_unit<double>(I) * sdown);
      i = 1;
    }
    for (; i < nscaled; i++) {
      Y += pick(c, i) * (carry_unit<double>(I + i) * sdown);
      Y += (pick(p, i - 1) - bin<double>(I + i - 1)) * sdown;
    }
    if (i == fold) {
      Y += (pick(p, fold - 1) - bin<double>(I + fold - 1)) * sdown;
      return Y * sup;
    }
    if (__builtin_isinf(Y * sup)) return Y * sup;
    Y *= sup;
    for (; i < fold; i++) {
      Y += pick(c, i) * carry_unit<double>(I + i);
      Y += pick(p, i - 1) - bin<double>(I + i - 1);
    }
    Y += pick(p, fold - 1) - bin<double>(I + fold - 1);
    return Y;
  }
  Y += c[0] * carry_unit<double>(I);
RB_UNROLL_CAP
  for (int i = 1; i < CAP; ++i) {
    if (i < fold) {
      Y += c[i] * carry_unit<double>(I + i);
      Y += p[i - 1] - bin<double>(I + i - 1);
    }
  }
  Y += pick(p, fold - 1) - bin<double>(I + fold - 1);
  return Y;
}

// reference src/binned/ssmconv.c: binned float -> float, accumulated in
// double in a fixed order (every term exact in double), rounded once.
template <int CAP>
RB_HD float convert(int fold, const float (&p)[CAP], const float (&c)[CAP]) {
  if (nan_or_inf(p[0])) return p[0];
  if (p[0] == 0.0f) return 0.0f;
  const int I = dmindex(p[0]);
  double Y = 0.0;
  int i;
  if (I == 0) {
    Y += (double)c[0] * (double)carry_unit<float>(0) * (double)Fmt<float>::expansion;
    Y += (double)pick(c, 1) * (double)carry_unit<float>(1);
    Y += (double)(p[0] - bin<float>(0)) * (double)Fmt<float>::expansion;
    i = 2;
  } else {
    Y += (double)c[0] * (double)carry_unit<float>(I);
    i = 1;
  }
  for (; i < fold; i++) {
    Y += (double)pick(c, i) * (double)carry_unit<float>(I + i);
    Y += (double)(pick(p, i - 1) - bin<float>(I + i - 1));
  }
  Y += (double)(pick(p, fold - 1) - bin<float>(I + fold - 1));
  return (float)Y;
}

template <int CAP>
RB_HD double to_double(int fold, const double (&p)[CAP], const double (&c)[CAP]) {
  return convert(fold, p, c);
}

// reference src/binned/dmdrescale.c:20-37, smsrescale.c: divide primaries by
// (X/scaleY)^2.
template <int CAP, class T>
RB_HD void rescale(int fold, T X, T scaleY, T (&p)[CAP], T (&c)[CAP]) {
  if (X == scaleY || X == T(0) || scaleY == T(0)) return;
  T r = X / scaleY;
  r *= r;
  for (int i = 0; i < fold; ++i) {
    T v = pick(p, i) / r;
    if constexpr (CAP > 8) {
      p[i] = v;
    } else {
#pragma unroll
      for (int t = 0; t < CAP; ++t)
        if (t == i) p[t] = v;
    }
    if (v == T(0)) {
      // dmdupdate(fold - i, 0.0, p + i, c + i): slots i.. become the last bins (empty)
RB_UNROLL_CAP
      for (int t = 0; t < CAP; ++t)
        if (t >= i && t < fold) { p[t] = bin<T>(Fmt<T>::max_index + (t - i)); c[t] = T(0); }
      return;
    }
  }
}

// reference src/binned/dmdmaddsq.c:26-41, smsmaddsq.c: (scaleY, Y) += (scaleX, X).
// A non-finite top primary on either side carries the class through p[0]
// as the plain add does.
// reference src/binned/zmdrescale.c: both components of a complex binned
// value (p0/c0 real part, p1/c1 imaginary part) divided slot by slot by
// (X/scaleY)^2. Its zero test reads the interleaved primaries at i -- slot
// i/2 of component i%2 -- and on a zero it applies zmdupdate(fold - i, 0.0)
// to slots i.. of both components: a component whose slot i is zero gets the
// empty bins from i on (dmdupdate.c on an empty value); a nonzero one keeps
// its index (0.0 has the largest).
template <int CAP, class T>
RB_HD void zrescale(int fold, T X, T scaleY, T (&p0)[CAP], T (&c0)[CAP], T (&p1)[CAP], T (&c1)[CAP]) {
  if (X == scaleY || X == T(0) || scaleY == T(0)) return;
  T r = X / scaleY;
  r *= r;
  for (int i = 0; i < fold; ++i) {
    const T v0 = pick(p0, i) / r, v1 = pick(p1, i) / r;
RB_UNROLL_CAP
    for (int t = 0; t < CAP; ++t)
      if (t == i) {
        p0[t] = v0;
        p1[t] = v1;
      }
    const T z = (i & 1) ? pick(p1, i >> 1) : pick(p0, i >> 1);
    if (z == T(0)) {
RB_UNROLL_CAP
      for (int t = 0; t < CAP; ++t)
        if (t >= i && t < fold) {
          if (v0 == T(0)) {
            p0[t] = bin<T>(Fmt<T>::max_index + (t - i));
            c0[t] = T(0);
          }
          if (v1 == T(0)) {
            p1[t] = bin<T>(Fmt<T>::max_index + (t - i));
            c1[t] = T(0);
          }
        }
      return;
    }
  }
}

template <int CAP, class T>
RB_HD T add_scaled(int fold, T sx, const T (&px)[CAP], const T (&cx)[CAP], T sy, T (&py)[CAP], T (&cy)[CAP]) {
  if (sx > sy) {
    rescale(fold, sx, sy, py, cy);
    add(fold, py, cy, px, cx);
    return sx;
  }
  if (sx == sy) {
    add(fold, py, cy, px, cx);
    return sx;
  }
  T tp[CAP], tc[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) { tp[j] = px[j]; tc[j] = cx[j]; }
  rescale(fold, sy, sx, tp, tc);
  add(fold, py, cy, tp, tc);
  return sy;
}

}  // namespace rb
