// binned_core.hpp -- the binned FP64 arithmetic, written once for host and
// device (gfx950) code.
//
// A binned value of fold F holds F primaries p[0..F) and F carries c[0..F).
// p[j] accumulates slices of every term that fall in bin (I + j), where I is
// the index of the value; bins are 40 bits wide (reference
// include/binned.h:84). The slice that each bin receives from a term is the
// round-to-nearest of the remaining term at that bin's ulp, with ties made
// impossible by forcing the term's last mantissa bit to one. That slice does
// not depend on the accumulator's contents or on the order of terms, so any
// lane layout, tiling or merge tree gives the same bits, provided:
//   * every accumulator is renormalised at least every 2048 deposits,
//   * merges use the exact add below on renormalised operands,
//   * conversion follows ddmconv's fixed operation order.
// All code here must be compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RB_HD __host__ __device__ __forceinline__
// Loops over a binned value's slots: fully unrolled (registers) for the
// small capacities, plain loops (private memory) for the general one.
#define RB_UNROLL_CAP _Pragma("unroll (CAP <= 8 ? CAP : 1)")

namespace rb {

constexpr int kWidth = 40;        // reference include/binned.h:84
constexpr int kMaxIndex = 51;     // reference include/binned.h:105
constexpr int kMaxFold = 52;      // reference include/binned.h:123
constexpr int kEndurance = 2048;  // reference include/binned.h:142
constexpr double kCompression = 1.0 / (1 << 14);  // reference include/binned.h:184
constexpr double kExpansion = 1.0 * (1 << 14);    // reference include/binned.h:206

RB_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
RB_HD double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
// biased exponent field (reference src/common/common.h:43-47)
RB_HD int biased_exp(double x) { return (int)((bits(x) >> 52) & 0x7ff); }
// exponent all ones (reference src/common/common.h:59-63)
RB_HD bool nan_or_inf(double x) { return biased_exp(x) == 0x7ff; }
// the deposit's "|1": last mantissa bit forced to one (one v_or_b32 on gfx950)
RB_HD double set_lsb(double x) { return from_bits(bits(x) | 1ull); }

// Bin value of index i: 1.5*2^1023 for i = 0, 1.5*2^(1037-40i) for 1..51 and
// bin 51 repeated beyond (reference src/binned/dmbins.c:23-39). Built from
// bits: no table, no memory traffic.
RB_HD double bin(int i) {
  i = i > kMaxIndex ? kMaxIndex : i;
  const int e = (i == 0) ? 2046 : 2060 - kWidth * i;
  return from_bits(((uint64_t)e << 52) | (1ull << 51));
}
// bin(i)/6 = 2^(exponent(bin(i)) - 2), the weight of one carry unit; exact,
// so equal to the reference's division (src/binned/ddmconv.c:47-75).
RB_HD double carry_unit(int i) {
  i = i > kMaxIndex ? kMaxIndex : i;
  const int e = (i == 0) ? 2046 : 2060 - kWidth * i;
  return from_bits((uint64_t)(e - 2) << 52);
}
// Smallest magnitude whose index is below I: |t| < admit_limit(I) <=> dindex(t) >= I.
// 2^(1024-40I) for I >= 1 and +inf for I = 0 (from src/binned/dindex.c:40).
RB_HD double admit_limit(int I) { return from_bits((uint64_t)(2047 - kWidth * I) << 52); }

// reference src/binned/dindex.c:19-41. Zero and subnormals map to 51 (the
// frexp branch of the reference yields at least 51 for every subnormal, then
// caps at 51); NaN/Inf map to 0 through C's truncating division.
RB_HD int dindex(double x) {
  const int e = biased_exp(x);
  if (e == 0) return kMaxIndex;
  return (2046 - e) / kWidth;
}
// reference src/binned/dmindex.c:22-38
RB_HD int dmindex(double p0) { return (2060 - biased_exp(p0)) / kWidth; }
// reference src/binned/dmindex0.c:19-31
RB_HD bool dmindex0(double p0) { return biased_exp(p0) == 2046; }

// reference src/binned/dscale.c:25-31: a power of two 2^(e-1); e keeps C's
// truncating '%' on a possibly negative operand.
RB_HD double dscale(double x) {
  int e = biased_exp(x);
  e = e < kWidth ? kWidth : e;
  e -= (e - 1022 - 1) % kWidth;
  e -= 1022;
  return from_bits((uint64_t)(e - 1 + 1023) << 52);
}

// Register-array helpers. CAP is the compile-time capacity; `fold` <= CAP is
// the fold in use (a constant after inlining for the fixed-fold kernels).
template <int CAP>
RB_HD double pick(const double (&a)[CAP], int k) {
  if constexpr (CAP > 8) {
    return (k >= 0 && k < CAP) ? a[k] : 0.0;
  } else {
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < CAP; ++t) v = (t == k) ? a[t] : v;
    return v;
  }
}

template <int CAP>
RB_HD void set_zero(int fold, double (&p)[CAP], double (&c)[CAP]) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = 0.0; c[j] = 0.0; }
}

// Fresh primaries of index I (reference src/binned/dmdupdate.c:36-41).
template <int CAP>
RB_HD void set_bins(int fold, int I, double (&p)[CAP], double (&c)[CAP]) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = bin(I + j); c[j] = 0.0; }
}

// Raise a non-empty value to index I < current: slots move down by
// shift = current - I, the lowest fall off, fresh top bins come in
// (reference src/binned/dmdupdate.c:42-55).
template <int CAP>
RB_HD void raise_index(int fold, int shift, int I, double (&p)[CAP], double (&c)[CAP]) {
  if (shift <= 0) return;
  double np[CAP], nc[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      if (j >= shift) {
        np[j] = pick(p, j - shift);
        nc[j] = pick(c, j - shift);
      } else {
        np[j] = bin(I + j);
        nc[j] = 0.0;
      }
    }
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = np[j]; c[j] = nc[j]; }
}

// reference src/binned/dmdupdate.c:24-56 (update for a term of magnitude x)
template <int CAP>
RB_HD void update(int fold, double x, double (&p)[CAP], double (&c)[CAP]) {
  if (nan_or_inf(p[0])) return;
  const int xi = dindex(x);
  if (p[0] == 0.0) {
    set_bins(fold, xi, p, c);
    return;
  }
  raise_index(fold, dmindex(p[0]) - xi, xi, p, c);
}

// One term into primaries that admit it. `top` selects the index-0 path that
// deposits the first slice of x*2^-14 (reference src/binned/dmddeposit.c:35-70
// and the scalar loops of src/binnedBLAS/dmdsum.ccog:5969-6192).
template <int CAP>
RB_HD void deposit(int fold, bool top, double x, double (&p)[CAP]) {
  if (top) {
    const double m = p[0];
    const double q = m + set_lsb(x * kCompression);
    p[0] = q;
    const double r = (m - q) * (kExpansion * 0.5);
    x = (x + r) + r;
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold - 1 && !(top && j == 0)) {
      const double m = p[j];
      const double q = m + set_lsb(x);
      p[j] = q;
      x = x + (m - q);
    }
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j == fold - 1) p[j] = p[j] + set_lsb(x);
}

// reference src/binned/dmrenorm.c:70-85: move the two bits below the leading
// one into the carry so p returns to [1.5, 1.75) * ufp.
template <int CAP>
RB_HD void renorm(int fold, double (&p)[CAP], double (&c)[CAP]) {
  if (p[0] == 0.0 || nan_or_inf(p[0])) return;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      uint64_t u = bits(p[j]);
      c[j] += (double)((int)((u >> 50) & 3) - 2);
      u &= ~(1ull << 50);
      u |= 1ull << 51;
      p[j] = from_bits(u);
    }
  }
}

// Add of two renormalised values that share index I (the common case inside
// a wave): exact, then renormalised. Equal to the general add below when the
// indices agree.
template <int CAP>
RB_HD void add_same_index(int fold, int I, double (&py)[CAP], double (&cy)[CAP], const double (&px)[CAP],
                          const double (&cx)[CAP]) {
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      py[j] = py[j] + (px[j] - bin(I + j));
      cy[j] = cy[j] + cx[j];
    }
  }
  renorm(fold, py, cy);
}

// reference src/binned/dmdmadd.c:27-74: Y += X for renormalised X, Y.
template <int CAP>
RB_HD void add(int fold, double (&py)[CAP], double (&cy)[CAP], const double (&px)[CAP], const double (&cx)[CAP]) {
  if (px[0] == 0.0) return;
  if (py[0] == 0.0) {
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
    double np[CAP], nc[CAP];
RB_UNROLL_CAP
    for (int j = 0; j < CAP; ++j) {
      if (j < fold) {
        if (j >= shift) {
          np[j] = px[j] + (pick(py, j - shift) - bin(yi + j - shift));
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
        py[j] += pick(px, j + shift) - bin(xi + j + shift);
        cy[j] += pick(cx, j + shift);
      }
    }
  }
  renorm(fold, py, cy);
}

// reference src/binned/ddmconv.c:23-81: binned -> double in a fixed order;
// scaled by 2^-66 while the index is at most 3 so top bins cannot overflow.
template <int CAP>
RB_HD double to_double(int fold, const double (&p)[CAP], const double (&c)[CAP]) {
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
      Y += c[0] * (carry_unit(0) * sdown * kExpansion);
      Y += pick(c, 1) * (carry_unit(1) * sdown);
      Y += (p[0] - bin(0)) * sdown * kExpansion;
      i = 2;
    } else {
      Y += c[0] * (carry_unit(I) * sdown);
      i = 1;
    }
    for (; i < nscaled; i++) {
      Y += pick(c, i) * (carry_unit(I + i) * sdown);
      Y += (pick(p, i - 1) - bin(I + i - 1)) * sdown;
    }
    if (i == fold) {
      Y += (pick(p, fold - 1) - bin(I + fold - 1)) * sdown;
      return Y * sup;
    }
    if (__builtin_isinf(Y * sup)) return Y * sup;
    Y *= sup;
    for (; i < fold; i++) {
      Y += pick(c, i) * carry_unit(I + i);
      Y += pick(p, i - 1) - bin(I + i - 1);
    }
    Y += pick(p, fold - 1) - bin(I + fold - 1);
    return Y;
  }
  Y += c[0] * carry_unit(I);
RB_UNROLL_CAP
  for (int i = 1; i < CAP; ++i) {
    if (i < fold) {
      Y += c[i] * carry_unit(I + i);
      Y += p[i - 1] - bin(I + i - 1);
    }
  }
  Y += pick(p, fold - 1) - bin(I + fold - 1);
  return Y;
}

// reference src/binned/dmdrescale.c:20-37: divide primaries by (X/scaleY)^2.
template <int CAP>
RB_HD void rescale(int fold, double X, double scaleY, double (&p)[CAP], double (&c)[CAP]) {
  if (X == scaleY || X == 0.0 || scaleY == 0.0) return;
  double r = X / scaleY;
  r *= r;
  for (int i = 0; i < fold; ++i) {
    double v = pick(p, i) / r;
    if constexpr (CAP > 8) {
      p[i] = v;
    } else {
#pragma unroll
      for (int t = 0; t < CAP; ++t)
        if (t == i) p[t] = v;
    }
    if (v == 0.0) {
      // dmdupdate(fold - i, 0.0, p + i, c + i): slots i.. become bins 51.. (empty)
RB_UNROLL_CAP
      for (int t = 0; t < CAP; ++t)
        if (t >= i && t < fold) { p[t] = bin(kMaxIndex + (t - i)); c[t] = 0.0; }
      return;
    }
  }
}

// reference src/binned/dmdmaddsq.c:26-41: (scaleY, Y) += (scaleX, X).
// A non-finite top primary on either side carries the class through p[0]
// as the plain add does.
template <int CAP>
RB_HD double add_scaled(int fold, double sx, const double (&px)[CAP], const double (&cx)[CAP], double sy,
                        double (&py)[CAP], double (&cy)[CAP]) {
  if (sx > sy) {
    rescale(fold, sx, sy, py, cy);
    add(fold, py, cy, px, cx);
    return sx;
  }
  if (sx == sy) {
    add(fold, py, cy, px, cx);
    return sx;
  }
  double tp[CAP], tc[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) { tp[j] = px[j]; tc[j] = cx[j]; }
  rescale(fold, sy, sx, tp, tc);
  add(fold, py, cy, tp, tc);
  return sy;
}

}  // namespace rb
