// efdc_fmt.h -- Fortran edit descriptors E17.8 and I10 as the Fortran
// runtime writes them, one value per call, for the restart writer's kernel
// (efdc_io.hip).  The functions are host-and-device so that the CPU tests
// can run the same code over millions of values.
//
// Ew.d (w = 17, d = 8): 0.d1..d8 with the exponent as E+ee / E-ee, or
// +eee / -eee when it has three digits, right-justified; the digits are the
// binary value rounded to nearest, ties to even.  The value is scaled by
// 10^(8-E), held as a double-double pair from efdc_pow10.h (relative error
// below 2^-105, so the scaled value is within 2^-76 of exact); the rounding
// direction is taken when the scaled value lies farther than 2^-60 from a
// half-integer, and otherwise (ties and near-ties), or for magnitudes
// outside [1e-280, 1e280], fmt_e17 returns false and the caller formats
// the value from its exact decimal expansion.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>

#include "efdc_pow10.h"

namespace efdc_fmt {

constexpr int E_W = 17;  // E17.8
constexpr int I_W = 10;  // I10

// a >= 10^j, exactly: hi + lo orders every double as 10^j does (checked
// when the table is generated), and a - hi is exact where it is evaluated
__host__ __device__ __forceinline__ bool ge_pow10(const double (*t)[2], double a, int j) {
  const double hi = t[j - EFDC_POW10_MIN][0], lo = t[j - EFDC_POW10_MIN][1];
  if (a >= 2.0 * hi) return true;
  if (a < 0.5 * hi) return false;
  return (a - hi) >= lo;
}

// "  -0.dddddddd" and the exponent, at fixed positions (the text is
// right-justified in 17 and always 14 or 15 characters long)
__host__ __device__ __forceinline__ void put_e17(char *o, bool neg, uint32_t N, int E) {
  o[0] = ' ';
  o[1] = ' ';
  o[2] = neg ? '-' : ' ';
  o[3] = '0';
  o[4] = '.';
#pragma unroll
  for (int i = 7; i >= 0; --i) {
    o[5 + i] = (char)('0' + N % 10u);
    N /= 10u;
  }
  const int ae = E < 0 ? -E : E;
  if (ae <= 99) {
    o[13] = 'E';
    o[14] = E < 0 ? '-' : '+';
  } else {
    o[13] = E < 0 ? '-' : '+';
    o[14] = (char)('0' + ae / 100);
  }
  o[15] = (char)('0' + (ae / 10) % 10);
  o[16] = (char)('0' + ae % 10);
}

// NaN, Inf, -Inf right-justified in 17
__host__ __device__ __forceinline__ void put_special(char *o, int which) {
#pragma unroll
  for (int i = 0; i < 13; ++i) o[i] = ' ';
  if (which == 0) {
    o[13] = ' ';
    o[14] = 'N';
    o[15] = 'a';
    o[16] = 'N';
  } else {
    o[13] = which < 0 ? '-' : ' ';
    o[14] = 'I';
    o[15] = 'n';
    o[16] = 'f';
  }
}

// E17.8 of v into o[0..16]; false: the caller formats v
__host__ __device__ inline bool fmt_e17(const double (*t)[2], double v, char *o) {
  if (std::isnan(v)) {
    put_special(o, 0);
    return true;
  }
  const bool neg = std::signbit(v);
  const double a = std::fabs(v);
  if (std::isinf(a)) {
    put_special(o, neg ? -1 : 1);
    return true;
  }
  if (a == 0.0) {
    put_e17(o, neg, 0, 0);
    return true;
  }
  if (!(a >= 1e-280 && a <= 1e280)) return false;
  // With e2 = floor(log2 a) and E = floor(e2 log10 2) + 1 (the integer
  // form is exact for |e2| < 1650): 10^(E-1) <= 2^e2 <= a, and
  // a < 2^(e2+1) < 2 * 10^E, so a < 10^E or 10^E <= a < 10^(E+1): one
  // exact comparison fixes E.  The three table rows are loaded together.
  int64_t bits;
  __builtin_memcpy(&bits, &a, sizeof bits);
  const int e2 = (int)((bits >> 52) & 0x7ff) - 1023;
  int E = ((e2 * 78913) >> 18) + 1;
  const double h0 = t[8 - E - EFDC_POW10_MIN][0], l0 = t[8 - E - EFDC_POW10_MIN][1];
  const double h1 = t[7 - E - EFDC_POW10_MIN][0], l1 = t[7 - E - EFDC_POW10_MIN][1];
  const bool up = ge_pow10(t, a, E);
  if (up) ++E;
  const double hi = up ? h1 : h0, lo = up ? l1 : l0;
  // x = a * 10^k in [1e7, 1e8) is p + t with |x - (p + t)| < 2^-76
  const double p = a * hi;
  const double e = std::fma(a, hi, -p);
  const double q = a * lo;
  const double tt = e + q;
  const double f = std::floor(p);
  const double d = ((p - f) - 0.5) + tt;  // x - (f + 1/2); the first two steps are exact
  if (std::fabs(d) <= 0x1p-60) return false;
  uint32_t N = (uint32_t)f + (d > 0.0 ? 1u : 0u);
  if (N >= 100000000u) {
    N = 10000000u;
    ++E;
  }
  put_e17(o, neg, N, E);
  return true;
}

// I10 of n into o[0..9] (ten asterisks when it does not fit)
__host__ __device__ inline void fmt_i10(long long n, char *o) {
  if (n > 9999999999LL || n < -999999999LL) {
#pragma unroll
    for (int i = 0; i < I_W; ++i) o[i] = '*';
    return;
  }
  const bool neg = n < 0;
  unsigned long long u = neg ? (unsigned long long)(-n) : (unsigned long long)n;
  bool sign_due = neg;
#pragma unroll
  for (int i = I_W - 1; i >= 0; --i) {
    char ch;
    if (i == I_W - 1 || u != 0) {
      ch = (char)('0' + (int)(u % 10));
      u /= 10;
    } else if (sign_due) {
      ch = '-';
      sign_due = false;
    } else {
      ch = ' ';
    }
    o[i] = ch;
  }
}

}  // namespace efdc_fmt
