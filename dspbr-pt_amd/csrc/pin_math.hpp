// Pinned float transcendentals: one implementation of sin, cos, exp, log,
// pow, atan2 and acos, compiled into the gfx950 kernels and into the CPU
// oracle's float build, so both sides produce the same bits.
//
// GLSL leaves the precision of these functions to the implementation, so the
// reference has no single answer for them; this file is the engine's choice.
// Every function takes float arguments, evaluates in IEEE double with only
// +, -, *, / and explicit fused multiply-adds (fma: one IEEE operation,
// rounded once; no library call on the GPU, the C library's fma on the
// host) and rounds once to float, so any compiler that keeps IEEE semantics
// (-ffp-contract=off, no fast-math) gives identical results on the host and
// on the GPU. Accuracy:
// the double result is within ~1e-13 relative of the exact value, so the
// float result is the correctly rounded one except within ~1e-13 relative of
// a rounding boundary (at most one ulp off there).
//
// Methods (all published):
//  * sin/cos: Cody & Waite range reduction by pi/2 split into 28 + 28 + 53
//    bits for |x| < 2^25, Payne & Hanek reduction with the bits of 2/pi above
//    that; Taylor polynomials of degree 15 / 16 on |r| <= pi/4.
//  * exp: Cody & Waite reduction by ln 2 (42 + 53 bits), Taylor degree 12 on
//    |r| <= ln2/2, scaling by 2^k through the exponent field.
//  * log: x = m 2^e with m in [sqrt(2)/2, sqrt(2)], log m = 2 atanh((m-1)/(m+1))
//    by its series to s^17.
//  * pow: for an integer y with |y| <= 8, binary powering in double (x^2 is
//    exact), its reciprocal for y < 0; otherwise exp(y log|x|) in double;
//    with the C99 (Annex F) special cases.
//  * atan2/acos: atan by reduction to [0,1] (reciprocal) and to
//    |u| <= 1/16 around k/8 (atan(z) = atan(k/8) + atan((z-k/8)/(1+z k/8))),
//    series to u^13; acos(x) = atan2(sqrt((1-x)(1+x)), x), the double square
//    root by one Newton step from the correctly rounded float root.
// The constants are derived in exact arithmetic by gen_pin_math.py.
#pragma once
#include <stdint.h>

// PM_FN: small helpers, always inlined. PM_BIG: the general functions, one
// out-of-line copy per kernel (inlining each call site of the double cores
// would keep their 64-bit constants and temporaries live beside the
// caller's state). PM_COLD: the huge-argument reduction.
#ifndef PM_INLINE_BIG
#define PM_INLINE_BIG 0
#endif
// PM_CONST: a result that depends on the arguments only, so repeated calls
// with equal arguments may be merged.
#define PM_CONST __attribute__((const))
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PM_FN __host__ __device__ inline __attribute__((always_inline))
#define PM_COLD __host__ __device__ __attribute__((noinline))
#if PM_INLINE_BIG
#define PM_BIG PM_FN
#else
#define PM_BIG PM_COLD
#endif
#else
#define PM_FN inline __attribute__((always_inline))
#define PM_COLD inline __attribute__((noinline))
#define PM_BIG PM_COLD
#endif

namespace pm {
// ---- BEGIN GENERATED CONSTANTS (gen_pin_math.py) ----
constexpr double PM_PI = 0x1.921fb54442d18p+1;
constexpr double PM_PIO2 = 0x1.921fb54442d18p+0;
constexpr double PM_2_PI = 0x1.45f306dc9c883p-1;
constexpr double PM_PIO2_1 = 0x1.921fb54000000p+0;  // 28 leading bits of pi/2
constexpr double PM_PIO2_2 = 0x1.10b4610000000p-30;  // next 28 bits
constexpr double PM_PIO2_3 = 0x1.a62633145c06ep-58;  // the rest, rounded
constexpr double PM_LN2 = 0x1.62e42fefa39efp-1;
constexpr double PM_LN2_HI = 0x1.62e42fefa3800p-1;  // 42 leading bits of ln 2
constexpr double PM_LN2_LO = 0x1.ef35793c76730p-45;
constexpr double PM_INV_LN2 = 0x1.71547652b82fep+0;
constexpr double PM_SQRT2 = 0x1.6a09e667f3bcdp+0;  // split point of the log reduction
// fraction bits of 2/pi, 32 per word, most significant first
constexpr uint32_t PM_TOP0 = 0xa2f9836eu;
constexpr uint32_t PM_TOP1 = 0x4e441529u;
constexpr uint32_t PM_TOP2 = 0xfc2757d1u;
constexpr uint32_t PM_TOP3 = 0xf534ddc0u;
constexpr uint32_t PM_TOP4 = 0xdb629599u;
constexpr uint32_t PM_TOP5 = 0x3c439041u;
constexpr uint32_t PM_TOP6 = 0xfe5163abu;
constexpr uint32_t PM_TOP7 = 0xdebbc561u;
// atan(k/8)
constexpr double PM_ATAN8_0 = 0x0.0p+0;
constexpr double PM_ATAN8_1 = 0x1.fd5ba9aac2f6ep-4;
constexpr double PM_ATAN8_2 = 0x1.f5b75f92c80ddp-3;
constexpr double PM_ATAN8_3 = 0x1.6f61941e4def1p-2;
constexpr double PM_ATAN8_4 = 0x1.dac670561bb4fp-2;
constexpr double PM_ATAN8_5 = 0x1.1e00babdefeb4p-1;
constexpr double PM_ATAN8_6 = 0x1.4978fa3269ee1p-1;
constexpr double PM_ATAN8_7 = 0x1.700a7c5784634p-1;
constexpr double PM_ATAN8_8 = 0x1.921fb54442d18p-1;
// sin r = r + r^3 (S3 + r^2 (S5 + ...)), S_n = (-1)^((n-1)/2) / n!
constexpr double PM_S3 = -0x1.5555555555555p-3;
constexpr double PM_S5 = 0x1.1111111111111p-7;
constexpr double PM_S7 = -0x1.a01a01a01a01ap-13;
constexpr double PM_S9 = 0x1.71de3a556c734p-19;
constexpr double PM_S11 = -0x1.ae64567f544e4p-26;
constexpr double PM_S13 = 0x1.6124613a86d09p-33;
constexpr double PM_S15 = -0x1.ae7f3e733b81fp-41;
// cos r = 1 + r^2 (C2 + r^2 (C4 + ...)), C_n = (-1)^(n/2) / n!
constexpr double PM_C2 = -0x1.0000000000000p-1;
constexpr double PM_C4 = 0x1.5555555555555p-5;
constexpr double PM_C6 = -0x1.6c16c16c16c17p-10;
constexpr double PM_C8 = 0x1.a01a01a01a01ap-16;
constexpr double PM_C10 = -0x1.27e4fb7789f5cp-22;
constexpr double PM_C12 = 0x1.1eed8eff8d898p-29;
constexpr double PM_C14 = -0x1.93974a8c07c9dp-37;
constexpr double PM_C16 = 0x1.ae7f3e733b81fp-45;
// exp r = 1 + r + r^2 (E2 + r (E3 + ...)), E_n = 1 / n!
constexpr double PM_E2 = 0x1.0000000000000p-1;
constexpr double PM_E3 = 0x1.5555555555555p-3;
constexpr double PM_E4 = 0x1.5555555555555p-5;
constexpr double PM_E5 = 0x1.1111111111111p-7;
constexpr double PM_E6 = 0x1.6c16c16c16c17p-10;
constexpr double PM_E7 = 0x1.a01a01a01a01ap-13;
constexpr double PM_E8 = 0x1.a01a01a01a01ap-16;
constexpr double PM_E9 = 0x1.71de3a556c734p-19;
constexpr double PM_E10 = 0x1.27e4fb7789f5cp-22;
constexpr double PM_E11 = 0x1.ae64567f544e4p-26;
constexpr double PM_E12 = 0x1.1eed8eff8d898p-29;
// atanh s = s + s^3 (L3 + s^2 (L5 + ...)), L_n = 1 / n
constexpr double PM_L3 = 0x1.5555555555555p-2;
constexpr double PM_L5 = 0x1.999999999999ap-3;
constexpr double PM_L7 = 0x1.2492492492492p-3;
constexpr double PM_L9 = 0x1.c71c71c71c71cp-4;
constexpr double PM_L11 = 0x1.745d1745d1746p-4;
constexpr double PM_L13 = 0x1.3b13b13b13b14p-4;
constexpr double PM_L15 = 0x1.1111111111111p-4;
constexpr double PM_L17 = 0x1.e1e1e1e1e1e1ep-5;
// atan u = u + u^3 (A3 + u^2 (A5 + ...)), A_n = (-1)^((n-1)/2) / n
constexpr double PM_A3 = -0x1.5555555555555p-2;
constexpr double PM_A5 = 0x1.999999999999ap-3;
constexpr double PM_A7 = -0x1.2492492492492p-3;
constexpr double PM_A9 = 0x1.c71c71c71c71cp-4;
constexpr double PM_A11 = -0x1.745d1745d1746p-4;
constexpr double PM_A13 = 0x1.3b13b13b13b14p-4;
// ---- END GENERATED CONSTANTS ----

// ---------------------------------------------------------------- bit helpers
PM_FN double d_from_bits(uint64_t u) { double d; __builtin_memcpy(&d, &u, 8); return d; }
PM_FN uint64_t bits_of_d(double d) { uint64_t u; __builtin_memcpy(&u, &d, 8); return u; }
PM_FN float f_from_bits(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
PM_FN uint32_t bits_of_f(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
PM_FN float f_inf() { return f_from_bits(0x7f800000u); }
PM_FN float f_nan() { return f_from_bits(0x7fc00000u); }
PM_FN float f_abs(float x) { return f_from_bits(bits_of_f(x) & 0x7fffffffu); }
PM_FN bool f_signbit(float x) { return (bits_of_f(x) >> 31) != 0u; }
// nearest integer (ties to even) of |z| < 2^51, as a double
PM_FN double rint_d(double z) { return (z + 0x1.8p52) - 0x1.8p52; }
PM_FN double fma_d(double a, double b, double c) { return __builtin_fma(a, b, c); }
// 2^k for -1022 <= k <= 1023
PM_FN double pow2i(int k) { return d_from_bits((uint64_t)(k + 1023) << 52); }

// ---------------------------------------------------------------- double cores
// e^z for |z| <= 700
PM_FN double exp_d(double z) {
  const double kd = rint_d(z * PM_INV_LN2);
  const double r = fma_d(-kd, PM_LN2_LO, fma_d(-kd, PM_LN2_HI, z));
  double p = PM_E12;
  p = fma_d(p, r, PM_E11);
  p = fma_d(p, r, PM_E10);
  p = fma_d(p, r, PM_E9);
  p = fma_d(p, r, PM_E8);
  p = fma_d(p, r, PM_E7);
  p = fma_d(p, r, PM_E6);
  p = fma_d(p, r, PM_E5);
  p = fma_d(p, r, PM_E4);
  p = fma_d(p, r, PM_E3);
  p = fma_d(p, r, PM_E2);
  const double e = 1.0 + fma_d(r * r, p, r);
  return e * pow2i((int)kd);
}

// ln x for a positive, finite, normal double
PM_FN double log_d(double x) {
  const uint64_t u = bits_of_d(x);
  int e = (int)(u >> 52) - 1023;
  double m = d_from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > PM_SQRT2) {
    m = m * 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = PM_L17;
  p = fma_d(p, s2, PM_L15);
  p = fma_d(p, s2, PM_L13);
  p = fma_d(p, s2, PM_L11);
  p = fma_d(p, s2, PM_L9);
  p = fma_d(p, s2, PM_L7);
  p = fma_d(p, s2, PM_L5);
  p = fma_d(p, s2, PM_L3);
  const double at = fma_d(s * s2, p, s);  // atanh(s)
  return fma_d((double)e, PM_LN2, 2.0 * at);
}

// atan(k/8) for k = 0..8
PM_FN double atan8(int k) {
  double a = PM_ATAN8_8;
  a = k == 7 ? PM_ATAN8_7 : a;
  a = k == 6 ? PM_ATAN8_6 : a;
  a = k == 5 ? PM_ATAN8_5 : a;
  a = k == 4 ? PM_ATAN8_4 : a;
  a = k == 3 ? PM_ATAN8_3 : a;
  a = k == 2 ? PM_ATAN8_2 : a;
  a = k == 1 ? PM_ATAN8_1 : a;
  a = k == 0 ? PM_ATAN8_0 : a;
  return a;
}

// atan t for t >= 0 (t may be +inf)
PM_FN double atan_d(double t) {
  const bool inv = t > 1.0;
  const double z = inv ? 1.0 / t : t;
  const double kd = rint_d(z * 8.0);
  const double c = kd * 0.125;
  const double u = (z - c) / (1.0 + z * c);
  const double u2 = u * u;
  double p = PM_A13;
  p = fma_d(p, u2, PM_A11);
  p = fma_d(p, u2, PM_A9);
  p = fma_d(p, u2, PM_A7);
  p = fma_d(p, u2, PM_A5);
  p = fma_d(p, u2, PM_A3);
  const double a = atan8((int)kd) + fma_d(u * u2, p, u);
  return inv ? PM_PIO2 - a : a;
}

// sqrt v for 0 <= v <= 1: one Newton step from the correctly rounded float root
PM_FN double sqrt_unit_d(double v) {
  if (v == 0.0) return 0.0;
  const double y = (double)__builtin_sqrtf((float)v);
  return 0.5 * (y + v / y);
}

// Taylor polynomials on |r| <= pi/4
PM_FN double sin_poly(double r) {
  const double r2 = r * r;
  double p = PM_S15;
  p = fma_d(p, r2, PM_S13);
  p = fma_d(p, r2, PM_S11);
  p = fma_d(p, r2, PM_S9);
  p = fma_d(p, r2, PM_S7);
  p = fma_d(p, r2, PM_S5);
  p = fma_d(p, r2, PM_S3);
  return fma_d(r * r2, p, r);
}
PM_FN double cos_poly(double r) {
  const double r2 = r * r;
  double p = PM_C16;
  p = fma_d(p, r2, PM_C14);
  p = fma_d(p, r2, PM_C12);
  p = fma_d(p, r2, PM_C10);
  p = fma_d(p, r2, PM_C8);
  p = fma_d(p, r2, PM_C6);
  p = fma_d(p, r2, PM_C4);
  p = fma_d(p, r2, PM_C2);
  return fma_d(r2, p, 1.0);
}

// 32 bits of 2/pi starting at fraction bit `pos` (1-based), pos <= 192
PM_FN uint32_t top_word(int i) {
  uint32_t w = 0u;
  w = i == 7 ? PM_TOP7 : w;
  w = i == 6 ? PM_TOP6 : w;
  w = i == 5 ? PM_TOP5 : w;
  w = i == 4 ? PM_TOP4 : w;
  w = i == 3 ? PM_TOP3 : w;
  w = i == 2 ? PM_TOP2 : w;
  w = i == 1 ? PM_TOP1 : w;
  w = i == 0 ? PM_TOP0 : w;
  return w;
}
PM_FN uint32_t top_bits32(int pos) {
  const int i = (pos - 1) >> 5, s = (pos - 1) & 31;
  const uint32_t a = top_word(i), b = top_word(i + 1);
  return s ? ((a << s) | (b >> (32 - s))) : a;
}

// Payne-Hanek: ax = (4j + q) pi/2 + r for a finite float ax >= 2^25
PM_COLD double reduce_big(float ax, int* qo) {
  const uint32_t b = bits_of_f(ax);
  const int e = (int)(b >> 23) - 150;  // ax = M 2^e, e >= 2
  const uint32_t M = (b & 0x7fffffu) | 0x800000u;
  const int p = e - 1;  // fraction bits of 2/pi before p only add multiples of 4
  const uint64_t lo = (uint64_t)M * top_bits32(p + 64);
  const uint64_t mid = (uint64_t)M * top_bits32(p + 32) + (lo >> 32);
  const uint64_t hi = (uint64_t)M * top_bits32(p) + (mid >> 32);
  // the low 96 bits of M * W are (ax * 2/pi mod 4) * 2^94
  const uint32_t h = (uint32_t)hi;
  int q = (int)(h >> 30);
  const uint64_t f = ((uint64_t)(h & 0x3fffffffu) << 34) | ((uint64_t)(uint32_t)mid << 2) | (uint64_t)((uint32_t)lo >> 30);
  double fd;
  if (f >> 63) {
    q = (q + 1) & 3;
    fd = -(double)(0ull - f);
  } else {
    fd = (double)f;
  }
  *qo = q;
  return fd * 0x1p-64 * PM_PIO2;
}

// ax = (4j + q) pi/2 + r, |r| <= pi/4 (up to rounding of the quotient)
PM_FN double reduce_pio2(float ax, int& q) {
  if (ax >= 33554432.0f) return reduce_big(ax, &q);
  const double x = (double)ax;
  const double kd = rint_d(x * PM_2_PI);
  q = (int)kd & 3;
  return fma_d(-kd, PM_PIO2_3, fma_d(-kd, PM_PIO2_2, fma_d(-kd, PM_PIO2_1, x)));
}

// ---------------------------------------------------------------- float API
PM_BIG PM_CONST float psin(float x) {
  const float ax = f_abs(x);
  if (!(ax < f_inf())) return f_nan();
  int q;
  const double r = reduce_pio2(ax, q);
  double v = (q & 1) ? cos_poly(r) : sin_poly(r);
  v = (q & 2) ? -v : v;
  return (float)(f_signbit(x) ? -v : v);
}

PM_BIG PM_CONST float pcos(float x) {
  const float ax = f_abs(x);
  if (!(ax < f_inf())) return f_nan();
  int q;
  const double r = reduce_pio2(ax, q);
  q += 1;
  double v = (q & 1) ? cos_poly(r) : sin_poly(r);
  return (float)((q & 2) ? -v : v);
}

// sin and cos of one argument, identical to psin(x) and pcos(x); returned by
// value (an out-of-line function writing through references would go
// through private memory on the GPU)
struct SinCos {
  float s, c;
};
PM_BIG PM_CONST SinCos sincos_pair(float x) {
  const float ax = f_abs(x);
  if (!(ax < f_inf())) return SinCos{f_nan(), f_nan()};
  int q;
  const double r = reduce_pio2(ax, q);
  const double sp = sin_poly(r), cp = cos_poly(r);
  double vs = (q & 1) ? cp : sp;
  vs = (q & 2) ? -vs : vs;
  double vc = (q & 1) ? sp : cp;
  vc = ((q + 1) & 2) ? -vc : vc;
  return SinCos{(float)(f_signbit(x) ? -vs : vs), (float)vc};
}
PM_FN void psincos(float x, float& s, float& c) {
  const SinCos r = sincos_pair(x);
  s = r.s;
  c = r.c;
}

PM_BIG PM_CONST float pexp(float x) {
  if (x != x) return x;
  if (x > 89.0f) return f_inf();
  if (x < -104.0f) return 0.0f;
  return (float)exp_d((double)x);
}

PM_BIG PM_CONST float plog(float x) {
  if (x != x) return x;
  if (x <= 0.0f) return x == 0.0f ? -f_inf() : f_nan();
  if (x == f_inf()) return x;
  return (float)log_d((double)x);
}

// y finite: is y an integer / an odd integer
PM_FN bool f_is_int(float y) {
  if (f_abs(y) >= 16777216.0f) return true;
  return (float)(int)y == y;
}
PM_FN bool f_is_odd(float y) {
  if (f_abs(y) >= 16777216.0f) return false;
  const int i = (int)y;
  return (float)i == y && (i & 1) != 0;
}

// x^n for 1 <= n <= 8 by binary powering (the first square is exact)
PM_FN double powi_d(double x, int n) {
  double r = (n & 1) ? x : 1.0, b = x;
  for (n >>= 1; n; n >>= 1) {
    b = b * b;
    if (n & 1) r = r * b;
  }
  return r;
}

// pow for y not an integer of magnitude <= 8 (nor 0)
PM_BIG PM_CONST float ppow_general(float x, float y) {
  if (x == 1.0f) return 1.0f;
  if (x != x || y != y) return f_nan();
  const float ax = f_abs(x), ay = f_abs(y);
  if (ay == f_inf()) {
    if (ax == 1.0f) return 1.0f;
    return ((ax > 1.0f) == (y > 0.0f)) ? f_inf() : 0.0f;
  }
  const bool odd = f_is_odd(y);
  if (ax == 0.0f) {
    if (y < 0.0f) return (odd && f_signbit(x)) ? -f_inf() : f_inf();
    return odd ? x : 0.0f;
  }
  if (ax == f_inf()) {
    const float r = y < 0.0f ? 0.0f : f_inf();
    return (odd && x < 0.0f) ? -r : r;
  }
  bool neg = false;
  if (x < 0.0f) {
    if (!f_is_int(y)) return f_nan();
    neg = odd;
  }
  const double z = (double)y * log_d((double)ax);
  float r;
  if (z > 200.0) r = f_inf();
  else if (z < -200.0) r = 0.0f;
  else r = (float)exp_d(z);
  return neg ? -r : r;
}

PM_FN float ppow(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (f_abs(y) <= 8.0f && (float)(int)y == y) {
    // integer exponent: exact signs, zeros and infinities follow from the
    // products; NaN stays NaN; x = 1 gives 1
    const double p = powi_d((double)x, y < 0.0f ? -(int)y : (int)y);
    return (float)(y < 0.0f ? 1.0 / p : p);
  }
  return ppow_general(x, y);
}


PM_BIG PM_CONST float patan2(float y, float x) {
  if (x != x || y != y) return f_nan();
  const bool xs = f_signbit(x), ys = f_signbit(y);
  const float ax = f_abs(x), ay = f_abs(y);
  double a;
  if (ay == 0.0f) a = xs ? PM_PI : 0.0;
  else if (ax == 0.0f) a = PM_PIO2;
  else if (ay == f_inf()) a = (ax == f_inf()) ? (xs ? 0.75 * PM_PI : 0.5 * PM_PIO2) : PM_PIO2;
  else if (ax == f_inf()) a = xs ? PM_PI : 0.0;
  else {
    a = atan_d((double)ay / (double)ax);
    a = xs ? PM_PI - a : a;
  }
  const float r = (float)a;
  return ys ? -r : r;
}

PM_BIG PM_CONST float pacos(float x) {
  if (!(f_abs(x) <= 1.0f)) return f_nan();
  const double d = (double)x;
  const double s = sqrt_unit_d((1.0 - d) * (1.0 + d));
  double a;
  if (d == 0.0) a = PM_PIO2;
  else {
    a = atan_d(s / (d < 0.0 ? -d : d));
    a = d < 0.0 ? PM_PI - a : a;
  }
  return (float)a;
}

}  // namespace pm
