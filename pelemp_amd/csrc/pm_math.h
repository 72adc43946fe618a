// pm_math.h -- project-owned FP64 elementary functions.
//
// The spray update branches on values produced by transcendental functions:
// the substep count nsub = int(dt/tau)+1, the convergence test of the
// Spalding heat-transfer fixed point (|B_T - B_T_old| > 1e-7), the
// boiling/min-mass deletion tests.  A one-ulp difference between two libm
// implementations can flip such a decision, so the GPU kernels and the CPU
// checker both evaluate exp/log/log1p/pow/cbrt/sin/cos/atan/acos with the
// code below.  Every routine is a fixed sequence of IEEE-754 double
// operations (+ - * / sqrt fma rint ldexp frexp), each of which is correctly
// rounded on both x86-64 (SSE2/FMA3) and gfx950, so the results are bitwise
// identical on host and device when the including translation unit is
// compiled with -ffp-contract=off (no implicit fusing).
//
// Methods (published, none copied from a vendor library):
//  * exp: Cody-Waite argument reduction x = k ln2 + r, |r| <= ln2/2, with a
//    two-part ln2, then e^r = 1 + r s(r) with s of degree 10 from a
//    Chebyshev fit of (e^r - 1)/r in 50-digit arithmetic (error < 1.9e-17),
//    in Horner form with fma, then scaling by 2^k with ldexp.
//    (W. J. Cody and W. Waite, Software Manual for the Elementary Functions,
//    Prentice-Hall 1980.)
//  * log: x = 2^e m, m in [sqrt(1/2), sqrt(2)); f = m-1 (exact by Sterbenz),
//    s = f/(2+f), log(1+f) = 2 atanh(s) = f - s (f - R(s^2)) with the
//    atanh series R(z) = sum_{i>=1} 2 z^i/(2i+1), evaluated as z Q(z) with
//    Q a degree-6 Chebyshev fit of the series' quotient (error < 3.2e-16).
//  * log1p: D. Goldberg, "What every computer scientist should know about
//    floating-point arithmetic", ACM Comput. Surv. 23 (1991), thm. 4:
//    log1p(x) = x                         if 1+x == 1
//             = log(1+x) * x / ((1+x)-1)  otherwise.
//  * pow(x,y), x > 0: exp(y log x).
//  * cbrt: x = 2^(3q) m, m in [1/2, 4); u ~ m^(-1/3) from a degree-5
//    weighted least-squares fit (max rel. error 3.0e-3) and three
//    division-free Newton steps u <- u + u (1 - m u^3)/3; then y = m u^2
//    and one Newton step on y, y <- y + (m - y^3) u^2/3, with m - y^3
//    formed in double-double (the fma error of y^2).
//  * sin/cos: Cody-Waite reduction by pi/2 with a three-part pi/2 (valid for
//    |x| < 2^20 pi/2), Taylor polynomials on |r| <= pi/4.
//  * atan: reduction to |t| <= tan(pi/12) with atan(x) = pi/6 + atan((x -
//    1/sqrt3)/(1 + x/sqrt3)) and atan(x) = pi/2 - atan(1/x), then the
//    Taylor series of atan.  atan2 and acos are built from atan and sqrt.
//
// Accuracy against correctly rounded results is a few ulp; it is checked in
// tests/test_pm_math.py.  These routines are not a replacement for a full
// libm: arguments are the finite, moderate ones the droplet physics
// produces; NaN, infinities and zero are handled, huge trig arguments are
// not reduced exactly.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PM_HD __host__ __device__ inline
// Functions that take pointers to per-droplet register arrays must be
// inlined: an out-of-line call forces those arrays into scratch memory.
#define PM_HDI __host__ __device__ __attribute__((always_inline)) inline
#else
#define PM_HD inline
#define PM_HDI inline
#endif

namespace pm {

constexpr double kInf = __builtin_huge_val();
constexpr double kPi = 0x1.921fb54442d18p+1;
constexpr double kPiHalf = 0x1.921fb54442d18p+0;
// ln2 split: high part has 11 trailing zero bits so k*LN2_HI is exact for
// |k| < 2^11.
constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;
constexpr double kLn2Lo = 0x1.ef35793c76730p-45;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;

PM_HD double exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return kInf;
  if (x < -745.1332191019412) return 0.0;
  const double k = std::rint(x * kInvLn2);
  double r = std::fma(-k, kLn2Hi, x);
  r = std::fma(-k, kLn2Lo, r);
  // exp(r) = 1 + r s(r) on |r| <= ln2/2, s of degree 10 from a Chebyshev
  // fit in 50-digit arithmetic (|error of s| < 1.9e-17)
  double p = 0x1.af4ecd2129492p-26;
  p = std::fma(p, r, 0x1.28a38365c4751p-22);
  p = std::fma(p, r, 0x1.71de01c3fc936p-19);
  p = std::fma(p, r, 0x1.a019a5aef5532p-16);
  p = std::fma(p, r, 0x1.a01a01ac0a9eep-13);
  p = std::fma(p, r, 0x1.6c16c17f7ee1dp-10);
  p = std::fma(p, r, 0x1.111111111009cp-7);
  p = std::fma(p, r, 0x1.5555555551fd6p-5);
  p = std::fma(p, r, 0x1.5555555555557p-3);
  p = std::fma(p, r, 0x1.0000000000005p-1);
  p = std::fma(p, r, 1.0);
  p = std::fma(p, r, 1.0);
  return std::ldexp(p, static_cast<int>(k));
}

PM_HD double log(double x) {
  if (!(x > 0.0)) {
    if (x == 0.0) return -kInf;
    return __builtin_nan("");  // negative or NaN
  }
  if (x == kInf) return x;
  int e = 0;
  double m = std::frexp(x, &e);  // m in [1/2, 1), exact for subnormals too
  if (m < kSqrtHalf) {
    m = m + m;
    e -= 1;
  }
  const double f = m - 1.0;  // exact (Sterbenz)
  const double s = f / (2.0 + f);
  const double z = s * s;
  // R(z) = z Q(z), Q(z) = sum_{i>=1} 2/(2i+1) z^(i-1) on 0 <= z <=
  // (3 - 2 sqrt2)^2, here of degree 6 from a Chebyshev fit in 50-digit
  // arithmetic (|error of Q| < 3.2e-16)
  double R = 0x1.2b5f68a50d903p-3;
  R = std::fma(R, z, 0x1.39fdcceb4bb45p-3);
  R = std::fma(R, z, 0x1.7462b91b8df65p-3);
  R = std::fma(R, z, 0x1.c71c62d5e53e0p-3);
  R = std::fma(R, z, 0x1.2492492dfd86cp-2);
  R = std::fma(R, z, 0x1.9999999995273p-2);
  R = std::fma(R, z, 0x1.5555555555558p-1);
  R = R * z;
  const double l1p = f - s * (f - R);
  const double de = static_cast<double>(e);
  return de * kLn2Hi + (de * kLn2Lo + l1p);
}

PM_HD double log1p(double x) {
  if (!(x == x)) return x;
  if (x == kInf) return x;
  const double u = 1.0 + x;
  if (u == 1.0) return x;
  return pm::log(u) * (x / (u - 1.0));
}

// x^y for x >= 0 (negative bases are not used by the droplet physics and
// return NaN unless y == 0).
PM_HD double pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (!(x == x) || !(y == y)) return x + y;
  if (x == 1.0) return 1.0;
  if (x == 0.0) return (y > 0.0) ? 0.0 : kInf;
  if (x < 0.0) return __builtin_nan("");
  if (x == kInf) return (y > 0.0) ? kInf : 0.0;
  return pm::exp(y * pm::log(x));
}

PM_HD double cbrt(double x) {
  if (x == 0.0 || !(x == x) || x == kInf || x == -kInf) return x;
  const double ax = std::fabs(x);
  int e = 0;
  double m = std::frexp(ax, &e);  // ax = m 2^e, m in [1/2,1)
  // e = 3q + r with r in {0,1,2}
  int q = (e >= 0) ? (e / 3) : -((-e + 2) / 3);
  int r = e - 3 * q;
  m = std::ldexp(m, r);  // m in [1/2, 4)
  // u ~ m^(-1/3): a degree-5 fit on [1/2, 4) (relative error < 2^-8.3),
  // then three Newton steps u <- u + u (1 - m u^3) / 3, each squaring the
  // relative error, with no division
  constexpr double kThird = 0x1.5555555555555p-2;
  double u = -0x1.7004e55df38c6p-8;
  u = std::fma(u, m, 0x1.30ee0814c4b64p-4);
  u = std::fma(u, m, -0x1.8e381fe895553p-2);
  u = std::fma(u, m, 0x1.06cf19aa0aa15p+0);
  u = std::fma(u, m, -0x1.80bf35f68a8d4p+0);
  u = std::fma(u, m, 0x1.cb88b5e0994ebp+0);
  for (int it = 0; it < 3; ++it) {
    const double u2 = u * u;
    const double res = std::fma(-(m * u), u2, 1.0);  // 1 - m u^3
    u = std::fma(u * res, kThird, u);
  }
  // y = m u^2 ~ m^(1/3), then one Newton step on y with the residual
  // m - y^3 formed to twice the working precision (the exact error of y^2
  // from an fma) and 1 / (3 y^2) ~ u^2 / 3
  const double uu = u * u;
  double y = m * uu;
  const double y2 = y * y;
  const double y2e = std::fma(y, y, -y2);
  const double d = std::fma(-y2e, y, std::fma(-y2, y, m));
  y = std::fma(d * uu, kThird, y);
  y = std::ldexp(y, q);
  return (x < 0.0) ? -y : y;
}

// pi/2 in three parts; the first two have 33 trailing zero bits so that
// k*P1 and k*P2 are exact for |k| < 2^20.
constexpr double kPio2_1 = 0x1.921fa00000000p+0;
constexpr double kPio2_2 = 0x1.5444200000000p-20;
constexpr double kPio2_3 = 0x1.a308d313198a3p-41;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;

// sin(r), cos(r) on |r| <= pi/4 (Taylor, truncation < 2e-17)
PM_HD double sin_kernel(double r) {
  // sum_{n=0..9} (-1)^n r^(2n+1)/(2n+1)!, Horner in z = r^2
  const double z = r * r;
  double p = -0x1.2f49b46814157p-57;           // -1/19!
  p = std::fma(p, z, 0x1.952c77030ad4ap-49);   // +1/17!
  p = std::fma(p, z, -0x1.ae7f3e733b81fp-41);  // -1/15!
  p = std::fma(p, z, 0x1.6124613a86d09p-33);   // +1/13!
  p = std::fma(p, z, -0x1.ae64567f544e4p-26);  // -1/11!
  p = std::fma(p, z, 0x1.71de3a556c734p-19);   // +1/9!
  p = std::fma(p, z, -0x1.a01a01a01a01ap-13);  // -1/7!
  p = std::fma(p, z, 0x1.1111111111111p-7);    // +1/5!
  p = std::fma(p, z, -0x1.5555555555555p-3);   // -1/3!
  return std::fma(p * z, r, r);
}

PM_HD double cos_kernel(double r) {
  // sum_{n=0..9} (-1)^n r^(2n)/(2n)!
  const double z = r * r;
  double p = -0x1.6827863b97d97p-53;           // -1/18!
  p = std::fma(p, z, 0x1.ae7f3e733b81fp-45);   // +1/16!
  p = std::fma(p, z, -0x1.93974a8c07c9dp-37);  // -1/14!
  p = std::fma(p, z, 0x1.1eed8eff8d898p-29);   // +1/12!
  p = std::fma(p, z, -0x1.27e4fb7789f5cp-22);  // -1/10!
  p = std::fma(p, z, 0x1.a01a01a01a01ap-16);   // +1/8!
  p = std::fma(p, z, -0x1.6c16c16c16c17p-10);  // -1/6!
  p = std::fma(p, z, 0x1.5555555555555p-5);    // +1/4!
  p = std::fma(p, z, -0.5);                    // -1/2!
  return std::fma(p, z, 1.0);
}

// Reduce x to r in [-pi/4, pi/4] and quadrant n (mod 4).
PM_HD double reduce_pio2(double x, int& n) {
  const double k = std::rint(x * kTwoOverPi);
  double r = std::fma(-k, kPio2_1, x);
  r = std::fma(-k, kPio2_2, r);
  r = std::fma(-k, kPio2_3, r);
  const long long kl = static_cast<long long>(k);
  n = static_cast<int>(kl & 3);
  return r;
}

PM_HD double sin(double x) {
  if (!(x == x) || x == kInf || x == -kInf) return __builtin_nan("");
  int n = 0;
  const double r = reduce_pio2(x, n);
  switch (n) {
    case 0: return sin_kernel(r);
    case 1: return cos_kernel(r);
    case 2: return -sin_kernel(r);
    default: return -cos_kernel(r);
  }
}

PM_HD double cos(double x) {
  if (!(x == x) || x == kInf || x == -kInf) return __builtin_nan("");
  int n = 0;
  const double r = reduce_pio2(x, n);
  switch (n) {
    case 0: return cos_kernel(r);
    case 1: return -sin_kernel(r);
    case 2: return -cos_kernel(r);
    default: return sin_kernel(r);
  }
}

// sin(x) and cos(x) from one reduction: the same values as sin and cos
PM_HD void sincos(double x, double& s, double& c) {
  if (!(x == x) || x == kInf || x == -kInf) {
    s = c = __builtin_nan("");
    return;
  }
  int n = 0;
  const double r = reduce_pio2(x, n);
  const double sk = sin_kernel(r), ck = cos_kernel(r);
  switch (n) {
    case 0: s = sk; c = ck; break;
    case 1: s = ck; c = -sk; break;
    case 2: s = -sk; c = -ck; break;
    default: s = -ck; c = sk; break;
  }
}

constexpr double kPiSixth = 0x1.0c152382d7366p-1;
constexpr double kInvSqrt3 = 0x1.279a74590331cp-1;
constexpr double kTanPi12 = 0x1.126145e9ecd56p-2;  // 2 - sqrt(3)

// atan(t) for |t| <= tan(pi/12) ~ 0.268: Taylor to t^39 (truncation < 1e-24)
PM_HD double atan_kernel(double t) {
  const double z = t * t;
  double p = -1.0 / 39.0;
  p = std::fma(p, z, 1.0 / 37.0);
  p = std::fma(p, z, -1.0 / 35.0);
  p = std::fma(p, z, 1.0 / 33.0);
  p = std::fma(p, z, -1.0 / 31.0);
  p = std::fma(p, z, 1.0 / 29.0);
  p = std::fma(p, z, -1.0 / 27.0);
  p = std::fma(p, z, 1.0 / 25.0);
  p = std::fma(p, z, -1.0 / 23.0);
  p = std::fma(p, z, 1.0 / 21.0);
  p = std::fma(p, z, -1.0 / 19.0);
  p = std::fma(p, z, 1.0 / 17.0);
  p = std::fma(p, z, -1.0 / 15.0);
  p = std::fma(p, z, 1.0 / 13.0);
  p = std::fma(p, z, -1.0 / 11.0);
  p = std::fma(p, z, 1.0 / 9.0);
  p = std::fma(p, z, -1.0 / 7.0);
  p = std::fma(p, z, 1.0 / 5.0);
  p = std::fma(p, z, -1.0 / 3.0);
  return std::fma(p * z, t, t);
}

PM_HD double atan(double x) {
  if (!(x == x)) return x;
  const bool neg = x < 0.0;
  double a = std::fabs(x);
  bool inv = false;
  if (a > 1.0) {
    a = 1.0 / a;
    inv = true;
  }
  double res;
  if (a > kTanPi12) {
    const double t = (a - kInvSqrt3) / std::fma(a, kInvSqrt3, 1.0);
    res = kPiSixth + atan_kernel(t);
  } else {
    res = atan_kernel(a);
  }
  if (inv) res = kPiHalf - res;
  return neg ? -res : res;
}

PM_HD double atan2(double y, double x) {
  if (!(x == x) || !(y == y)) return x + y;
  if (x == 0.0 && y == 0.0) return (std::signbit(x)) ? (std::signbit(y) ? -kPi : kPi) : y;
  if (x == 0.0) return (y > 0.0) ? kPiHalf : -kPiHalf;
  double r;
  if (std::fabs(y) <= std::fabs(x)) {
    r = pm::atan(y / x);
    if (x < 0.0) r = (y >= 0.0 && !std::signbit(y)) ? r + kPi : r - kPi;
  } else {
    r = ((y > 0.0) ? kPiHalf : -kPiHalf) - pm::atan(x / y);
  }
  return r;
}

// acos(x) = 2 atan( sqrt((1-x)/(1+x)) ), written with sqrt(1-x)/sqrt(1+x)
PM_HD double acos(double x) {
  if (!(x >= -1.0 && x <= 1.0)) return __builtin_nan("");
  if (x == -1.0) return kPi;
  return 2.0 * pm::atan(std::sqrt(1.0 - x) / std::sqrt(1.0 + x));
}

// asin(x) = atan( x / sqrt((1-x)(1+x)) ), exact +-pi/2 at |x| = 1
PM_HD double asin(double x) {
  if (!(x >= -1.0 && x <= 1.0)) return __builtin_nan("");
  if (x == 1.0) return kPiHalf;
  if (x == -1.0) return -kPiHalf;
  return pm::atan(x / std::sqrt((1.0 - x) * (1.0 + x)));
}

// tan(x) = sin(x) / cos(x)
PM_HD double tan(double x) { return pm::sin(x) / pm::cos(x); }

}  // namespace pm
