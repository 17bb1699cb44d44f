// Deterministic arithmetic shared by the gfx950 kernels and the CPU replay of
// the sampler.  Every function here is built only from IEEE-754 operations
// that are correctly rounded on both the host (x86-64, glibc) and gfx950
// (add, mul, fma, div, rint, ldexp, fabs), with floating-point contraction
// disabled, so a host build and a device build return bit-identical results.
// That is what makes Metropolis accept/reject decisions replayable on the CPU.
//
// Methods (all published, none derived from a vendor math library):
//   * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//     easy as 1, 2, 3", SC'11.  Multipliers 0xD2511F53/0xCD9E8D57, Weyl key
//     increments 0x9E3779B9/0xBB67AE85, ten rounds.
//   * exp: Cody & Waite two-constant argument reduction x = k ln2 + r,
//     |r| <= ln2/2, then the Taylor polynomial of e^r evaluated by Horner
//     (degree 13 in double, 7 in float), then ldexp.
//   * log1p on [0,1]: one halving step to f in [sqrt(1/2)-1, sqrt(2)-1],
//     then log1p(f) = 2 atanh(s), s = f/(2+f), as an odd series in s.
//   * logcosh: log cosh x = |x| - ln2 + log1p(exp(-2|x|)) for |x| <= 12 and
//     |x| - ln2 beyond, the same branch point as Utilities/Utility.hpp:15-24.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define YQ_HD __host__ __device__
#else
#define YQ_HD
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace yq {

// ---------------------------------------------------------------- Philox4x32-10
struct u32x4 {
  uint32_t x, y, z, w;
};

YQ_HD inline u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__clang__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * ctr.x;
    const uint64_t p1 = (uint64_t)M1 * ctr.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    u32x4 n;
    n.x = hi1 ^ ctr.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ ctr.w ^ k1;
    n.w = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// Counter layout of every draw in this library: (lo, hi) of a 64-bit event
// index, the chain (or other stream) id, and a domain tag that separates
// independent uses of one seed.
enum Domain : uint32_t {
  kDomSweep = 0u,      // Metropolis proposals
  kDomRandomize = 1u,  // initial spin configurations
  kDomParams = 2u,     // parameter initialisation
  kDomUniform = 3u,    // uniform configurations for batch inputs
  kDomExact = 4u,      // draws of the exact (inverse-CDF) sampler
  kDomMix = 5u,        // replica swaps of parallel tempering
};

YQ_HD inline u32x4 draw(uint64_t seed, uint64_t event, uint32_t stream,
                        uint32_t domain) {
  u32x4 c;
  c.x = (uint32_t)event;
  c.y = (uint32_t)(event >> 32);
  c.z = stream;
  c.w = domain;
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// floor(x * n / 2^32): an index in [0, n) from one 32-bit word.
YQ_HD inline uint32_t index_from(uint32_t x, uint32_t n) {
  return (uint32_t)(((uint64_t)x * (uint64_t)n) >> 32);
}
// 53-bit uniform double in [0, 1) from two words.
YQ_HD inline double u01_double(uint32_t a, uint32_t b) {
  const uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)v * (1.0 / 9007199254740992.0);
}
// 24-bit uniform float in [0, 1) from one word.
YQ_HD inline float u01_float(uint32_t a) {
  return (float)(a >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------- constants
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLog2e = 1.4426950408889634074;
constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // 0x3dea39ef35793c76
constexpr double kSqrt2m1 = 0.41421356237309504880;

constexpr float kLn2f = 0.693147180559945f;
constexpr float kLog2ef = 1.44269504088896f;
constexpr float kLn2Hif = 0.693145751953125f;     // 0x3f317200 (low 12 bits 0)
constexpr float kLn2Lof = 1.42860682030941723e-6f;  // ln2 - kLn2Hif
constexpr float kSqrt2m1f = 0.414213562373095f;

// ---------------------------------------------------------------- exp
YQ_HD inline double exp_d(double x) {
  if (!(x > -708.0)) return (x != x) ? x : 0.0;  // NaN passes, no subnormals
  if (x > 709.0) return HUGE_VAL;
  const double k = rint(x * kLog2e);
  double r = fma(-k, kLn2Hi, x);
  r = fma(-k, kLn2Lo, r);
  double p = 1.0 / 6227020800.0;       // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);    // 1/12!
  p = fma(p, r, 1.0 / 39916800.0);     // 1/11!
  p = fma(p, r, 1.0 / 3628800.0);      // 1/10!
  p = fma(p, r, 1.0 / 362880.0);       // 1/9!
  p = fma(p, r, 1.0 / 40320.0);        // 1/8!
  p = fma(p, r, 1.0 / 5040.0);         // 1/7!
  p = fma(p, r, 1.0 / 720.0);          // 1/6!
  p = fma(p, r, 1.0 / 120.0);          // 1/5!
  p = fma(p, r, 1.0 / 24.0);           // 1/4!
  p = fma(p, r, 1.0 / 6.0);            // 1/3!
  p = fma(p, r, 0.5);                  // 1/2!
  p = fma(p, r, 1.0);                  // 1/1!
  p = fma(p, r, 1.0);                  // 1/0!
  return ldexp(p, (int)k);
}

YQ_HD inline float exp_f(float x) {
  if (!(x > -87.0f)) return (x != x) ? x : 0.0f;
  if (x > 88.0f) return HUGE_VALF;
  const float k = rintf(x * kLog2ef);
  float r = fmaf(-k, kLn2Hif, x);
  r = fmaf(-k, kLn2Lof, r);
  float p = 1.0f / 5040.0f;            // 1/7!
  p = fmaf(p, r, 1.0f / 720.0f);
  p = fmaf(p, r, 1.0f / 120.0f);
  p = fmaf(p, r, 1.0f / 24.0f);
  p = fmaf(p, r, 1.0f / 6.0f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// ---------------------------------------------------------------- log1p on [0,1]
YQ_HD inline double log1p01_d(double y) {
  double f, off;
  if (y > kSqrt2m1) {
    f = (y - 1.0) * 0.5;  // log(1+y) = ln2 + log1p((y-1)/2)
    off = kLn2;
  } else {
    f = y;
    off = 0.0;
  }
  const double s = f / (2.0 + f);
  const double z = s * s;
  double q = 1.0 / 21.0;
  q = fma(q, z, 1.0 / 19.0);
  q = fma(q, z, 1.0 / 17.0);
  q = fma(q, z, 1.0 / 15.0);
  q = fma(q, z, 1.0 / 13.0);
  q = fma(q, z, 1.0 / 11.0);
  q = fma(q, z, 1.0 / 9.0);
  q = fma(q, z, 1.0 / 7.0);
  q = fma(q, z, 1.0 / 5.0);
  q = fma(q, z, 1.0 / 3.0);
  const double t = 2.0 * s;
  return off + fma(t, z * q, t);
}

YQ_HD inline float log1p01_f(float y) {
  float f, off;
  if (y > kSqrt2m1f) {
    f = (y - 1.0f) * 0.5f;
    off = kLn2f;
  } else {
    f = y;
    off = 0.0f;
  }
  const float s = f / (2.0f + f);
  const float z = s * s;
  float q = 1.0f / 9.0f;
  q = fmaf(q, z, 1.0f / 7.0f);
  q = fmaf(q, z, 1.0f / 5.0f);
  q = fmaf(q, z, 1.0f / 3.0f);
  const float t = 2.0f * s;
  return off + fmaf(t, z * q, t);
}

// ---------------------------------------------------------------- logcosh
YQ_HD inline double logcosh_d(double x) {
  const double xp = fabs(x);
  if (xp <= 12.0) return (xp - kLn2) + log1p01_d(exp_d(-2.0 * xp));
  return xp - kLn2;
}

YQ_HD inline float logcosh_f(float x) {
  const float xp = fabsf(x);
  if (xp <= 12.0f) return (xp - kLn2f) + log1p01_f(exp_f(-2.0f * xp));
  return xp - kLn2f;
}

// ---------------------------------------------------------------- complex
// sin and cos: x = k pi/2 + r with a three-part Cody-Waite split of pi/2 (the
// fdlibm constants pio2_1, pio2_2, pio2_3), |r| <= pi/4, Taylor polynomials
// to r^17 / r^18, quadrant from k mod 4.  Accurate for |x| < 2^20.
YQ_HD inline void sincos_d(double x, double* s, double* c) {
  const double k = rint(x * 0.63661977236758134308);  // 2/pi
  double r = fma(-k, 1.57079632673412561417e+00, x);
  r = fma(-k, 6.07710050630396597660e-11, r);
  r = fma(-k, 2.02226624871116645580e-21, r);
  const double z = r * r;
  double ps = 1.0 / 355687428096000.0;  // 1/17!
  ps = fma(ps, z, -1.0 / 1307674368000.0);
  ps = fma(ps, z, 1.0 / 6227020800.0);
  ps = fma(ps, z, -1.0 / 39916800.0);
  ps = fma(ps, z, 1.0 / 362880.0);
  ps = fma(ps, z, -1.0 / 5040.0);
  ps = fma(ps, z, 1.0 / 120.0);
  ps = fma(ps, z, -1.0 / 6.0);
  const double sr = fma(r * z, ps, r);
  double pc = 1.0 / 6402373705728000.0;  // 1/18!
  pc = fma(pc, z, -1.0 / 20922789888000.0);
  pc = fma(pc, z, 1.0 / 87178291200.0);
  pc = fma(pc, z, -1.0 / 479001600.0);
  pc = fma(pc, z, 1.0 / 3628800.0);
  pc = fma(pc, z, -1.0 / 40320.0);
  pc = fma(pc, z, 1.0 / 720.0);
  pc = fma(pc, z, -1.0 / 24.0);
  pc = fma(pc, z, 0.5);
  const double cr = fma(-z, pc, 1.0);
  const int q = ((int)(int64_t)k) & 3;
  if (q == 0) { *s = sr; *c = cr; }
  else if (q == 1) { *s = cr; *c = -sr; }
  else if (q == 2) { *s = -sr; *c = -cr; }
  else { *s = -cr; *c = sr; }
}

YQ_HD inline void sincos_f(float x, float* s, float* c) {
  const float k = rintf(x * 0.636619772f);
  float r = fmaf(-k, 1.5703125f, x);                // pi/2 in three float parts
  r = fmaf(-k, 4.8382679233327508e-4f, r);
  r = fmaf(-k, 2.5633440682570896e-12f, r);
  const float z = r * r;
  float ps = 1.0f / 362880.0f;  // to r^9: the r^11 term is below 2e-9 relative
  ps = fmaf(ps, z, -1.0f / 5040.0f);
  ps = fmaf(ps, z, 1.0f / 120.0f);
  ps = fmaf(ps, z, -1.0f / 6.0f);
  const float sr = fmaf(r * z, ps, r);
  float pc = 1.0f / 40320.0f;
  pc = fmaf(pc, z, -1.0f / 720.0f);
  pc = fmaf(pc, z, 1.0f / 24.0f);
  pc = fmaf(pc, z, -0.5f);
  const float cr = fmaf(z, pc, 1.0f);
  const int q = ((int)k) & 3;
  if (q == 0) { *s = sr; *c = cr; }
  else if (q == 1) { *s = cr; *c = -sr; }
  else if (q == 2) { *s = -sr; *c = -cr; }
  else { *s = -cr; *c = sr; }
}

// log1p(f) + e ln2 for f in [sqrt(1/2) - 1, sqrt(2) - 1]: 2 atanh(s),
// s = f / (2 + f), as an odd series in s.
YQ_HD inline double log1p_red_d(double f, double e) {
  const double s = f / (2.0 + f);
  const double z = s * s;
  double q = 1.0 / 21.0;
  q = fma(q, z, 1.0 / 19.0);
  q = fma(q, z, 1.0 / 17.0);
  q = fma(q, z, 1.0 / 15.0);
  q = fma(q, z, 1.0 / 13.0);
  q = fma(q, z, 1.0 / 11.0);
  q = fma(q, z, 1.0 / 9.0);
  q = fma(q, z, 1.0 / 7.0);
  q = fma(q, z, 1.0 / 5.0);
  q = fma(q, z, 1.0 / 3.0);
  const double t = 2.0 * s;
  return fma(e, kLn2Hi, fma(e, kLn2Lo, fma(t, z * q, t)));
}
YQ_HD inline float log1p_red_f(float f, float e) {
  const float s = f / (2.0f + f);
  const float z = s * s;
  float q = 1.0f / 9.0f;
  q = fmaf(q, z, 1.0f / 7.0f);
  q = fmaf(q, z, 1.0f / 5.0f);
  q = fmaf(q, z, 1.0f / 3.0f);
  const float t = 2.0f * s;
  return fmaf(e, kLn2f, fmaf(t, z * q, t));
}

// log v for v > 0: v = f 2^e with f in [sqrt(1/2), sqrt(2)) (f - 1 exact).
YQ_HD inline double log_pos_d(double v) {
  int ex;
  double mnt = frexp(v, &ex);  // [0.5, 1)
  if (mnt < 0.70710678118654752440) {
    mnt = mnt * 2.0;
    ex -= 1;
  }
  return log1p_red_d(mnt - 1.0, (double)ex);
}
YQ_HD inline float log_pos_f(float v) {
  int ex;
  float mnt = frexpf(v, &ex);
  if (mnt < 0.707106781f) {
    mnt = mnt * 2.0f;
    ex -= 1;
  }
  return log1p_red_f(mnt - 1.0f, (float)ex);
}

// log(1 + w) for w in (-1, inf): the series directly when |w| <= 1/4, else
// the reduction of 1 + w as in log_pos.
YQ_HD inline double log1p_any_d(double w) {
  if (!(w > -1.0)) return w == -1.0 ? -HUGE_VAL : (w - w) / (w - w);
  if (fabs(w) <= 0.25) return log1p_red_d(w, 0.0);
  return log_pos_d(1.0 + w);
}
YQ_HD inline float log1p_any_f(float w) {
  if (!(w > -1.0f)) return w == -1.0f ? -HUGE_VALF : (w - w) / (w - w);
  if (fabsf(w) <= 0.25f) return log1p_red_f(w, 0.0f);
  return log_pos_f(1.0f + w);
}

// expm1 for the tanh of small arguments: the exp polynomial without its
// constant term when |x| < ln2/2, exp(x) - 1 (no cancellation) otherwise.
YQ_HD inline double expm1_d(double x) {
  if (!(fabs(x) < 0.34657359027997264)) return exp_d(x) - 1.0;
  double p = 1.0 / 87178291200.0;  // 1/14!
  p = fma(p, x, 1.0 / 6227020800.0);
  p = fma(p, x, 1.0 / 479001600.0);
  p = fma(p, x, 1.0 / 39916800.0);
  p = fma(p, x, 1.0 / 3628800.0);
  p = fma(p, x, 1.0 / 362880.0);
  p = fma(p, x, 1.0 / 40320.0);
  p = fma(p, x, 1.0 / 5040.0);
  p = fma(p, x, 1.0 / 720.0);
  p = fma(p, x, 1.0 / 120.0);
  p = fma(p, x, 1.0 / 24.0);
  p = fma(p, x, 1.0 / 6.0);
  p = fma(p, x, 0.5);
  return fma(p * x, x, x);
}
YQ_HD inline float expm1_f(float x) {
  if (!(fabsf(x) < 0.34657359f)) return exp_f(x) - 1.0f;
  float p = 1.0f / 40320.0f;  // 1/8!
  p = fmaf(p, x, 1.0f / 5040.0f);
  p = fmaf(p, x, 1.0f / 720.0f);
  p = fmaf(p, x, 1.0f / 120.0f);
  p = fmaf(p, x, 1.0f / 24.0f);
  p = fmaf(p, x, 1.0f / 6.0f);
  p = fmaf(p, x, 0.5f);
  return fmaf(p * x, x, x);
}

// Re log cosh(x + iy) = logcosh(x) + (1/2) log q with
// q = |cosh(x+iy)|^2 / cosh^2 x = 1 - sin^2 y sech^2 x = tanh^2 x + cos^2 y sech^2 x
// (the real part of Utility.hpp:28-37).  With w = -sin^2 y sech^2 x, log q is
// log1p(w) while w > -1/2 and log of the positive sum otherwise (near the
// zeros of cosh, where 1 + w would cancel); sech^2 x = 4e/(1+e)^2 with
// e = exp(-2|x|), tanh|x| = -u/(2+u) with u = expm1(-2|x|).
YQ_HD inline double logcosh_re_d(double x, double y) {
  double sy, cy;
  sincos_d(y, &sy, &cy);
  const double e = exp_d(-2.0 * fabs(x));
  const double ic2 = 4.0 * e / ((1.0 + e) * (1.0 + e));
  const double w = -(sy * sy) * ic2;
  double lq;
  if (w > -0.5) {
    lq = log1p_any_d(w);
  } else {
    const double u = expm1_d(-2.0 * fabs(x));
    const double t = -u / (2.0 + u);
    lq = log_pos_d(t * t + (cy * cy) * ic2);
  }
  return logcosh_d(x) + 0.5 * lq;
}
YQ_HD inline float logcosh_re_f(float x, float y) {
  float sy, cy;
  sincos_f(y, &sy, &cy);
  const float e = exp_f(-2.0f * fabsf(x));
  const float ic2 = 4.0f * e / ((1.0f + e) * (1.0f + e));
  const float w = -(sy * sy) * ic2;
  float lq;
  if (w > -0.5f) {
    lq = log1p_any_f(w);
  } else {
    const float u = expm1_f(-2.0f * fabsf(x));
    const float t = -u / (2.0f + u);
    lq = log_pos_f(t * t + (cy * cy) * ic2);
  }
  return logcosh_f(x) + 0.5f * lq;
}

// Both parts of log cosh(x + iy): the real part as logcosh_re (bit for bit)
// and the principal argument atan2(tanh x sin y, cos y).  The argument uses
// the platform tanh and atan2 and is not replayed bit for bit.
YQ_HD inline void logcosh_c_d(double x, double y, double* re, double* im) {
  double sy, cy;
  sincos_d(y, &sy, &cy);
  *re = logcosh_re_d(x, y);
  *im = atan2(tanh(x) * sy, cy);
}
YQ_HD inline void logcosh_c_f(float x, float y, float* re, float* im) {
  float sy, cy;
  sincos_f(y, &sy, &cy);
  *re = logcosh_re_f(x, y);
  *im = atan2f(tanhf(x) * sy, cy);
}

// tanh(x + iy) = (sinh 2x + i sin 2y) / (cosh 2x + cos 2y), with numerator and
// denominator scaled by 2 exp(-2|x|): with q = expm1(-2|x|) (so e = 1 + q),
// re = sgn(x) (-q (2 + q)) / d, im = 4 e sin y cos y / d and
// d = 1 + e^2 + 2 e cos 2y written as q^2 + 4 e cos^2 y (no cancellation near
// the poles x = 0, cos y = 0).
YQ_HD inline void tanh_c_d(double x, double y, double* re, double* im) {
  double sy, cy;
  sincos_d(y, &sy, &cy);
  const double q = expm1_d(-2.0 * fabs(x));
  const double e = 1.0 + q;
  const double d = q * q + 4.0 * e * (cy * cy);
  *re = copysign(-q * (2.0 + q), x) / d;
  *im = 4.0 * e * (sy * cy) / d;
}

template <typename R>
struct Math;
template <>
struct Math<double> {
  YQ_HD static double logcosh(double x) { return logcosh_d(x); }
  YQ_HD static double exp(double x) { return exp_d(x); }
};
template <>
struct Math<float> {
  YQ_HD static float logcosh(float x) { return logcosh_f(x); }
  YQ_HD static float exp(float x) { return exp_f(x); }
};

// Metropolis acceptance of Samplers/LocalSweeper.hpp:44-46 with the
// replayable exp: p = min(1, exp(beta*2*delta)); accept iff u < p.
YQ_HD inline bool accept_d(double beta, double delta, double u) {
  const double lp = beta * 2.0 * delta;
  if (lp >= 0.0) return true;  // p == 1 and u < 1 always
  return u < exp_d(lp);
}
YQ_HD inline bool accept_f(float beta, float delta, float u) {
  const float lp = beta * 2.0f * delta;
  if (lp >= 0.0f) return true;
  return u < exp_f(lp);
}

}  // namespace yq
