// Deterministic f64 transcendentals and the Philox4x32-10 stream, usable on
// host and device.  Every function here is evaluated with IEEE-754 +, -, *, /
// and sqrt only (no fused multiply-add: the library is compiled with
// -ffp-contract=off), in a fixed order, so the host build, the gfx950 build and
// a NumPy evaluation of the same formulas return identical bits.  That is what
// makes collision masks, SEE spawns and injected particle counts reproducible
// between the GPU and a CPU check.
//
// exp : Cody-Waite reduction x = k*ln2 + r (ln2_hi has 32 significant bits so
//       k*ln2_hi is exact), Taylor polynomial of degree 13 on |r| <= ln2/2,
//       scale by 2^k in two exact power-of-two multiplications.
// log : x = m*2^e, m in [sqrt(1/2), sqrt(2)); log m = 2*atanh(s),
//       s = (m-1)/(m+1), odd series to s^21.
// sincos(2*pi*u), u in [0,1): exact quadrant split of 4u, Taylor on [0,pi/4].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BEMCS_HD __host__ __device__ __forceinline__

namespace bemcs {

constexpr double kInvLn2 = 1.4426950408889634;
constexpr double kLn2Hi = 0.6931471803691238;      // 0x1.62e42fee00000p-1
constexpr double kLn2Lo = 1.9082149292705877e-10;  // ln2 - kLn2Hi
constexpr double kSqrtHalf = 0.7071067811865476;
constexpr double kHalfPi = 1.5707963267948966;
constexpr double kInvLn10 = 0.43429448190325176;
constexpr double kLn10 = 2.302585092994045684;

BEMCS_HD double pow2i(int k) {  // 2^k for -1022 <= k <= 1023, exact
  uint64_t bits = (uint64_t)(k + 1023) << 52;
  double d;
  __builtin_memcpy(&d, &bits, 8);
  return d;
}

BEMCS_HD double det_exp(double x) {
  if (x != x) return x;
  double xc = x < -1100.0 ? -1100.0 : (x > 1100.0 ? 1100.0 : x);
  double k = __builtin_rint(xc * kInvLn2);
  double r = (xc - k * kLn2Hi) - k * kLn2Lo;
  double p = 1.6059043836821613e-10;
  p = p * r + 2.08767569878681e-09;
  p = p * r + 2.505210838544172e-08;
  p = p * r + 2.755731922398589e-07;
  p = p * r + 2.7557319223985893e-06;
  p = p * r + 2.48015873015873e-05;
  p = p * r + 0.0001984126984126984;
  p = p * r + 0.001388888888888889;
  p = p * r + 0.008333333333333333;
  p = p * r + 0.041666666666666664;
  p = p * r + 0.16666666666666666;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  int ki = (int)k;
  int a = ki >> 1;  // floor division, as NumPy's >> on int64
  int b = ki - a;
  return (p * pow2i(a)) * pow2i(b);
}

BEMCS_HD double det_log(double x) {
  if (x != x || x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  if (x == __builtin_inf()) return x;
  int e;
  double m = __builtin_frexp(x, &e);
  if (m < kSqrtHalf) {
    m = m * 2.0;
    e = e - 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double R = 0.047619047619047616;
  R = R * z + 0.05263157894736842;
  R = R * z + 0.058823529411764705;
  R = R * z + 0.06666666666666667;
  R = R * z + 0.07692307692307693;
  R = R * z + 0.09090909090909091;
  R = R * z + 0.1111111111111111;
  R = R * z + 0.14285714285714285;
  R = R * z + 0.2;
  R = R * z + 0.3333333333333333;
  R = R * z;
  double two_s = 2.0 * s;
  double log_m = two_s + two_s * R;
  double ef = (double)e;
  return ef * kLn2Hi + (ef * kLn2Lo + log_m);
}

// sin and cos of 2*pi*u for u in [0, 1)
BEMCS_HD void det_sincos_2pi(double u, double* s_out, double* c_out) {
  double t = 4.0 * u;
  double q = __builtin_floor(t);
  double f = t - q;
  bool flip = f > 0.5;
  double g = flip ? 1.0 - f : f;
  double a = g * kHalfPi;
  double a2 = a * a;
  double P = 2.8114572543455206e-15;
  P = P * a2 + -7.647163731819816e-13;
  P = P * a2 + 1.6059043836821613e-10;
  P = P * a2 + -2.505210838544172e-08;
  P = P * a2 + 2.7557319223985893e-06;
  P = P * a2 + -0.0001984126984126984;
  P = P * a2 + 0.008333333333333333;
  P = P * a2 + -0.16666666666666666;
  double s = a + (a * a2) * P;
  double Q = 4.779477332387385e-14;
  Q = Q * a2 + -1.1470745597729725e-11;
  Q = Q * a2 + 2.08767569878681e-09;
  Q = Q * a2 + -2.755731922398589e-07;
  Q = Q * a2 + 2.48015873015873e-05;
  Q = Q * a2 + -0.001388888888888889;
  Q = Q * a2 + 0.041666666666666664;
  Q = Q * a2 + -0.5;
  double c = 1.0 + a2 * Q;
  if (flip) {
    double tmp = s;
    s = c;
    c = tmp;
  }
  int qi = ((int)q) & 3;
  if (qi == 0) { *s_out = s;  *c_out = c; }
  else if (qi == 1) { *s_out = c;  *c_out = -s; }
  else if (qi == 2) { *s_out = -s; *c_out = -c; }
  else { *s_out = -c; *c_out = s; }
}

// ---------------------------------------------------------------- Philox
struct u32x4 {
  uint32_t x, y, z, w;
};

BEMCS_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// purpose tags: the second key word of every stream
enum : uint32_t {
  P_INIT_ION = 0x10,
  P_INIT_ELEC = 0x11,
  P_HOLD_ION = 0x12,
  P_HOLD_ELEC = 0x13,
  P_INJ_COUNT = 0x20,
  P_INJ_ION = 0x21,
  P_NEUT = 0x22,
  P_RF_COUNT = 0x23,
  P_RF_ELEC = 0x24,
  P_SEE = 0x30,
  P_MCC = 0x40,
};

struct u01x2 {
  double a, b;
};

// two uniforms in [0,1) from block (id, step, blk, 0) of stream (seed, purpose)
BEMCS_HD u01x2 uniforms(uint32_t seed, uint32_t purpose, uint32_t id, uint32_t step,
                        uint32_t blk) {
  u32x4 c{id, step, blk, 0u};
  u32x4 r = philox4x32_10(c, seed, purpose);
  uint64_t a = ((uint64_t)r.y << 32) | r.x;
  uint64_t b = ((uint64_t)r.w << 32) | r.z;
  const double scale = 1.1102230246251565e-16;  // 2^-53
  u01x2 o;
  o.a = (double)(a >> 11) * scale;
  o.b = (double)(b >> 11) * scale;
  return o;
}

// Box-Muller: two standard normals from one block
BEMCS_HD u01x2 normals(uint32_t seed, uint32_t purpose, uint32_t id, uint32_t step,
                       uint32_t blk) {
  u01x2 u = uniforms(seed, purpose, id, step, blk);
  double r = __builtin_sqrt(-2.0 * det_log(1.0 - u.a));
  double s, c;
  det_sincos_2pi(u.b, &s, &c);
  u01x2 o;
  o.a = r * c;
  o.b = r * s;
  return o;
}

// Poisson(lam) as a sum of ceil(lam/16) Knuth product samples; uniform j is
// half (j & 1) of block (j >> 1, step, 0).
BEMCS_HD int64_t poisson_count(uint32_t seed, uint32_t purpose, uint32_t step, double lam) {
  if (!(lam > 0.0)) return 0;
  double nch = __builtin_ceil(lam / 16.0);
  int64_t nchunk = (int64_t)nch;
  double lam_c = lam / (double)nchunk;
  double L = det_exp(-lam_c);
  int64_t count = 0;
  uint64_t j = 0;
  for (int64_t ch = 0; ch < nchunk; ++ch) {
    double p = 1.0;
    for (;;) {
      u01x2 u = uniforms(seed, purpose, (uint32_t)(j >> 1), step, 0u);
      double v = (j & 1) ? u.b : u.a;
      j += 1;
      p = p * v;
      if (p <= L) break;
      count += 1;
    }
  }
  return count;
}

}  // namespace bemcs
