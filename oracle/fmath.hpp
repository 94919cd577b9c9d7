// ORACLE — test infrastructure only (see directory.hpp).
//
// Scalar math the restatement uses. For float, every function is built from IEEE basic
// operations (+ - * /, sqrt, floor, min, max), so a GPU that evaluates the same expression
// trees without fused multiply-add produces the same bits. For double, the C library is
// used: the FP64 restatement is the error reference, not a bit-exact twin.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

// exp(x) in single precision: Cody & Waite's two-constant reduction x = k ln2 + r,
// |r| <= ln2/2, then the degree-7 Taylor polynomial of e^r in Horner form, then an
// exact scaling by 2^k (two steps below the normal range). Relative error < 2 ulp.
inline float fexp(float x) {
  if (!(x < 88.7228f)) return x != x ? x : INFINITY;
  if (x < -103.972f) return 0.0f;
  float k = std::floor(x * 1.44269502f + 0.5f);
  float r = x - k * 0.693145752f;       // ln2 high part: 0.693145751953125, 16 bits
  r = r - k * 1.42860677e-06f;          // ln2 low part
  float p = 1.98412701e-04f;            // 1/5040
  p = p * r + 1.38888892e-03f;          // 1/720
  p = p * r + 8.33333377e-03f;          // 1/120
  p = p * r + 4.16666679e-02f;          // 1/24
  p = p * r + 1.66666672e-01f;          // 1/6
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int ki = (int)k;
  if (ki > 127) {  // 2^128 is not a float: fold one factor of two into p
    p = p * 2.0f;
    ki -= 1;
  }
  if (ki > -126) {
    uint32_t bits = (uint32_t)(ki + 127) << 23;
    float s;
    std::memcpy(&s, &bits, 4);
    return p * s;
  }
  uint32_t b1 = (uint32_t)(ki + 64 + 127) << 23;  // 2^(k+64), then 2^-64
  uint32_t b2 = (uint32_t)(127 - 64) << 23;
  float s1, s2;
  std::memcpy(&s1, &b1, 4);
  std::memcpy(&s2, &b2, 4);
  return (p * s1) * s2;
}
inline double fexp(double x) { return std::exp(x); }

// 2^x in single precision for -126 <= x <= 127: k = the nearest integer, r = (x - k) ln2
// (x - k is exact, |r| <= ln2/2), the polynomial of fexp for e^r, and an exact 2^k.
inline float fexp2(float x) {
  float k = std::floor(x + 0.5f);
  float r = (x - k) * 0.693147182f;
  float p = 1.98412701e-04f;
  p = p * r + 1.38888892e-03f;
  p = p * r + 8.33333377e-03f;
  p = p * r + 4.16666679e-02f;
  p = p * r + 1.66666672e-01f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  uint32_t bits = (uint32_t)((int)k + 127) << 23;
  float s;
  std::memcpy(&s, &bits, 4);
  return p * s;
}

// Storage rounding of the mixed-precision multigrid: float to bfloat16 and to IEEE
// binary16, round to nearest even (binary16 subnormals kept, overflow to infinity), and
// back to float.
inline float roundBF16(float v) {
  uint32_t b;
  std::memcpy(&b, &v, 4);
  b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
  std::memcpy(&v, &b, 4);
  return v;
}
inline float roundF16(float v) {
  uint32_t b;
  std::memcpy(&b, &v, 4);
  uint32_t sign = b & 0x80000000u, mag = b & 0x7fffffffu;
  float out;
  if (mag >= 0x477ff000u) {  // 65520 and up round to infinity
    uint32_t inf = sign | 0x7f800000u;
    std::memcpy(&out, &inf, 4);
    return out;
  }
  if (mag < 0x38800000u) {  // below 2^-14: multiples of 2^-24
    float a = std::fabs(v) * 16777216.0f;
    float n = std::nearbyint(a);  // round to nearest even
    out = n * 5.9604644775390625e-08f;
    return sign ? -out : out;
  }
  mag = (mag + 0xfffu + ((mag >> 13) & 1u)) & ~0x1fffu;
  uint32_t r = sign | mag;
  std::memcpy(&out, &r, 4);
  return out;
}

inline float fsqrt(float x) { return std::sqrt(x); }
inline double fsqrt(double x) { return std::sqrt(x); }
template <class R> inline R ffloor(R x) { return std::floor(x); }
template <class R> inline R fmin(R a, R b) { return b < a ? b : a; }
template <class R> inline R fmax(R a, R b) { return a < b ? b : a; }
template <class R> inline R fabs_(R a) { return a < R(0) ? -a : a; }
template <class R> inline R clampv(R x, R lo, R hi) { return fmin(fmax(x, lo), hi); }
// WGSL mix(e1, e2, e3) = e1 * (1 - e3) + e2 * e3.
template <class R> inline R mixv(R a, R b, R t) { return a * (R(1) - t) + b * t; }
template <class R> inline R fractv(R x) { return x - std::floor(x); }
// WGSL smoothstep: t * t * (3 - 2 t), t = clamp((x - e0) / (e1 - e0), 0, 1).
template <class R> inline R smoothstepv(R e0, R e1, R x) {
  R t = clampv((x - e0) / (e1 - e0), R(0), R(1));
  return t * t * (R(3) - R(2) * t);
}
template <class R> inline R length3(R x, R y, R z) { return fsqrt(x * x + y * y + z * z); }

}  // namespace oracle
