// nsx_device.h -- device primitives of the simulation core (gfx950, wave64).
//
// Everything here is evaluated one IEEE operation at a time (the library is compiled with
// -ffp-contract=off), so a CPU restatement written with the same operation order reproduces
// every result bit for bit.  Transcendentals are the project's own, written from published
// methods (Cody-Waite reduction + Taylor polynomials, odd atanh series for log); no vendor
// math routine is reproduced.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsx {

// Philox4x32-10 counter stream ids (third counter word).
constexpr uint32_t kStreamConn = 0x434F4E4Eu;   // synapse generator
constexpr uint32_t kStreamDrive = 0x44525645u;  // per-step external drive
constexpr uint32_t kStreamParam = 0x50415241u;  // per-neuron parameter jitter
constexpr int32_t kNeverStep = -1073741824;     // "no spike yet" (-2^30)

struct U4 {
  uint32_t x, y, z, w;
};

// Philox4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64-bit products with a
// Weyl-sequence key schedule.  Each product is one 64-bit multiply (v_mad_u64_u32 on the device,
// one instruction for both halves).
__host__ __device__ inline U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                     uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

__host__ __device__ inline U4 draw(uint64_t seed, uint64_t idx, uint32_t c1, uint32_t stream) {
  return philox((uint32_t)idx, c1, stream, (uint32_t)(idx >> 32), (uint32_t)seed,
                (uint32_t)(seed >> 32));
}

// [0,1) with 24 random bits
__host__ __device__ inline float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }

// exp(x): k = rint(x log2 e), r = (x - k ln2_hi) - k ln2_lo, degree-7 Taylor polynomial,
// scaled by 2^k.  x < -86 -> 0, x > 88 clamps.
__device__ inline float expf_(float x) {
  if (x < -86.0f) return 0.0f;
  x = fminf(x, 88.0f);
  float k = __builtin_rintf(x * 1.44269504088896341f);
  float r = x - k * 0.693145751953125f;
  r = r - k * 1.42860682030941723212e-6f;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return __builtin_ldexpf(p, (int)k);
}

// log(x), x positive normal: m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)).
__device__ inline float logf_(float x) {
  uint32_t b = __float_as_uint(x);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float z = s * s;
  float p = 1.0f / 11.0f;
  p = p * z + 1.0f / 9.0f;
  p = p * z + 1.0f / 7.0f;
  p = p * z + 1.0f / 5.0f;
  p = p * z + 1.0f / 3.0f;
  p = p * z;
  float s2 = s + s;
  float lm = s2 * p;
  lm = s2 + lm;
  float fe = (float)e;
  float lo = fe * 1.42860682030941723212e-6f;
  lo = lm + lo;
  float hi = fe * 0.693145751953125f;
  return hi + lo;
}

// cos(pi x), x in [0, 2]: x = q/2 + f with |f| <= 1/4; Taylor sin/cos of pi f; quadrant q.
__device__ inline float cospi_(float x) {
  float q = __builtin_rintf(x + x);
  float f = x - 0.5f * q;
  float y = f * 3.14159265358979f;
  float y2 = y * y;
  float s = 1.0f / 362880.0f;
  s = s * y2 - 1.0f / 5040.0f;
  s = s * y2 + 1.0f / 120.0f;
  s = s * y2 - 1.0f / 6.0f;
  s = s * y2;
  s = s * y;
  s = y + s;
  float c = -1.0f / 3628800.0f;
  c = c * y2 + 1.0f / 40320.0f;
  c = c * y2 - 1.0f / 720.0f;
  c = c * y2 + 1.0f / 24.0f;
  c = c * y2 - 0.5f;
  c = c * y2 + 1.0f;
  int qi = ((int)q) & 3;
  float r = qi == 0 ? c : (qi == 1 ? -s : (qi == 2 ? -c : s));
  return r;
}

// Box-Muller (1958) standard normal from two 32-bit draws.
__device__ inline float normal_(uint32_t r0, uint32_t r1) {
  float u1 = (float)((r0 >> 8) + 1u) * 5.9604644775390625e-8f;
  float u2 = (float)(r1 >> 8) * 5.9604644775390625e-8f;
  float l = logf_(u1);
  l = -2.0f * l;
  float rad = __builtin_sqrtf(l);
  float c = cospi_(u2 + u2);
  return rad * c;
}

// Fixed point: arrivals are int64 sums of rint(w * 2^S); read back as float(double(sum)*2^-S).
__device__ inline long long quantize(float w, double scale) {
  return (long long)__builtin_rint((double)w * scale);
}
__device__ inline float dequantize(long long a, double inv_scale) {
  return (float)((double)a * inv_scale);
}

__device__ inline float clampf_(float x, float lo, float hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

}  // namespace nsx
