// Random streams and sampling primitives of the wavefront renderer
// (host + device, bit-identical on both).
//
// bajo draws every path stage from its own Philox stream, keyed by the
// render seed with subsequence (path_id << 32) | stage
// (rt/wavefront_contract.mojo:107-134, rt/common.mojo:37-40), and consumes
// the four floats of each Philox block in order (core/random.mojo:7-36).
// bajo's generator is the Mojo stdlib Random[10]; its exact
// (seed, subsequence) -> counter/key mapping is not available here, so this
// file defines its own: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// with key = seed and counter = (block, 0, subsequence lo, subsequence hi),
// floats = (x >> 8) * 2^-24.
//
// sin/cos of 2*pi*u use a Taylor polynomial on a quarter-turn-reduced
// argument with explicit FMAs, so the device and the CPU oracle agree to
// the bit; accuracy is within 2 ulp of libm on [0, 1).
#pragma once

#include "bajo_common.h"

namespace bj {

// ---- Philox4x32-10 ---------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

BJ_HD void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

BJ_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, c.x, hi0, lo0);
    mulhilo(M1, c.z, hi1, lo1);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

BJ_HD float u32_to_unit(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// One path stage's stream: four floats per Philox block, consumed in order.
struct StageRng {
  uint32_t k0, k1, s0, s1, block;
  float buf[4];
  int used;

  BJ_HD StageRng(uint64_t seed, uint32_t path_id, uint32_t stage)
      : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), s0(stage), s1(path_id), block(0), used(4) {}

  BJ_HD float next() {
    if (used >= 4) {
      U4 r = philox4x32_10(U4{block, 0u, s0, s1}, k0, k1);
      ++block;
      buf[0] = u32_to_unit(r.x);
      buf[1] = u32_to_unit(r.y);
      buf[2] = u32_to_unit(r.z);
      buf[3] = u32_to_unit(r.w);
      used = 0;
    }
    return buf[used++];
  }
};

// rt/wavefront_contract.mojo:33-39,107-134: stage domains in bits 30-31.
constexpr uint32_t kRngDomainShift = 30;
constexpr uint32_t kRngBsdf = 0, kRngRoulette = 1, kRngLight = 2;
BJ_HD uint32_t rng_bsdf_stage(uint32_t bounce) { return (kRngBsdf << kRngDomainShift) | (bounce + 1u); }
BJ_HD uint32_t rng_roulette_stage(uint32_t bounce) { return (kRngRoulette << kRngDomainShift) | (bounce + 1u); }
BJ_HD uint32_t rng_light_stage(uint32_t bounce) { return (kRngLight << kRngDomainShift) | (bounce + 1u); }

// ---- sin / cos of 2*pi*u ------------------------------------------------------
// u is reduced to r = u - q/4 with q = nearest quarter turn (exact in FP32 for
// u in [0, 1)), x = 2*pi*r in [-pi/4, pi/4]; Taylor series to x^9 / x^10.
BJ_HD void sincos_turns(float u, float& s, float& c) {
  const float q = rintf(u * 4.0f);
  const float r = u - q * 0.25f;
  const float x = r * 6.28318530717958647692f;
  const float x2 = x * x;
  float ps = fmaf(x2, 2.75573192239858906526e-6f, -1.98412698412698412698e-4f);
  ps = fmaf(x2, ps, 8.33333333333333333333e-3f);
  ps = fmaf(x2, ps, -1.66666666666666666667e-1f);
  const float sn = fmaf(x * x2, ps, x);
  float pc = fmaf(x2, -2.75573192239858906526e-7f, 2.48015873015873015873e-5f);
  pc = fmaf(x2, pc, -1.38888888888888888889e-3f);
  pc = fmaf(x2, pc, 4.16666666666666666667e-2f);
  pc = fmaf(x2, pc, -0.5f);
  const float cs = fmaf(x2, pc, 1.0f);
  const int quadrant = ((int)q) & 3;
  switch (quadrant) {
    case 0: s = sn; c = cs; break;
    case 1: s = cs; c = -sn; break;
    case 2: s = -sn; c = -cs; break;
    default: s = -cs; c = sn; break;
  }
}

// ---- small helpers of bajo's renderer -----------------------------------------
constexpr float kInvPi = 0.3183098861837907f;
constexpr float kT_MinSecondary = 0.001f;

// rt/common.mojo:24-33.
BJ_HD v3 sky_color(v3 d) {
  v3 u = normalize(d);
  float a = 0.5f * (u.y + 1.0f);
  return add(scale(mk(1.0f, 1.0f, 1.0f), 1.0f - a), scale(mk(0.5f, 0.7f, 1.0f), a));
}

BJ_HD float length2(v3 v) { return dot(v, v); }

}  // namespace bj
