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

// ---- x^y for x in [0, 1] and y >= 0 (the Phong lobe of bajo's metal) --------
// log2 x = k + ln(m) / ln 2 with x = m 2^k, m in [sqrt(1/2), sqrt(2)), and
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), by the odd series of atanh to
// s^11 (|s| <= 0.172); 2^t with t = n + f, n = rint(t), |f| <= 1/2, and
// e^(f ln 2) by its Taylor series to degree 8.  Plain FP32 operations and
// fmaf only, so the host restatement gives the same bits.  Results below
// 2^-125 are returned as 0.
BJ_HD float log2_pos(float x) {
  uint32_t b = __builtin_bit_cast(uint32_t, x);
  int k = 0;
  if (b < 0x00800000u) {  // subnormal input: scale into the normal range
    x *= 8388608.0f;
    b = __builtin_bit_cast(uint32_t, x);
    k = -23;
  }
  k += (int)(b >> 23) - 127;
  float m = __builtin_bit_cast(float, (b & 0x007FFFFFu) | 0x3F800000u);
  if (m > 1.41421356237309504880f) {
    m *= 0.5f;
    k += 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float p = fmaf(s2, 0.0909090909090909090909f, 0.111111111111111111111f);
  p = fmaf(s2, p, 0.142857142857142857143f);
  p = fmaf(s2, p, 0.2f);
  p = fmaf(s2, p, 0.333333333333333333333f);
  const float ln_m = 2.0f * fmaf(s * s2, p, s);
  return fmaf(ln_m, 1.44269504088896340736f, (float)k);
}

BJ_HD float exp2_nonpos(float t) {
  if (t < -125.0f) return 0.0f;
  const float n = rintf(t);
  const float x = (t - n) * 0.693147180559945309417f;
  float p = fmaf(x, 2.48015873015873015873e-5f, 1.98412698412698412698e-4f);
  p = fmaf(x, p, 1.38888888888888888889e-3f);
  p = fmaf(x, p, 8.33333333333333333333e-3f);
  p = fmaf(x, p, 4.16666666666666666667e-2f);
  p = fmaf(x, p, 1.66666666666666666667e-1f);
  p = fmaf(x, p, 0.5f);
  p = fmaf(x, p, 1.0f);
  p = fmaf(x, p, 1.0f);
  return p * __builtin_bit_cast(float, (uint32_t)((int)n + 127) << 23);
}

BJ_HD float pow_unit(float x, float y) {
  if (!(x > 0.0f)) return y == 0.0f ? 1.0f : 0.0f;
  if (y == 0.0f) return 1.0f;
  const float t = y * log2_pos(x);
  return exp2_nonpos(t > 0.0f ? 0.0f : t);
}

// ---- bajo's metal and dielectric (rt/shading.mojo) -----------------------------
constexpr float kTwoPi = 6.28318530717958647692f;

// reflect / refract / reflectance, rt/shading.mojo:226-260.
BJ_HD v3 reflect(v3 v, v3 n) { return sub(v, scale(n, 2.0f * dot(v, n))); }
BJ_HD v3 refract(v3 uv, v3 n, float eta) {
  const float cos_theta = fmin_sel(dot(mk(-uv.x, -uv.y, -uv.z), n), 1.0f);
  const v3 perp = scale(add(uv, scale(n, cos_theta)), eta);
  const v3 par = scale(n, -sqrtf(fabsf(1.0f - length2(perp))));
  return add(perp, par);
}
BJ_HD float reflectance(float cosine, float ref_idx) {
  const float root = (1.0f - ref_idx) / (1.0f + ref_idx);
  const float r2 = root * root;
  const float x = 1.0f - cosine;
  const float x2 = x * x;
  const float x5 = x2 * x2 * x;
  return fmaf(1.0f - r2, x5, r2);
}

struct BsdfEval {
  v3 value;
  float pdf;
  bool delta;
};
struct BsdfSample {
  v3 direction, weight;
  float pdf;
  bool delta, ok;
};

// _evaluate_metal, rt/shading.mojo:30-57: a normalized Phong lobe around the
// mirror direction, exponent 2 / fuzz^2 - 2; fuzz <= 1e-4 is a mirror (delta).
BJ_HD BsdfEval eval_metal(v3 ray_d, v3 n, v3 albedo, float fuzz, v3 out_dir) {
  const bool smooth = fuzz <= 1.0e-4f;
  const float sf = smooth ? 1.0f : fuzz;
  const v3 dir = normalize(out_dir);
  const v3 refl = normalize(reflect(normalize(ray_d), n));
  const bool surface_valid = dot(n, dir) > 0.0f;
  const float lobe_cos = fmax_sel(dot(refl, dir), 0.0f);
  const bool lobe_valid = lobe_cos > 0.0f;
  const float e = fmax_sel(2.0f / (sf * sf) - 2.0f, 0.0f);
  const float lobe = pow_unit(lobe_cos, e);
  const bool valid = !smooth && surface_valid && lobe_valid;
  const float vs = (e + 2.0f) * lobe / kTwoPi;
  const float pdf = (e + 1.0f) * lobe / kTwoPi;
  return BsdfEval{valid ? scale(albedo, vs) : mk(0.0f, 0.0f, 0.0f), valid ? pdf : 0.0f, smooth};
}

// _sample_metal, rt/shading.mojo:120-160 (cos / sin of 2 pi v by sincos_turns).
BJ_HD BsdfSample sample_metal(v3 ray_d, v3 n, v3 albedo, float fuzz, float ru, float rv) {
  const v3 refl = normalize(reflect(normalize(ray_d), n));
  const bool smooth = fuzz <= 1.0e-4f;
  const float sf = smooth ? 1.0f : fuzz;
  const float e = fmax_sel(2.0f / (sf * sf) - 2.0f, 0.0f);
  const float cos_t = pow_unit(ru, 1.0f / (e + 1.0f));
  const float sin_t = sqrtf(fmax_sel(1.0f - cos_t * cos_t, 0.0f));
  float sp, cp;
  sincos_turns(rv, sp, cp);
  const v3 helper = fabsf(refl.y) > 0.99f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
  const v3 tangent = normalize(cross(helper, refl));
  const v3 bitangent = cross(refl, tangent);
  const v3 rough = normalize(add(add(scale(tangent, cp * sin_t), scale(bitangent, sp * sin_t)), scale(refl, cos_t)));
  const v3 dir = smooth ? refl : rough;
  const BsdfEval ev = eval_metal(ray_d, n, albedo, fuzz, dir);
  const float scos = fmax_sel(dot(n, dir), 0.0f);
  const bool rough_ok = ev.pdf > 0.0f && scos > 0.0f;
  const float safe_pdf = rough_ok ? ev.pdf : 1.0f;
  const float k = scos / safe_pdf;
  BsdfSample out;
  out.direction = dir;
  out.weight = smooth ? albedo : scale(ev.value, k);
  out.pdf = smooth ? 1.0f : ev.pdf;
  out.delta = smooth;
  out.ok = smooth ? dot(refl, n) > 0.0f : rough_ok;
  return out;
}

// _sample_dielectric, rt/shading.mojo:163-196: Schlick reflection
// probability, total internal reflection, weight 1, always delta.
BJ_HD BsdfSample sample_dielectric(v3 ray_d, v3 n, float ior, bool front, float ru) {
  const float ri = front ? 1.0f / ior : ior;
  const v3 u = normalize(ray_d);
  const float cos_t = fmin_sel(dot(mk(-u.x, -u.y, -u.z), n), 1.0f);
  const float sin_t = sqrtf(fmax_sel(1.0f - cos_t * cos_t, 0.0f));
  const bool cannot = ri * sin_t > 1.0f;
  const float rp = reflectance(cos_t, ri);
  const bool refl = cannot || rp > ru;
  const v3 dir = normalize(refl ? reflect(u, n) : refract(u, n, ri));
  BsdfSample out;
  out.direction = dir;
  out.weight = mk(1.0f, 1.0f, 1.0f);
  out.pdf = refl ? (cannot ? 1.0f : rp) : 1.0f - rp;
  out.delta = true;
  out.ok = true;
  return out;
}

}  // namespace bj
