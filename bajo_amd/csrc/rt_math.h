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

// ---- GGX microfacet conductor (no bajo counterpart) ----------------------------
// Trowbridge-Reitz (GGX) normal distribution of width alpha = the material's
// roughness, Smith height-correlated masking-shadowing
// G2 = 1 / (1 + Lambda(wo) + Lambda(wi)), Schlick Fresnel with F0 = the
// material's specular color.  Directions are drawn from the distribution of
// visible normals (Heitz 2018, "Sampling the GGX Distribution of Visible
// Normals", JCGT 7(4)): pdf(wi) = G1(wo) D(h) / (4 cos_o), weight
// f cos_i / pdf = F G2 / G1(wo).  The local frame around the shading normal
// is Duff et al. 2017 ("Building an Orthonormal Basis, Revisited", JCGT
// 6(1)).  Plain FP32 operations in a fixed order, so the CPU restatement
// gives the same bits.
constexpr float kPi = 3.14159265358979323846f;

struct Frame {
  v3 t, b, n;
};
BJ_HD Frame frame_of(v3 n) {
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  Frame f;
  f.t = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  f.b = mk(b, sign + n.y * n.y * a, -n.y);
  f.n = n;
  return f;
}
BJ_HD v3 to_local(const Frame& f, v3 v) { return mk(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
BJ_HD v3 to_world(const Frame& f, v3 v) { return add(add(scale(f.t, v.x), scale(f.b, v.y)), scale(f.n, v.z)); }

// D(h) for cos_h = n.h > 0.
BJ_HD float ggx_d(float a2, float cos_h) {
  const float c2 = cos_h * cos_h;
  const float d = c2 * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * d * d);
}
// Smith Lambda(w) for cos_w = n.w > 0.
BJ_HD float ggx_lambda(float a2, float cos_w) {
  const float c2 = cos_w * cos_w;
  const float t2 = fmax_sel(1.0f - c2, 0.0f) / c2;
  return 0.5f * (sqrtf(1.0f + a2 * t2) - 1.0f);
}
BJ_HD v3 schlick_f0(v3 f0, float c) {
  const float x = 1.0f - fmin_sel(fmax_sel(c, 0.0f), 1.0f);
  const float x2 = x * x;
  const float x5 = x2 * x2 * x;
  return mk(fmaf(1.0f - f0.x, x5, f0.x), fmaf(1.0f - f0.y, x5, f0.y), fmaf(1.0f - f0.z, x5, f0.z));
}

// f(wo, wi) and pdf(wi) in the local frame (z = shading normal).
BJ_HD BsdfEval eval_ggx_local(v3 wo, v3 wi, v3 f0, float alpha) {
  BsdfEval out;
  out.value = mk(0.0f, 0.0f, 0.0f);
  out.pdf = 0.0f;
  out.delta = false;
  if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) return out;
  const float a2 = alpha * alpha;
  const v3 h = normalize(add(wo, wi));
  const float d = ggx_d(a2, h.z);
  const float lo = ggx_lambda(a2, wo.z), li = ggx_lambda(a2, wi.z);
  const float g2 = 1.0f / (1.0f + lo + li);
  const float g1o = 1.0f / (1.0f + lo);
  const v3 f = schlick_f0(f0, dot(wo, h));
  out.value = scale(f, d * g2 / (4.0f * wo.z * wi.z));
  out.pdf = g1o * d / (4.0f * wo.z);
  return out;
}

// World-space evaluation toward `dir` for a ray arriving along ray_d at a
// surface with oriented normal n (n . ray_d < 0).
BJ_HD BsdfEval eval_ggx(v3 ray_d, v3 n, v3 f0, float alpha, v3 dir) {
  const Frame fr = frame_of(n);
  const v3 wo = to_local(fr, mk(-ray_d.x, -ray_d.y, -ray_d.z));
  const v3 wi = to_local(fr, dir);
  return eval_ggx_local(normalize(wo), normalize(wi), f0, alpha);
}

// Visible-normal sample (u1, u2 in [0, 1)); ok = false when the reflected
// direction falls below the surface (the sample is absorbed).
BJ_HD BsdfSample sample_ggx(v3 ray_d, v3 n, v3 f0, float alpha, float u1, float u2) {
  BsdfSample out;
  out.direction = n;
  out.weight = mk(0.0f, 0.0f, 0.0f);
  out.pdf = 0.0f;
  out.delta = false;
  out.ok = false;
  const Frame fr = frame_of(n);
  const v3 wo = normalize(to_local(fr, mk(-ray_d.x, -ray_d.y, -ray_d.z)));
  if (!(wo.z > 0.0f)) return out;
  const v3 vh = normalize(mk(alpha * wo.x, alpha * wo.y, wo.z));
  const float lensq = vh.x * vh.x + vh.y * vh.y;
  const float inv_len = lensq > 0.0f ? 1.0f / sqrtf(lensq) : 0.0f;
  const v3 t1v = lensq > 0.0f ? mk(-vh.y * inv_len, vh.x * inv_len, 0.0f) : mk(1.0f, 0.0f, 0.0f);
  const v3 t2v = cross(vh, t1v);
  const float r = sqrtf(u1);
  float sp, cp;
  sincos_turns(u2, sp, cp);
  const float t1 = r * cp;
  float t2 = r * sp;
  const float sh = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - sh) * sqrtf(fmax_sel(1.0f - t1 * t1, 0.0f)) + sh * t2;
  const float tz = sqrtf(fmax_sel(1.0f - t1 * t1 - t2 * t2, 0.0f));
  const v3 nh = add(add(scale(t1v, t1), scale(t2v, t2)), scale(vh, tz));
  const v3 h = normalize(mk(alpha * nh.x, alpha * nh.y, fmax_sel(nh.z, 0.0f)));
  const float oh = dot(wo, h);
  const v3 wi = sub(scale(h, 2.0f * oh), wo);
  if (!(wi.z > 0.0f) || !(oh > 0.0f)) return out;
  const float a2 = alpha * alpha;
  const float lo = ggx_lambda(a2, wo.z), li = ggx_lambda(a2, wi.z);
  const v3 f = schlick_f0(f0, oh);
  out.weight = scale(f, (1.0f + lo) / (1.0f + lo + li));
  out.pdf = ggx_d(a2, h.z) / ((1.0f + lo) * 4.0f * wo.z);
  out.direction = normalize(to_world(fr, wi));
  out.ok = out.pdf > 0.0f;
  return out;
}

}  // namespace bj
