// Learned path guiding: the network's input encoding and its output, a
// mixture of 8 von Mises-Fisher lobes over directions (host + device).
// bajo has no counterpart (its renderer samples only the BSDF,
// rt/gpu/path_shading.mojo:463-535, rt/shading.mojo:92-118).
//
// Network: 9 inputs (position normalized to the scene box, shading normal,
// outgoing direction), each encoded at 12 frequencies as
// sin / cos(pi 2^f x) -> 216 features (frequencies 0, 4, 8 from the series,
// the others by doubling); three hidden layers of 64 ReLU units;
// 40 outputs = 8 lobes x (mean direction 3, log concentration 1, logit 1).
// Feature k = 24 d + 2 f + c (input d, frequency f, c = 0 sin / 1 cos).
//
// Lobe j: mean mu = u / |u| of its three outputs, concentration
// kappa = exp(clamp(s, -4, 7)), weight = softmax of the logits.  Density
// vMF(w) = kappa / (2 pi (1 - e^(-2 kappa))) exp(kappa (mu.w - 1)), sampled
// by inverting the CDF of mu.w (Jakob 2012, "Numerically stable sampling of
// the von Mises Fisher distribution on S^2"):
// mu.w = 1 + ln(u + (1 - u) e^(-2 kappa)) / kappa.
//
// Every function is plain FP32 arithmetic with explicit fmaf and this
// library's own exp / log / sin / cos series, so the CPU restatement gives
// the same bits.
#pragma once

#include "rt_math.h"

namespace bj {

constexpr int kGuideInputs = 9;
constexpr int kGuideFreqs = 12;
constexpr int kGuideFeatures = kGuideInputs * kGuideFreqs * 2;  // 216
constexpr int kGuideK0 = 224;      // features padded to the MFMA K step; column 216 is the bias input 1
constexpr int kGuideHidden = 64;
constexpr int kGuideLobes = 8;
constexpr int kGuideOut = 5 * kGuideLobes;  // 40
constexpr int kGuideOutPad = 48;
constexpr float kGuideLogKappaMin = -4.0f, kGuideLogKappaMax = 7.0f;

// Parameter vector (FP32 master copy): W1[64][216], b1[64], W2[64][64], b2,
// W3[64][64], b3, W4[40][64], b4[40]; W[j][k] maps input k to output j.
constexpr int kGuideOffW1 = 0;
constexpr int kGuideOffB1 = kGuideOffW1 + kGuideHidden * kGuideFeatures;
constexpr int kGuideOffW2 = kGuideOffB1 + kGuideHidden;
constexpr int kGuideOffB2 = kGuideOffW2 + kGuideHidden * kGuideHidden;
constexpr int kGuideOffW3 = kGuideOffB2 + kGuideHidden;
constexpr int kGuideOffB3 = kGuideOffW3 + kGuideHidden * kGuideHidden;
constexpr int kGuideOffW4 = kGuideOffB3 + kGuideHidden;
constexpr int kGuideOffB4 = kGuideOffW4 + kGuideOut * kGuideHidden;
constexpr int kGuideParams = kGuideOffB4 + kGuideOut;  // 24,808

// e^x for x in [-87, 87] (larger x is clamped): 2^t, t = x log2 e = n + r,
// |r| <= 1/2, e^(r ln 2) by its Taylor series to degree 8 (as exp2_nonpos);
// below -87 it is 0.
BJ_HD float exp_f(float x) {
  if (x < -87.0f) return 0.0f;
  if (x > 87.0f) x = 87.0f;
  const float t = x * 1.44269504088896340736f;
  const float n = rintf(t);
  const float y = (t - n) * 0.693147180559945309417f;
  float p = fmaf(y, 2.48015873015873015873e-5f, 1.98412698412698412698e-4f);
  p = fmaf(y, p, 1.38888888888888888889e-3f);
  p = fmaf(y, p, 8.33333333333333333333e-3f);
  p = fmaf(y, p, 4.16666666666666666667e-2f);
  p = fmaf(y, p, 1.66666666666666666667e-1f);
  p = fmaf(y, p, 0.5f);
  p = fmaf(y, p, 1.0f);
  p = fmaf(y, p, 1.0f);
  return p * __builtin_bit_cast(float, (uint32_t)((int)n + 127) << 23);
}

// ln x for x > 0 (log2_pos times ln 2).
BJ_HD float ln_f(float x) { return log2_pos(x) * 0.693147180559945309417f; }

// Encoded pair (sin, cos)(pi 2^f x) of one input value: the argument in
// turns, x 2^(f-1), is reduced to its fraction first (exact scaling).
BJ_HD void guide_encode_pair(float x, int f, float& s, float& c) {
  const float y = x * __builtin_bit_cast(float, (uint32_t)(f - 1 + 127) << 23);
  const float u = y - floorf(y);
  sincos_turns(u, s, c);
}

// The four pairs of frequencies 4 b .. 4 b + 3: the first from the series,
// the next three by the double-angle formulas sin 2a = 2 sin a cos a,
// cos 2a = (cos a - sin a)(cos a + sin a), in this order of FP32 operations
// (each doubling at most doubles the first pair's error, well below the
// BF16 features' precision).
BJ_HD void guide_encode_group(float x, int b, float (&s)[4], float (&c)[4]) {
  guide_encode_pair(x, 4 * b, s[0], c[0]);
#pragma unroll
  for (int t = 1; t < 4; ++t) {
    s[t] = 2.0f * s[t - 1] * c[t - 1];
    c[t] = (c[t - 1] - s[t - 1]) * (c[t - 1] + s[t - 1]);
  }
}

// The 9 network inputs of a path vertex.
struct GuideInput {
  float v[kGuideInputs];
};
BJ_HD GuideInput guide_input(v3 p, v3 n, v3 wo, v3 lo, v3 inv_ext) {
  GuideInput g;
  const float q[3] = {(p.x - lo.x) * inv_ext.x, (p.y - lo.y) * inv_ext.y, (p.z - lo.z) * inv_ext.z};
  for (int a = 0; a < 3; ++a) g.v[a] = fmin_sel(fmax_sel(q[a], 0.0f), 1.0f);
  g.v[3] = n.x; g.v[4] = n.y; g.v[5] = n.z;
  g.v[6] = wo.x; g.v[7] = wo.y; g.v[8] = wo.z;
  return g;
}

struct GuideLobe {
  v3 mu;
  float kappa, log_norm;  // log_norm = ln(kappa / (2 pi (1 - e^(-2 kappa))))
};
struct GuideMixture {
  GuideLobe lobe[kGuideLobes];
  float weight[kGuideLobes];
};

// ln(2 pi)
constexpr float kLn2Pi = 1.83787706640934548356f;

BJ_HD GuideLobe guide_lobe(const float* o) {
  GuideLobe l;
  const float len2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
  if (len2 > 1.0e-24f) {
    const float inv = 1.0f / sqrtf(len2);
    l.mu = mk(o[0] * inv, o[1] * inv, o[2] * inv);
  } else {
    l.mu = mk(0.0f, 0.0f, 1.0f);
  }
  const float s = fmin_sel(fmax_sel(o[3], kGuideLogKappaMin), kGuideLogKappaMax);
  l.kappa = exp_f(s);
  const float e2 = exp_f(-2.0f * l.kappa);
  l.log_norm = s - kLn2Pi - ln_f(1.0f - e2);
  return l;
}

// Decodes the 40 outputs of one vertex.
BJ_HD void guide_decode(const float* o, GuideMixture& m) {
  float amax = o[4];
#pragma unroll
  for (int j = 1; j < kGuideLobes; ++j) amax = fmax_sel(amax, o[5 * j + 4]);
  float sum = 0.0f;
#pragma unroll
  for (int j = 0; j < kGuideLobes; ++j) {
    m.lobe[j] = guide_lobe(o + 5 * j);
    m.weight[j] = exp_f(o[5 * j + 4] - amax);
    sum += m.weight[j];
  }
  const float inv = 1.0f / sum;
#pragma unroll
  for (int j = 0; j < kGuideLobes; ++j) m.weight[j] *= inv;
}

BJ_HD float vmf_pdf(const GuideLobe& l, v3 w) { return exp_f(fmaf(l.kappa, dot(l.mu, w) - 1.0f, l.log_norm)); }

BJ_HD float guide_pdf(const GuideMixture& m, v3 w) {
  float p = 0.0f;
#pragma unroll
  for (int j = 0; j < kGuideLobes; ++j) p = fmaf(m.weight[j], vmf_pdf(m.lobe[j], w), p);
  return p;
}

// A direction of lobe j from two uniforms.
BJ_HD v3 vmf_sample(const GuideLobe& l, float u1, float u2) {
  const float e2 = exp_f(-2.0f * l.kappa);
  float w = 1.0f + ln_f(fmaf(1.0f - u1, e2, u1)) / l.kappa;
  w = fmin_sel(fmax_sel(w, -1.0f), 1.0f);
  const float r = sqrtf(fmax_sel(1.0f - w * w, 0.0f));
  float sp, cp;
  sincos_turns(u2, sp, cp);
  const Frame fr = frame_of(l.mu);
  const v3 d = add(add(scale(fr.t, r * cp), scale(fr.b, r * sp)), scale(l.mu, w));
  return normalize(d);
}

// A direction of the mixture decoded from outputs `o`: lobe by u_lobe
// against the cumulative weights, that lobe decoded again from its outputs
// (same bits as m.lobe[sel]; picking a lobe out of `m` by a run-time index
// would put the mixture in scratch memory).
BJ_HD v3 guide_sample(const GuideMixture& m, const float* o, float u_lobe, float u1, float u2) {
  int sel = kGuideLobes - 1;
  float c = 0.0f;
  bool found = false;
#pragma unroll
  for (int j = 0; j < kGuideLobes - 1; ++j) {
    c += m.weight[j];
    const bool take = !found && u_lobe < c;
    sel = take ? j : sel;
    found = found || take;
  }
  return vmf_sample(guide_lobe(o + 5 * sel), u1, u2);
}

// Stage of the guide's random numbers (domain 3 beside BSDF 0, roulette 1,
// light 2): u_choose (BSDF or guide), u_lobe, u1, u2 = one Philox block.
constexpr uint32_t kRngGuide = 3;
BJ_HD uint32_t rng_guide_stage(uint32_t bounce) { return (kRngGuide << kRngDomainShift) | (bounce + 1u); }

// One-sample MIS of BSDF and guide sampling, each chosen with probability
// 1/2: the density of the combined strategy.
constexpr float kGuideMix = 0.5f;
BJ_HD float guide_mix_pdf(float bsdf_pdf, float guide_pdf_v) {
  return fmaf(kGuideMix, guide_pdf_v, (1.0f - kGuideMix) * bsdf_pdf);
}

}  // namespace bj
