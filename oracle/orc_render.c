/*
 * CPU oracle -- test infrastructure only (see orc_core.h).
 *
 * Restates bajo's wavefront path tracer for Lambertian + emissive triangle
 * scenes, one path at a time, in the stage order the GPU wavefront uses
 * (extend -> shadow per bounce):
 *   - stage streams, (path_id << 32) | stage   rt/wavefront_contract.mojo:107-134, rt/common.mojo:37-40
 *   - primary ray (lens disk, jitter)           rt/gpu/common_kernels.mojo:22-94, bvh/camera.mojo:156-208
 *   - sky on miss                               rt/common.mojo:24-33
 *   - oriented normal                           rt/geometry.mojo:20-30
 *   - NEE candidate (alias pick, area sample)   rt/gpu/path_shading.mojo:261-403
 *   - emission weights (PATH / NEE / MIS)       rt/gpu/path_shading.mojo:570-610, rt/lighting.mojo:7-38
 *   - Lambertian sample + roulette              rt/shading.mojo:92-118, rt/common.mojo:43-67
 *   - AO ray / NORMALS                          rt/gpu/path_shading.mojo:651-667, :556-563
 *   - light table (Walker alias, LIFO lists)    rt/types.mojo:207-247, :490-506, :527-537
 *   - resolve                                   rt/gpu/common_kernels.mojo:97-136
 * The random generator is this repository's own Philox4x32-10 stream
 * definition (bajo's Mojo stdlib Random[10] mapping is not available).
 *
 * orc_render_fp64 is an independent double-precision estimator of the same
 * images (own PCG32 streams, brute-force double intersection), used to check
 * radiance statistically.
 */
#include <math.h>
#include <time.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "orc_internal.h"

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---------------------------------- */
void orc_philox4x32_10(const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

typedef struct { uint32_t k0, k1, s0, s1, block; float buf[4]; int used; } orc_rng;

static void rng_init(orc_rng *r, uint64_t seed, uint32_t path_id, uint32_t stage) {
    r->k0 = (uint32_t)seed; r->k1 = (uint32_t)(seed >> 32); r->s0 = stage; r->s1 = path_id;
    r->block = 0; r->used = 4;
}
static float rng_next(orc_rng *r) {
    if (r->used >= 4) {
        uint32_t c[4] = {r->block, 0u, r->s0, r->s1}, o[4];
        orc_philox4x32_10(c, r->k0, r->k1, o);
        r->block++;
        for (int i = 0; i < 4; ++i) r->buf[i] = (float)(o[i] >> 8) * (1.0f / 16777216.0f);
        r->used = 0;
    }
    return r->buf[r->used++];
}

void orc_rng_stage_floats(uint64_t seed, uint32_t path_id, uint32_t stage, uint32_t count, float *out) {
    orc_rng r;
    rng_init(&r, seed, path_id, stage);
    for (uint32_t i = 0; i < count; ++i) out[i] = rng_next(&r);
}

static uint32_t stage_bsdf(uint32_t b) { return (0u << 30) | (b + 1u); }
static uint32_t stage_roulette(uint32_t b) { return (1u << 30) | (b + 1u); }
static uint32_t stage_light(uint32_t b) { return (2u << 30) | (b + 1u); }

/* sin/cos(2 pi u): quarter-turn reduction, Taylor series with FMAs. */
void orc_sincos_turns(float u, float *s, float *c) {
    float q = rintf(u * 4.0f);
    float r = u - q * 0.25f;
    float x = r * 6.28318530717958647692f;
    float x2 = x * x;
    float ps = fmaf(x2, 2.75573192239858906526e-6f, -1.98412698412698412698e-4f);
    ps = fmaf(x2, ps, 8.33333333333333333333e-3f);
    ps = fmaf(x2, ps, -1.66666666666666666667e-1f);
    float sn = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, -2.75573192239858906526e-7f, 2.48015873015873015873e-5f);
    pc = fmaf(x2, pc, -1.38888888888888888889e-3f);
    pc = fmaf(x2, pc, 4.16666666666666666667e-2f);
    pc = fmaf(x2, pc, -0.5f);
    float cs = fmaf(x2, pc, 1.0f);
    switch (((int)q) & 3) {
        case 0: *s = sn; *c = cs; break;
        case 1: *s = cs; *c = -sn; break;
        case 2: *s = -sn; *c = -cs; break;
        default: *s = -cs; *c = sn; break;
    }
}

/* ---- x^y for x in [0, 1], y >= 0 (the metal's Phong lobe) --------------------
 * log2 via x = m 2^k, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) by
 * its odd series to s^11; 2^t via t = n + f, n = rint(t), e^(f ln 2) by its
 * Taylor series to degree 8; results below 2^-125 are 0.  (The renderer's
 * own definition; FP32 operations and fmaf only.) */
float orc_log2_pos(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    int k = 0;
    if (b < 0x00800000u) { x *= 8388608.0f; memcpy(&b, &x, 4); k = -23; }
    k += (int)(b >> 23) - 127;
    uint32_t mb = (b & 0x007FFFFFu) | 0x3F800000u;
    float m;
    memcpy(&m, &mb, 4);
    if (m > 1.41421356237309504880f) { m *= 0.5f; k += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = fmaf(s2, 0.0909090909090909090909f, 0.111111111111111111111f);
    p = fmaf(s2, p, 0.142857142857142857143f);
    p = fmaf(s2, p, 0.2f);
    p = fmaf(s2, p, 0.333333333333333333333f);
    float ln_m = 2.0f * fmaf(s * s2, p, s);
    return fmaf(ln_m, 1.44269504088896340736f, (float)k);
}

static float orc_exp2_nonpos(float t) {
    if (t < -125.0f) return 0.0f;
    float n = rintf(t);
    float x = (t - n) * 0.693147180559945309417f;
    float p = fmaf(x, 2.48015873015873015873e-5f, 1.98412698412698412698e-4f);
    p = fmaf(x, p, 1.38888888888888888889e-3f);
    p = fmaf(x, p, 8.33333333333333333333e-3f);
    p = fmaf(x, p, 4.16666666666666666667e-2f);
    p = fmaf(x, p, 1.66666666666666666667e-1f);
    p = fmaf(x, p, 0.5f);
    p = fmaf(x, p, 1.0f);
    p = fmaf(x, p, 1.0f);
    uint32_t eb = (uint32_t)((int)n + 127) << 23;
    float sc;
    memcpy(&sc, &eb, 4);
    return p * sc;
}

float orc_pow_unit(float x, float y) {
    if (!(x > 0.0f)) return y == 0.0f ? 1.0f : 0.0f;
    if (y == 0.0f) return 1.0f;
    float t = y * orc_log2_pos(x);
    return orc_exp2_nonpos(t > 0.0f ? 0.0f : t);
}

/* ---- metal and dielectric, rt/shading.mojo:30-260 ------------------------------ */
#define TWO_PI_F 6.28318530717958647692f

static orc_v3 orc_reflect(orc_v3 v, orc_v3 n) { return orc_sub(v, orc_scale(n, 2.0f * orc_dot(v, n))); }

static orc_v3 orc_refract(orc_v3 uv, orc_v3 n, float eta) {
    float cos_t = orc_fmin(orc_dot(orc_v(-uv.x, -uv.y, -uv.z), n), 1.0f);
    orc_v3 perp = orc_scale(orc_add(uv, orc_scale(n, cos_t)), eta);
    orc_v3 par = orc_scale(n, -sqrtf(fabsf(1.0f - orc_dot(perp, perp))));
    return orc_add(perp, par);
}

float orc_reflectance(float cosine, float ref_idx) {
    float root = (1.0f - ref_idx) / (1.0f + ref_idx);
    float r2 = root * root;
    float x = 1.0f - cosine;
    float x2 = x * x;
    float x5 = x2 * x2 * x;
    return fmaf(1.0f - r2, x5, r2);
}

typedef struct { orc_v3 value; float pdf; int delta; } orc_eval;
typedef struct { orc_v3 dir, weight; float pdf; int delta, ok; } orc_sample;

/* _evaluate_metal, rt/shading.mojo:30-57 */
static orc_eval eval_metal(orc_v3 ray_d, orc_v3 n, orc_v3 albedo, float fuzz, orc_v3 out_dir) {
    int smooth = fuzz <= 1.0e-4f;
    float sf = smooth ? 1.0f : fuzz;
    orc_v3 dir = orc_normalize(out_dir);
    orc_v3 refl = orc_normalize(orc_reflect(orc_normalize(ray_d), n));
    int surface_valid = orc_dot(n, dir) > 0.0f;
    float lobe_cos = orc_fmax(orc_dot(refl, dir), 0.0f);
    int lobe_valid = lobe_cos > 0.0f;
    float e = orc_fmax(2.0f / (sf * sf) - 2.0f, 0.0f);
    float lobe = orc_pow_unit(lobe_cos, e);
    int valid = !smooth && surface_valid && lobe_valid;
    float vs = (e + 2.0f) * lobe / TWO_PI_F;
    float pdf = (e + 1.0f) * lobe / TWO_PI_F;
    orc_eval r = {valid ? orc_scale(albedo, vs) : orc_v(0, 0, 0), valid ? pdf : 0.0f, smooth};
    return r;
}

/* _sample_metal, rt/shading.mojo:120-160 */
static orc_sample sample_metal(orc_v3 ray_d, orc_v3 n, orc_v3 albedo, float fuzz, float ru, float rv) {
    orc_v3 refl = orc_normalize(orc_reflect(orc_normalize(ray_d), n));
    int smooth = fuzz <= 1.0e-4f;
    float sf = smooth ? 1.0f : fuzz;
    float e = orc_fmax(2.0f / (sf * sf) - 2.0f, 0.0f);
    float cos_t = orc_pow_unit(ru, 1.0f / (e + 1.0f));
    float sin_t = sqrtf(orc_fmax(1.0f - cos_t * cos_t, 0.0f));
    float sp, cp;
    orc_sincos_turns(rv, &sp, &cp);
    orc_v3 helper = fabsf(refl.y) > 0.99f ? orc_v(1, 0, 0) : orc_v(0, 1, 0);
    orc_v3 tangent = orc_normalize(orc_cross(helper, refl));
    orc_v3 bitangent = orc_cross(refl, tangent);
    orc_v3 rough = orc_normalize(orc_add(orc_add(orc_scale(tangent, cp * sin_t), orc_scale(bitangent, sp * sin_t)),
                                         orc_scale(refl, cos_t)));
    orc_v3 dir = smooth ? refl : rough;
    orc_eval ev = eval_metal(ray_d, n, albedo, fuzz, dir);
    float scos = orc_fmax(orc_dot(n, dir), 0.0f);
    int rough_ok = ev.pdf > 0.0f && scos > 0.0f;
    float safe_pdf = rough_ok ? ev.pdf : 1.0f;
    float k = scos / safe_pdf;
    orc_sample r;
    r.dir = dir;
    r.weight = smooth ? albedo : orc_scale(ev.value, k);
    r.pdf = smooth ? 1.0f : ev.pdf;
    r.delta = smooth;
    r.ok = smooth ? orc_dot(refl, n) > 0.0f : rough_ok;
    return r;
}

/* _sample_dielectric, rt/shading.mojo:163-196 */
static orc_sample sample_dielectric(orc_v3 ray_d, orc_v3 n, float ior, int front, float ru) {
    float ri = front ? 1.0f / ior : ior;
    orc_v3 u = orc_normalize(ray_d);
    float cos_t = orc_fmin(orc_dot(orc_v(-u.x, -u.y, -u.z), n), 1.0f);
    float sin_t = sqrtf(orc_fmax(1.0f - cos_t * cos_t, 0.0f));
    int cannot = ri * sin_t > 1.0f;
    float rp = orc_reflectance(cos_t, ri);
    int refl = cannot || rp > ru;
    orc_sample r;
    r.dir = orc_normalize(refl ? orc_reflect(u, n) : orc_refract(u, n, ri));
    r.weight = orc_v(1, 1, 1);
    r.pdf = refl ? (cannot ? 1.0f : rp) : 1.0f - rp;
    r.delta = 1;
    r.ok = 1;
    return r;
}

/* Test hooks: out = dir.xyz, weight.xyz, pdf, delta, ok (9 floats). */
void orc_sample_metal(const float *ray_d, const float *n, const float *albedo, float fuzz, float ru, float rv,
                      float *out) {
    orc_sample r = sample_metal(orc_v(ray_d[0], ray_d[1], ray_d[2]), orc_v(n[0], n[1], n[2]),
                                orc_v(albedo[0], albedo[1], albedo[2]), fuzz, ru, rv);
    float o[9] = {r.dir.x, r.dir.y, r.dir.z, r.weight.x, r.weight.y, r.weight.z, r.pdf, (float)r.delta, (float)r.ok};
    memcpy(out, o, sizeof(o));
}
void orc_eval_metal(const float *ray_d, const float *n, const float *albedo, float fuzz, const float *out_dir,
                    float *out) {
    orc_eval r = eval_metal(orc_v(ray_d[0], ray_d[1], ray_d[2]), orc_v(n[0], n[1], n[2]),
                            orc_v(albedo[0], albedo[1], albedo[2]), fuzz, orc_v(out_dir[0], out_dir[1], out_dir[2]));
    float o[5] = {r.value.x, r.value.y, r.value.z, r.pdf, (float)r.delta};
    memcpy(out, o, sizeof(o));
}
void orc_sample_dielectric(const float *ray_d, const float *n, float ior, int front, float ru, float *out) {
    orc_sample r = sample_dielectric(orc_v(ray_d[0], ray_d[1], ray_d[2]), orc_v(n[0], n[1], n[2]), ior, front, ru);
    float o[9] = {r.dir.x, r.dir.y, r.dir.z, r.weight.x, r.weight.y, r.weight.z, r.pdf, (float)r.delta, (float)r.ok};
    memcpy(out, o, sizeof(o));
}
void orc_reflect3(const float *v, const float *n, float *out) {
    orc_v3 r = orc_reflect(orc_v(v[0], v[1], v[2]), orc_v(n[0], n[1], n[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* ---- light table ------------------------------------------------------------- */
#define L_STRIDE 14
/* ---- GGX conductor (the library's BAJO_MAT_GGX; bajo has no counterpart) ------
 * Trowbridge-Reitz D of width alpha = roughness, Smith height-correlated G2,
 * Schlick F with F0 = the specular color, visible-normal sampling (Heitz
 * 2018, JCGT 7(4)), frame of Duff et al. 2017 (JCGT 6(1)).  Restated
 * operation for operation in FP32 so it yields the device's bits. */
#define PI_F 3.14159265358979323846f
typedef struct { orc_v3 t, b, n; } orc_frame;
static orc_frame frame_of(orc_v3 n) {
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    orc_frame f;
    f.t = orc_v(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f.b = orc_v(b, sign + n.y * n.y * a, -n.y);
    f.n = n;
    return f;
}
static orc_v3 to_local(orc_frame f, orc_v3 v) { return orc_v(orc_dot(v, f.t), orc_dot(v, f.b), orc_dot(v, f.n)); }
static orc_v3 to_world(orc_frame f, orc_v3 v) {
    return orc_add(orc_add(orc_scale(f.t, v.x), orc_scale(f.b, v.y)), orc_scale(f.n, v.z));
}
static float ggx_d(float a2, float cos_h) {
    float c2 = cos_h * cos_h;
    float d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (PI_F * d * d);
}
static float ggx_lambda(float a2, float cos_w) {
    float c2 = cos_w * cos_w;
    float t2 = orc_fmax(1.0f - c2, 0.0f) / c2;
    return 0.5f * (sqrtf(1.0f + a2 * t2) - 1.0f);
}
static orc_v3 schlick_f0(orc_v3 f0, float c) {
    float x = 1.0f - orc_fmin(orc_fmax(c, 0.0f), 1.0f);
    float x2 = x * x;
    float x5 = x2 * x2 * x;
    return orc_v(fmaf(1.0f - f0.x, x5, f0.x), fmaf(1.0f - f0.y, x5, f0.y), fmaf(1.0f - f0.z, x5, f0.z));
}
static orc_eval eval_ggx_local(orc_v3 wo, orc_v3 wi, orc_v3 f0, float alpha) {
    orc_eval out = {orc_v(0, 0, 0), 0.0f, 0};
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) return out;
    float a2 = alpha * alpha;
    orc_v3 h = orc_normalize(orc_add(wo, wi));
    float d = ggx_d(a2, h.z);
    float lo = ggx_lambda(a2, wo.z), li = ggx_lambda(a2, wi.z);
    float g2 = 1.0f / (1.0f + lo + li);
    float g1o = 1.0f / (1.0f + lo);
    orc_v3 f = schlick_f0(f0, orc_dot(wo, h));
    out.value = orc_scale(f, d * g2 / (4.0f * wo.z * wi.z));
    out.pdf = g1o * d / (4.0f * wo.z);
    return out;
}
static orc_eval eval_ggx(orc_v3 ray_d, orc_v3 n, orc_v3 f0, float alpha, orc_v3 dir) {
    orc_frame fr = frame_of(n);
    orc_v3 wo = to_local(fr, orc_v(-ray_d.x, -ray_d.y, -ray_d.z));
    orc_v3 wi = to_local(fr, dir);
    return eval_ggx_local(orc_normalize(wo), orc_normalize(wi), f0, alpha);
}
static orc_sample sample_ggx(orc_v3 ray_d, orc_v3 n, orc_v3 f0, float alpha, float u1, float u2) {
    orc_sample out = {n, orc_v(0, 0, 0), 0.0f, 0, 0};
    orc_frame fr = frame_of(n);
    orc_v3 wo = orc_normalize(to_local(fr, orc_v(-ray_d.x, -ray_d.y, -ray_d.z)));
    if (!(wo.z > 0.0f)) return out;
    orc_v3 vh = orc_normalize(orc_v(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    float inv_len = lensq > 0.0f ? 1.0f / sqrtf(lensq) : 0.0f;
    orc_v3 t1v = lensq > 0.0f ? orc_v(-vh.y * inv_len, vh.x * inv_len, 0.0f) : orc_v(1.0f, 0.0f, 0.0f);
    orc_v3 t2v = orc_cross(vh, t1v);
    float r = sqrtf(u1);
    float sp, cp;
    orc_sincos_turns(u2, &sp, &cp);
    float t1 = r * cp;
    float t2 = r * sp;
    float sh = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - sh) * sqrtf(orc_fmax(1.0f - t1 * t1, 0.0f)) + sh * t2;
    float tz = sqrtf(orc_fmax(1.0f - t1 * t1 - t2 * t2, 0.0f));
    orc_v3 nh = orc_add(orc_add(orc_scale(t1v, t1), orc_scale(t2v, t2)), orc_scale(vh, tz));
    orc_v3 h = orc_normalize(orc_v(alpha * nh.x, alpha * nh.y, orc_fmax(nh.z, 0.0f)));
    float oh = orc_dot(wo, h);
    orc_v3 wi = orc_sub(orc_scale(h, 2.0f * oh), wo);
    if (!(wi.z > 0.0f) || !(oh > 0.0f)) return out;
    float a2 = alpha * alpha;
    float lo = ggx_lambda(a2, wo.z), li = ggx_lambda(a2, wi.z);
    orc_v3 f = schlick_f0(f0, oh);
    out.weight = orc_scale(f, (1.0f + lo) / (1.0f + lo + li));
    out.pdf = ggx_d(a2, h.z) / ((1.0f + lo) * 4.0f * wo.z);
    out.dir = orc_normalize(to_world(fr, wi));
    out.ok = out.pdf > 0.0f;
    return out;
}
/* out: value.xyz, pdf */
void orc_eval_ggx(const float *ray_d, const float *n, const float *f0, float alpha, const float *dir, float *out) {
    orc_eval r = eval_ggx(orc_v(ray_d[0], ray_d[1], ray_d[2]), orc_v(n[0], n[1], n[2]), orc_v(f0[0], f0[1], f0[2]),
                          alpha, orc_v(dir[0], dir[1], dir[2]));
    out[0] = r.value.x; out[1] = r.value.y; out[2] = r.value.z; out[3] = r.pdf;
}
/* out: dir.xyz, weight.xyz, pdf, ok */
void orc_sample_ggx(const float *ray_d, const float *n, const float *f0, float alpha, float u1, float u2, float *out) {
    orc_sample r = sample_ggx(orc_v(ray_d[0], ray_d[1], ray_d[2]), orc_v(n[0], n[1], n[2]), orc_v(f0[0], f0[1], f0[2]),
                              alpha, u1, u2);
    out[0] = r.dir.x; out[1] = r.dir.y; out[2] = r.dir.z;
    out[3] = r.weight.x; out[4] = r.weight.y; out[5] = r.weight.z;
    out[6] = r.pdf; out[7] = (float)r.ok;
}

#define MAT_LAMB 0u
#define MAT_METAL 1u
#define MAT_DIEL 2u
#define MAT_EMIT 3u
#define MAT_GGX 4u
#define IDX_MASK ((1u << 28) - 1u)

void orc_alias_table(const float *w, uint32_t n, float total, float *prob, uint32_t *alias) {
    for (uint32_t i = 0; i < n; ++i) { prob[i] = 1.0f; alias[i] = 0; }
    if (n == 0 || !(total > 0.0f)) return;
    float *scaled = (float *)malloc(n * sizeof(float));
    uint32_t *small = (uint32_t *)malloc(n * sizeof(uint32_t)), *large = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t ns = 0, nl = 0;
    for (uint32_t i = 0; i < n; ++i) {
        alias[i] = i;
        scaled[i] = w[i] * (float)n / total;
        if (scaled[i] < 1.0f) small[ns++] = i; else large[nl++] = i;
    }
    while (ns > 0 && nl > 0) {
        uint32_t si = small[--ns], li = large[--nl];
        prob[si] = scaled[si];
        alias[si] = li;
        scaled[li] = scaled[li] + scaled[si] - 1.0f;
        if (scaled[li] < 1.0f) small[ns++] = li; else large[nl++] = li;
    }
    while (ns > 0) { uint32_t i = small[--ns]; prob[i] = 1.0f; alias[i] = i; }
    while (nl > 0) { uint32_t i = large[--nl]; prob[i] = 1.0f; alias[i] = i; }
    free(scaled); free(small); free(large);
}

typedef struct {
    float *fields;  /* 14 per light */
    uint32_t *kinds;
    uint32_t count;
    float total;
} lights_t;

/* SceneData._build_light_store (rt/types.mojo:490-530): emissive triangles
 * (weight area * mean radiance), then emissive spheres (weight
 * 4 pi r^2 * mean radiance, |r|); one alias table over both. */
static void build_lights_mixed(const float *v, uint64_t n, const uint32_t *surf, const float *sph, uint64_t ns,
                               const uint32_t *ssurf, const float *emissive, lights_t *L) {
    uint64_t cnt = 0;
    for (uint64_t i = 0; i < n; ++i) if ((surf[i] >> 28) == MAT_EMIT) cnt++;
    for (uint64_t i = 0; i < ns; ++i) if ((ssurf[i] >> 28) == MAT_EMIT) cnt++;
    float *w = (float *)malloc((cnt ? cnt : 1) * sizeof(float));
    uint32_t *tri = (uint32_t *)malloc((cnt ? cnt : 1) * sizeof(uint32_t));
    uint8_t *is_s = (uint8_t *)malloc(cnt ? cnt : 1);
    uint32_t k = 0;
    float total = 0.0f;
    for (uint64_t i = 0; i < n; ++i) {
        if ((surf[i] >> 28) != MAT_EMIT) continue;
        const float *t = v + 9 * i;
        orc_v3 p0 = orc_v(t[0], t[1], t[2]), p1 = orc_v(t[3], t[4], t[5]), p2 = orc_v(t[6], t[7], t[8]);
        orc_v3 cr = orc_cross(orc_sub(p1, p0), orc_sub(p2, p0));
        float area = 0.5f * sqrtf(orc_dot(cr, cr));
        const float *e = emissive + 3 * (surf[i] & IDX_MASK);
        float imp = orc_fmax((e[0] + e[1] + e[2]) / 3.0f, 0.0f);
        float wt = area * imp;
        if (wt > 0.0f) { w[k] = wt; tri[k] = (uint32_t)i; is_s[k] = 0; total += wt; k++; }
    }
    for (uint64_t i = 0; i < ns; ++i) {
        if ((ssurf[i] >> 28) != MAT_EMIT) continue;
        float r = fabsf(sph[4 * i + 3]);
        const float *e = emissive + 3 * (ssurf[i] & IDX_MASK);
        float imp = orc_fmax((e[0] + e[1] + e[2]) / 3.0f, 0.0f);
        float wt = 4.0f * 3.14159265358979f * r * r * imp;
        if (wt > 0.0f) { w[k] = wt; tri[k] = (uint32_t)i; is_s[k] = 1; total += wt; k++; }
    }
    float *prob = (float *)malloc((k ? k : 1) * sizeof(float));
    uint32_t *alias = (uint32_t *)malloc((k ? k : 1) * sizeof(uint32_t));
    orc_alias_table(w, k, total, prob, alias);
    L->fields = (float *)calloc((k ? k : 1) * L_STRIDE, sizeof(float));
    L->kinds = (uint32_t *)calloc(k ? k : 1, sizeof(uint32_t));
    for (uint32_t l = 0; l < k; ++l) {
        float *o = L->fields + (size_t)l * L_STRIDE;
        const float *e;
        if (is_s[l]) {
            const float *sp = sph + 4 * (uint64_t)tri[l];
            e = emissive + 3 * (ssurf[tri[l]] & IDX_MASK);
            o[0] = sp[0]; o[1] = sp[1]; o[2] = sp[2];
            o[9] = fabsf(sp[3]);
        } else {
            const float *t = v + 9 * (uint64_t)tri[l];
            e = emissive + 3 * (surf[tri[l]] & IDX_MASK);
            for (int j = 0; j < 9; ++j) o[j] = t[j];
        }
        o[10] = e[0]; o[11] = e[1]; o[12] = e[2];
        o[13] = prob[l];
        L->kinds[l] = (alias[l] << 4) | (is_s[l] ? 0u : 1u);
    }
    L->count = k;
    L->total = total;
    free(w); free(tri); free(is_s); free(prob); free(alias);
}


void orc_light_table(const float *v, uint64_t n, const uint32_t *surf, const float *emissive, float *fields,
                     uint32_t *kinds, uint32_t *count, float *total) {
    orc_light_table_mixed(v, n, surf, NULL, 0, NULL, emissive, fields, kinds, count, total);
}

void orc_light_table_mixed(const float *v, uint64_t n, const uint32_t *surf, const float *sph, uint64_t ns,
                           const uint32_t *ssurf, const float *emissive, float *fields, uint32_t *kinds,
                           uint32_t *count, float *total) {
    lights_t L;
    build_lights_mixed(v, n, surf, sph, ns, ssurf, emissive, &L);
    if (fields) memcpy(fields, L.fields, (size_t)L.count * L_STRIDE * sizeof(float));
    if (kinds) memcpy(kinds, L.kinds, (size_t)L.count * sizeof(uint32_t));
    *count = L.count;
    *total = L.total;
    free(L.fields); free(L.kinds);
}

/* ---- one path ------------------------------------------------------------------- */
typedef struct {
    orc_bvh bvh;
    const uint32_t *surf;
    const float *lamb, *metal, *diel, *emit;
    lights_t lights;
    orc_render_settings st;
    float cam[20];
    /* spheres (rt/gpu/scene_trace.mojo:113-193): their BVH over |r|, the
     * signed radii (r < 0 turns the normal inward) and their surfaces */
    uint64_t ns;
    orc_bvh sbvh;
    const float *srad;
    const uint32_t *ssurf;
} scene_t;

/* Closest hit of a mixed scene: spheres first, then triangles with t_max =
 * the sphere hit (a triangle must be strictly closer); *surf receives the
 * hit's surface id and the normal is the outward one of bajo's geometry. */
static orc_hit scene_closest(const scene_t *sc, orc_v3 o, orc_v3 d, float tmin, uint32_t *surf) {
    if (!sc->ns) {
        orc_hit h = orc_trace_closest(&sc->bvh, o, d, tmin, ORC_FLT_MAX, NULL);
        if (h.prim != ORC_EMPTY) *surf = sc->surf[h.prim];
        return h;
    }
    orc_hit hs = orc_trace_closest(&sc->sbvh, o, d, tmin, ORC_FLT_MAX, NULL);
    float tmax = hs.prim != ORC_EMPTY ? hs.t : ORC_FLT_MAX;
    orc_hit ht = sc->bvh.prim_count ? orc_trace_closest(&sc->bvh, o, d, tmin, tmax, NULL) : hs;
    if (sc->bvh.prim_count && ht.prim != ORC_EMPTY && ht.t < tmax) {
        *surf = sc->surf[ht.prim];
        return ht;
    }
    if (hs.prim != ORC_EMPTY) {
        if (sc->srad[hs.prim] < 0.0f) hs.n = orc_v(-hs.n.x, -hs.n.y, -hs.n.z);
        *surf = sc->ssurf[hs.prim];
    }
    return hs;
}

/* Occlusion of a mixed scene: spheres, then triangles. */
static int scene_occluded(const scene_t *sc, orc_v3 o, orc_v3 d, float tmin, float tmax) {
    if (sc->ns) {
        orc_hit hs = orc_trace_any(&sc->sbvh, o, d, tmin, tmax);
        if (hs.prim == ORC_EMPTY && hs.t == 0.0f) return 1;
        if (!sc->bvh.prim_count) return 0;
    }
    orc_hit h = orc_trace_any(&sc->bvh, o, d, tmin, tmax);
    return h.prim == ORC_EMPTY && h.t == 0.0f;
}

static float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b, d = a2 + b2;
    return d > 0.0f ? a2 / d : 0.0f;
}

static orc_v3 sky(orc_v3 d) {
    orc_v3 u = orc_normalize(d);
    float a = 0.5f * (u.y + 1.0f);
    return orc_add(orc_scale(orc_v(1.0f, 1.0f, 1.0f), 1.0f - a), orc_scale(orc_v(0.5f, 0.7f, 1.0f), a));
}

static void acc(float *s, orc_v3 c) { s[0] += c.x; s[1] += c.y; s[2] += c.z; }

/* Returns 1 with a shadow ray + contribution, rt/gpu/path_shading.mojo:261-403. */
static int direct_light(const scene_t *sc, uint32_t path_id, uint32_t bounce, orc_v3 point, orc_v3 normal,
                        uint32_t kind, uint32_t mi, orc_v3 ray_d, orc_v3 T, orc_v3 *dir_out, float *tmax_out,
                        orc_v3 *contrib) {
    const lights_t *L = &sc->lights;
    if (L->count == 0 || !(L->total > 0.0f)) return 0;
    orc_rng rng;
    rng_init(&rng, sc->st.seed, path_id, stage_light(bounce));
    float as = rng_next(&rng) * (float)L->count;
    uint32_t sel = (uint32_t)as;
    if (sel >= L->count) sel = L->count - 1;
    float frac = as - (float)sel;
    uint32_t packed = L->kinds[sel];
    if (frac > L->fields[sel * L_STRIDE + 13]) sel = packed >> 4;
    const float *f = L->fields + (size_t)sel * L_STRIDE;
    orc_v3 p0 = orc_v(f[0], f[1], f[2]), ln, lp;
    if ((L->kinds[sel] & 0xFu) == 0u) {
        /* sphere light: random_unit_vector (core/random.mojo:39-45) as the light normal */
        float ss, cs;
        orc_sincos_turns(rng_next(&rng), &ss, &cs);
        float z = 1.0f - 2.0f * rng_next(&rng);
        float r = sqrtf(1.0f - z * z);
        ln = orc_v(r * cs, r * ss, z);
        lp = orc_add(p0, orc_scale(ln, f[9]));
    } else {
        orc_v3 p1 = orc_v(f[3], f[4], f[5]), p2 = orc_v(f[6], f[7], f[8]);
        orc_v3 e1 = orc_sub(p1, p0), e2 = orc_sub(p2, p0);
        orc_v3 av = orc_cross(e1, e2);
        float tas = orc_dot(av, av);
        if (tas <= 0.0f) return 0;
        float sq = sqrtf(tas);
        ln = orc_v(av.x / sq, av.y / sq, av.z / sq);
        float root_u = sqrtf(rng_next(&rng));
        float b1 = root_u * (1.0f - rng_next(&rng));
        float b2 = root_u - b1;
        lp = orc_add(orc_add(p0, orc_scale(e1, b1)), orc_scale(e2, b2));
    }
    orc_v3 tl = orc_sub(lp, point);
    float d2 = orc_dot(tl, tl);
    if (d2 <= 1.0e-8f) return 0;
    float dist = sqrtf(d2);
    orc_v3 dir = orc_v(tl.x / dist, tl.y / dist, tl.z / dist);
    float scos = orc_fmax(orc_dot(normal, dir), 0.0f);
    float lcos = orc_fmax(orc_dot(ln, orc_v(-dir.x, -dir.y, -dir.z)), 0.0f);
    if (scos <= 0.0f || lcos <= 0.0f) return 0;
    float tmax = dist - 0.002f;
    if (tmax <= 0.001f) return 0;
    orc_v3 E = orc_v(f[10], f[11], f[12]);
    float light_pdf = d2 * (E.x + E.y + E.z) / (3.0f * lcos * L->total);
    orc_v3 value;
    float bsdf_pdf;
    if (kind == MAT_METAL) {
        const float *m = sc->metal + 4 * (size_t)mi;
        orc_eval ev = eval_metal(ray_d, normal, orc_v(m[0], m[1], m[2]), m[3], dir);
        value = ev.value;
        bsdf_pdf = ev.pdf;
    } else if (kind == MAT_GGX) {
        const float *g = sc->st.ggx + 4 * (size_t)mi;
        orc_eval ev = eval_ggx(ray_d, normal, orc_v(g[0], g[1], g[2]), g[3], dir);
        value = ev.value;
        bsdf_pdf = ev.pdf;
    } else {
        const float *a = sc->lamb + 3 * (size_t)mi;
        value = orc_scale(orc_v(a[0], a[1], a[2]), 0.3183098861837907f);
        bsdf_pdf = orc_fmax(orc_dot(normal, orc_normalize(dir)), 0.0f) * 0.3183098861837907f;
    }
    int ok = light_pdf > 0.0f && bsdf_pdf > 0.0f;
    float w = sc->st.algorithm == 4 ? power_heuristic(light_pdf, bsdf_pdf) : 1.0f;
    float s = ok ? scos * w / light_pdf : 0.0f;
    if (!(s > 0.0f)) return 0;
    *contrib = orc_scale(orc_mul(orc_mul(T, value), E), s);
    *dir_out = dir;
    *tmax_out = tmax;
    return 1;
}

static void render_path(const scene_t *sc, uint32_t path_id, float *sample, uint64_t *rays) {
    const orc_render_settings *st = &sc->st;
    const float *k = sc->cam;
    uint32_t pixel = path_id / st->spp, px = pixel % st->width, py = pixel / st->width;
    orc_rng rng;
    rng_init(&rng, st->seed, path_id, 0u);
    float s, c;
    orc_sincos_turns(rng_next(&rng), &s, &c);
    float rr = sqrtf(rng_next(&rng));
    float lu = rr * c, lv = rr * s;
    float pu = rng_next(&rng), pv = rng_next(&rng);
    float fw = (float)st->width, fh = (float)st->height;
    float aspect = fw / fh;
    float sx = (((float)px + pu) / fw) * 2.0f - 1.0f;
    float sy = 1.0f - (((float)py + pv) / fh) * 2.0f;
    orc_v3 origin = orc_v(k[0], k[1], k[2]);
    orc_v3 plane = orc_add(orc_add(orc_v(k[3], k[4], k[5]), orc_scale(orc_v(k[6], k[7], k[8]), sx * aspect * k[12])),
                           orc_scale(orc_v(k[9], k[10], k[11]), sy * k[12]));
    orc_v3 focal = orc_add(origin, orc_scale(plane, k[13]));
    orc_v3 ro = orc_add(orc_add(origin, orc_scale(orc_v(k[14], k[15], k[16]), lu)), orc_scale(orc_v(k[17], k[18], k[19]), lv));
    orc_v3 o = ro, d = orc_normalize(orc_sub(focal, ro));
    float tmin = 0.001f;
    orc_v3 T = orc_v(1.0f, 1.0f, 1.0f);
    float path_pdf = 0.0f;
    int delta = 1;
    int alg = st->algorithm;
    uint32_t depth = (alg == 1 || alg == 2) ? 1u : st->max_depth;
    sample[0] = sample[1] = sample[2] = 0.0f;
    for (uint32_t b = 0; b < depth; ++b) {
        if (rays) rays[0]++;
        uint32_t surf = 0;
        orc_hit h = scene_closest(sc, o, d, tmin, &surf);
        int found = h.prim != ORC_EMPTY && h.t < ORC_FLT_MAX;
        int shadow = 0, next = 0;
        orc_v3 sh_o = orc_v(0, 0, 0), sh_d = orc_v(0, 0, 0), sh_c = orc_v(0, 0, 0);
        float sh_tmax = 0.0f;
        orc_v3 no = orc_v(0, 0, 0), nd = orc_v(0, 0, 0), nT = T;
        float npdf = 0.0f;
        int ndelta = 0;
        if (!found) {
            if (alg == 0 || alg == 3 || alg == 4) acc(sample, orc_mul(T, sky(d)));
        } else {
            int front = orc_dot(d, h.n) < 0.0f;
            orc_v3 n = front ? h.n : orc_v(-h.n.x, -h.n.y, -h.n.z);
            orc_v3 point = orc_add(o, orc_scale(d, h.t));
            uint32_t kind = surf >> 28, mi = surf & IDX_MASK;
            if (alg == 1) {
                acc(sample, orc_scale(orc_add(n, orc_v(1.0f, 1.0f, 1.0f)), 0.5f));
            } else if (alg == 2) {
                orc_rng ar;
                rng_init(&ar, st->seed, path_id, 1u);
                float as_, ac_;
                orc_sincos_turns(rng_next(&ar), &as_, &ac_);
                float z = 1.0f - 2.0f * rng_next(&ar);
                float r = sqrtf(1.0f - z * z);
                orc_v3 dir = orc_v(r * ac_, r * as_, z);
                if (orc_dot(dir, n) < 0.0f) dir = orc_v(-dir.x, -dir.y, -dir.z);
                shadow = 1; sh_o = point; sh_d = dir; sh_tmax = 4.0f;
            } else {
                orc_v3 albedo = orc_v(0, 0, 0);
                if (kind == MAT_LAMB) albedo = orc_v(sc->lamb[3 * mi], sc->lamb[3 * mi + 1], sc->lamb[3 * mi + 2]);
                if ((alg == 3 || alg == 4) && (kind == MAT_LAMB || kind == MAT_METAL || kind == MAT_GGX)) {
                    shadow = direct_light(sc, path_id, b, point, n, kind, mi, d, T, &sh_d, &sh_tmax, &sh_c);
                    sh_o = point;
                }
                if (kind == MAT_EMIT) {
                    if (front) {
                        orc_v3 E = orc_v(sc->emit[3 * mi], sc->emit[3 * mi + 1], sc->emit[3 * mi + 2]);
                        float w = 1.0f;
                        if (alg == 3 && b > 0 && !delta) w = 0.0f;
                        if (alg == 4 && b > 0 && !delta) {
                            float lc = orc_fmax(orc_dot(n, orc_normalize(orc_v(-d.x, -d.y, -d.z))), 0.0f);
                            float lpdf = 0.0f;
                            if (lc > 0.0f && sc->lights.total > 0.0f) {
                                float d2 = h.t * h.t * orc_dot(d, d);
                                lpdf = d2 * (E.x + E.y + E.z) / (3.0f * lc * sc->lights.total);
                            }
                            w = power_heuristic(path_pdf, lpdf);
                        }
                        acc(sample, orc_scale(orc_mul(T, E), w));
                    }
                } else if (kind == MAT_LAMB) {
                    orc_rng br;
                    rng_init(&br, st->seed, path_id, stage_bsdf(b));
                    float ru = rng_next(&br), rv = rng_next(&br);
                    float ss, cc;
                    orc_sincos_turns(ru, &ss, &cc);
                    float z = 1.0f - 2.0f * rv;
                    float radius = sqrtf(orc_fmax(1.0f - z * z, 0.0f));
                    orc_v3 sd = orc_add(n, orc_v(radius * cc, radius * ss, z));
                    if (fabsf(sd.x) < 1e-8f && fabsf(sd.y) < 1e-8f && fabsf(sd.z) < 1e-8f) sd = n;
                    orc_v3 dir = orc_normalize(sd);
                    float pdf = orc_fmax(orc_dot(n, orc_normalize(dir)), 0.0f) * 0.3183098861837907f;
                    orc_v3 Tn = orc_mul(T, albedo);
                    int alive = 1;
                    if (b + 1u >= 5u) {
                        float mx = orc_fmax(Tn.x, orc_fmax(Tn.y, Tn.z));
                        if (mx <= 0.0f) alive = 0;
                        else {
                            float survival = orc_fmin(orc_fmax(mx, 0.05f), 0.95f);
                            orc_rng rr2;
                            rng_init(&rr2, st->seed, path_id, stage_roulette(b));
                            if (rng_next(&rr2) >= survival) alive = 0;
                            else Tn = orc_v(Tn.x / survival, Tn.y / survival, Tn.z / survival);
                        }
                    }
                    if (alive) {
                        next = 1;
                        no = orc_v(o.x + h.t * d.x, o.y + h.t * d.y, o.z + h.t * d.z);
                        nd = dir; nT = Tn; npdf = pdf; ndelta = 0;
                    }
                } else if (kind == MAT_METAL || kind == MAT_DIEL || kind == MAT_GGX) {
                    /* rt/gpu/path_shading.mojo:669-783: numbers drawn only where used */
                    orc_rng br;
                    rng_init(&br, st->seed, path_id, stage_bsdf(b));
                    orc_sample smp;
                    if (kind == MAT_GGX) {
                        const float *g = st->ggx + 4 * (size_t)mi;
                        float ru = rng_next(&br), rv = rng_next(&br);
                        smp = sample_ggx(d, n, orc_v(g[0], g[1], g[2]), g[3], ru, rv);
                    } else if (kind == MAT_METAL) {
                        const float *m = sc->metal + 4 * (size_t)mi;
                        float ru = 0.0f, rv = 0.0f;
                        if (m[3] > 1.0e-4f) { ru = rng_next(&br); rv = rng_next(&br); }
                        smp = sample_metal(d, n, orc_v(m[0], m[1], m[2]), m[3], ru, rv);
                    } else {
                        float ior = sc->diel[mi];
                        float ri = front ? 1.0f / ior : ior;
                        orc_v3 u = orc_normalize(d);
                        float cos_t = orc_fmin(orc_dot(orc_v(-u.x, -u.y, -u.z), n), 1.0f);
                        float sin_t = sqrtf(orc_fmax(1.0f - cos_t * cos_t, 0.0f));
                        float ru = ri * sin_t <= 1.0f ? rng_next(&br) : 0.0f;
                        smp = sample_dielectric(d, n, ior, front, ru);
                    }
                    if (smp.ok) {
                        orc_v3 Tn = orc_mul(T, smp.weight);
                        int alive = 1;
                        if (b + 1u >= 5u) {
                            float mx = orc_fmax(Tn.x, orc_fmax(Tn.y, Tn.z));
                            if (mx <= 0.0f) alive = 0;
                            else {
                                float survival = orc_fmin(orc_fmax(mx, 0.05f), 0.95f);
                                orc_rng rr2;
                                rng_init(&rr2, st->seed, path_id, stage_roulette(b));
                                if (rng_next(&rr2) >= survival) alive = 0;
                                else Tn = orc_v(Tn.x / survival, Tn.y / survival, Tn.z / survival);
                            }
                        }
                        if (alive) {
                            next = 1;
                            no = orc_v(o.x + h.t * d.x, o.y + h.t * d.y, o.z + h.t * d.z);
                            nd = smp.dir; nT = Tn; npdf = smp.pdf; ndelta = smp.delta;
                        }
                    }
                }
            }
        }
        if (shadow) {
            if (rays) rays[1]++;
            int occluded = scene_occluded(sc, sh_o, sh_d, 0.001f, sh_tmax);
            if (alg == 2) { float v = occluded ? 0.08f : 1.0f; acc(sample, orc_v(v, v, v)); }
            else if (!occluded) acc(sample, sh_c);
        }
        if (!next) break;
        o = no; d = nd; T = nT; path_pdf = npdf; tmin = 0.001f; delta = ndelta;
    }
}

typedef struct { const scene_t *sc; uint64_t first, last; float *samples; uint64_t rays[2]; } job_t;

static void *render_worker(void *p) {
    job_t *j = (job_t *)p;
    for (uint64_t id = j->first; id < j->last; ++id) render_path(j->sc, (uint32_t)id, j->samples + 3 * id, j->rays);
    return NULL;
}

int orc_render(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *lambertian,
               const float *metal, const float *dielectric, const float *emissive, const orc_render_settings *st,
               const float *cam20, float *pixels, uint64_t *ray_counts, int threads) {
    return orc_render_mixed(vertices, n, surfaces, NULL, 0, NULL, lambertian, metal, dielectric, emissive, st, cam20,
                            pixels, ray_counts, threads);
}

/* The sphere half of a mixed scene: the LBVH over the spheres' boxes
 * (|r|), collapsed to the render's node width; returns 0 on failure. */
typedef struct { float *abs4, *boxes, *nb; uint64_t *keys; uint32_t *ids, *meta, *parent, *counts; void *wide; } sph_bvh_t;

static int build_sphere_bvh(const float *sph, uint64_t ns, const orc_render_settings *st, orc_bvh *b, sph_bvh_t *t) {
    uint64_t m = ns > 1 ? ns - 1 : 1;
    memset(t, 0, sizeof(*t));
    t->abs4 = (float *)malloc(4 * ns * sizeof(float));
    t->boxes = (float *)malloc(6 * ns * sizeof(float));
    for (uint64_t i = 0; i < ns; ++i) {
        float r = fabsf(sph[4 * i + 3]);
        for (int k = 0; k < 3; ++k) {
            t->abs4[4 * i + k] = sph[4 * i + k];
            t->boxes[6 * i + k] = sph[4 * i + k] - r;
            t->boxes[6 * i + 3 + k] = sph[4 * i + k] + r;
        }
        t->abs4[4 * i + 3] = r;
    }
    t->keys = (uint64_t *)malloc(ns * sizeof(uint64_t));
    t->ids = (uint32_t *)malloc(ns * sizeof(uint32_t));
    t->meta = (uint32_t *)malloc(4 * m * sizeof(uint32_t));
    t->parent = (uint32_t *)malloc(ns * sizeof(uint32_t));
    t->counts = (uint32_t *)malloc(m * sizeof(uint32_t));
    t->nb = (float *)malloc(12 * m * sizeof(float));
    float scene12[12];
    uint32_t root = orc_build_lbvh_bounds(t->boxes, ns, t->keys, t->ids, t->meta, t->parent, t->counts, t->nb, scene12);
    memset(b, 0, sizeof(*b));
    b->leaf_ids = t->ids; b->node_meta = t->meta; b->node_bounds = t->nb; b->root = root;
    b->prim_count = ns; b->spheres = t->abs4;
    if (st->node_width >= 4) {
        t->wide = malloc(ns * 96);
        uint64_t wc = orc_collapse_wide_n(t->meta, t->nb, t->counts, ns, root, scene12,
                                          st->leaf_size ? (uint32_t)st->leaf_size : 4u, (uint32_t)st->node_width,
                                          t->wide, ns);
        if (wc == 0) return 0;
        b->wide = t->wide;
        b->wide_count = wc;
    }
    return 1;
}

static void free_sphere_bvh(sph_bvh_t *t) {
    free(t->abs4); free(t->boxes); free(t->nb); free(t->keys); free(t->ids); free(t->meta); free(t->parent);
    free(t->counts); free(t->wide);
}

/* Timing of the CPU baseline (not used by parity checks): renders repeat
 * g_timing_reps times after a warm-up and keep the best wall time. */
static int g_timing_reps = 0;
static double g_timing_best = 0.0;
void orc_render_set_timing(int reps) { g_timing_reps = reps; }
double orc_render_best_seconds(void) { return g_timing_best; }

int orc_render_mixed(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *spheres, uint64_t ns,
                     const uint32_t *sphere_surfaces, const float *lambertian, const float *metal,
                     const float *dielectric, const float *emissive, const orc_render_settings *st,
                     const float *cam20, float *pixels, uint64_t *ray_counts, int threads) {
    if (n == 0 && ns == 0) return -3;
    uint64_t m = n > 1 ? n - 1 : 1;
    uint64_t *keys = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *ids = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *meta = (uint32_t *)malloc(4 * m * sizeof(uint32_t));
    uint32_t *parent = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *counts = (uint32_t *)malloc(m * sizeof(uint32_t));
    float *nb = (float *)malloc(12 * m * sizeof(float));
    float scene12[12];
    uint32_t root = n ? orc_build_lbvh(vertices, n, keys, ids, meta, parent, counts, nb, scene12) : ORC_SENTINEL;
    if (st->method == 1 && n > 1) {
        /* H-PLOC topology over the same sorted keys, as a depth-first binary layout */
        uint64_t m2 = 2 * n - 1;
        float *tb = (float *)malloc(6 * n * sizeof(float));
        uint32_t *codes = (uint32_t *)malloc(n * sizeof(uint32_t));
        uint32_t *tp = (uint32_t *)malloc(5 * m2 * sizeof(uint32_t));
        float *tbox = (float *)malloc(6 * m2 * sizeof(float));
        orc_triangle_bounds(vertices, n, tb);
        for (uint64_t i = 0; i < n; ++i) codes[i] = (uint32_t)(keys[i] >> 32);
        uint32_t troot = orc_hploc_topology(tb, codes, ids, n, 8, 16, tp, tp + m2, tp + 2 * m2, tp + 3 * m2,
                                            tp + 4 * m2, tbox, NULL);
        uint32_t *num = (uint32_t *)malloc(m2 * sizeof(uint32_t));
        if (troot != ORC_SENTINEL) {
            /* the device builder's inner-node numbering (from the sorted ids, before they are reordered) */
            orc_hploc_lowest_leaf_numbering(tp + m2, tp + 2 * m2, ids, n, num);
            root = orc_hploc_binary(troot, tp + m2, tp + 2 * m2, tp + 3 * m2, tp + 4 * m2, tbox, n, num, meta, nb,
                                    ids, counts, parent);
        }
        free(tb); free(codes); free(tp); free(tbox); free(num);
        if (troot == ORC_SENTINEL) { free(keys); free(ids); free(meta); free(parent); free(counts); free(nb); return -2; }
    }
    scene_t sc;
    memset(&sc.bvh, 0, sizeof(sc.bvh));
    sc.bvh.vertices = vertices; sc.bvh.leaf_ids = ids; sc.bvh.node_meta = meta; sc.bvh.node_bounds = nb; sc.bvh.root = root;
    sc.bvh.prim_count = n;
    void *wide = NULL;
    if (st->node_width >= 4 && n > 0) {
        wide = malloc(n * 96);
        uint64_t wc = orc_collapse_wide_n(meta, nb, counts, n, root, scene12,
                                          st->leaf_size ? (uint32_t)st->leaf_size : 4u, (uint32_t)st->node_width, wide,
                                          n);
        if (wc == 0) { free(wide); free(keys); free(ids); free(meta); free(parent); free(counts); free(nb); return -1; }
        sc.bvh.wide = wide;
        sc.bvh.wide_count = wc;
    }
    sc.surf = surfaces; sc.lamb = lambertian; sc.metal = metal; sc.diel = dielectric; sc.emit = emissive; sc.st = *st;
    memcpy(sc.cam, cam20, sizeof(sc.cam));
    sc.ns = ns;
    sc.srad = spheres;
    sc.ssurf = sphere_surfaces;
    sph_bvh_t sph_store;
    memset(&sph_store, 0, sizeof(sph_store));
    memset(&sc.sbvh, 0, sizeof(sc.sbvh));
    if (ns && !build_sphere_bvh(spheres, ns, st, &sc.sbvh, &sph_store)) {
        free_sphere_bvh(&sph_store);
        free(wide); free(keys); free(ids); free(meta); free(parent); free(counts); free(nb);
        return -1;
    }
    if (ns) {
        /* signed radii are read per sphere from the caller's array: stride 4, radius at [3] */
        float *sr = (float *)malloc(ns * sizeof(float));
        for (uint64_t i = 0; i < ns; ++i) sr[i] = spheres[4 * i + 3];
        sc.srad = sr;
    }
    build_lights_mixed(vertices, n, surfaces, spheres, ns, sphere_surfaces, emissive, &sc.lights);
    uint64_t total = (uint64_t)st->width * st->height * st->spp;
    float *samples = (float *)malloc(3 * total * sizeof(float));
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    job_t jobs[256];
    uint64_t rays[2] = {0, 0};
    /* with timing on: one warm-up render, then g_timing_reps timed renders of
     * the same frame over the built hierarchy (the best is kept) */
    const int runs = g_timing_reps > 0 ? 1 + g_timing_reps : 1;
    g_timing_best = 0.0;
    for (int run = 0; run < runs; ++run) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (int t = 0; t < threads; ++t) {
            job_t j = {&sc, total * (uint64_t)t / (uint64_t)threads, total * (uint64_t)(t + 1) / (uint64_t)threads, samples, {0, 0}};
            jobs[t] = j;
            pthread_create(&tid[t], NULL, render_worker, &jobs[t]);
        }
        rays[0] = rays[1] = 0;
        for (int t = 0; t < threads; ++t) { pthread_join(tid[t], NULL); rays[0] += jobs[t].rays[0]; rays[1] += jobs[t].rays[1]; }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        const double dt = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        if (run > 0 || runs == 1) g_timing_best = (g_timing_best == 0.0 || dt < g_timing_best) ? dt : g_timing_best;
    }
    float kk = 1.0f / (float)st->spp;
    for (uint64_t p = 0; p < (uint64_t)st->width * st->height; ++p) {
        float r = 0, g = 0, b = 0;
        for (uint64_t s = p * st->spp; s < (p + 1) * st->spp; ++s) { r += samples[3 * s]; g += samples[3 * s + 1]; b += samples[3 * s + 2]; }
        pixels[3 * p] = r * kk; pixels[3 * p + 1] = g * kk; pixels[3 * p + 2] = b * kk;
    }
    if (ray_counts) { ray_counts[0] = rays[0]; ray_counts[1] = rays[1]; }
    free(samples); free(keys); free(ids); free(meta); free(parent); free(counts); free(nb); free(wide);
    free(sc.lights.fields); free(sc.lights.kinds);
    if (ns) free((void *)sc.srad);
    free_sphere_bvh(&sph_store);
    return 0;
}

/* ---- FP64 reference estimator ---------------------------------------------------- */
typedef struct { uint64_t state, inc; } pcg_t;
static uint32_t pcg_next(pcg_t *r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u), rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
static double pcg_d(pcg_t *r) { return (double)(pcg_next(r) >> 5) * (1.0 / 134217728.0); }

typedef struct { double x, y, z; } d3;
static d3 D(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static d3 dadd(d3 a, d3 b) { return D(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 dsub(d3 a, d3 b) { return D(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 dsc(d3 a, double s) { return D(a.x * s, a.y * s, a.z * s); }
static d3 dmul(d3 a, d3 b) { return D(a.x * b.x, a.y * b.y, a.z * b.z); }
static double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 dcross(d3 a, d3 b) { return D(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static d3 dnorm(d3 a) { double l = sqrt(ddot(a, a)); return l > 0 ? dsc(a, 1.0 / l) : a; }

typedef struct {
    const float *v; uint64_t n; const uint32_t *surf; const float *lamb, *metal, *diel, *emit;
    double *lw; uint32_t *ltri; uint32_t nl; double total;
    orc_render_settings st; const float *cam;
    const float *sph; uint64_t ns; const uint32_t *ssurf; uint8_t *lsph;  /* spheres (signed radii), light kind */
} scene64;

#define SPH_BIT (1ULL << 63)

/* Russian roulette from depth 5 in double (survival max(T) in [0.05, 0.95]). */
static int roulette64(pcg_t *rng, uint32_t b, d3 *T) {
    if (b + 1 < 5) return 1;
    double mx = fmax(T->x, fmax(T->y, T->z));
    if (mx <= 0) return 0;
    double surv = fmin(fmax(mx, 0.05), 0.95);
    if (pcg_d(rng) >= surv) return 0;
    *T = dsc(*T, 1.0 / surv);
    return 1;
}

/* Metal and dielectric scattering in double precision, written from the
 * definitions (not from the FP32 code): a Phong lobe of exponent
 * e = 2/fuzz^2 - 2 around the mirror direction sampled by cos = u^(1/(e+1)),
 * weight = albedo (e+2)/(e+1) cos_n (a mirror below fuzz 1e-4: weight albedo);
 * glass: Schlick's reflectance chooses reflection, else Snell refraction,
 * weight 1.  Returns 0 when the path is absorbed. */
static int scatter64(const scene64 *s, uint32_t kind, uint32_t mi, d3 d, d3 n, int front, double u1, double u2,
                     d3 *dir, d3 *weight) {
    d3 u = dnorm(d);
    double un = ddot(u, n);
    d3 refl = dnorm(dsub(u, dsc(n, 2 * un)));
    if (kind == MAT_METAL) {
        const float *m = s->metal + 4 * (size_t)mi;
        d3 alb = D(m[0], m[1], m[2]);
        double fuzz = m[3];
        if (fuzz <= 1.0e-4) {
            if (ddot(refl, n) <= 0) return 0;
            *dir = refl;
            *weight = alb;
            return 1;
        }
        double e = fmax(2.0 / (fuzz * fuzz) - 2.0, 0.0);
        double ct = pow(u1, 1.0 / (e + 1.0)), st = sqrt(fmax(1 - ct * ct, 0.0)), ph = 2 * M_PI * u2;
        d3 helper = fabs(refl.y) > 0.99 ? D(1, 0, 0) : D(0, 1, 0);
        d3 t = dnorm(dcross(helper, refl)), bt = dcross(refl, t);
        d3 w = dnorm(dadd(dadd(dsc(t, cos(ph) * st), dsc(bt, sin(ph) * st)), dsc(refl, ct)));
        double cn = ddot(n, w);
        if (cn <= 0 || ddot(refl, w) <= 0) return 0;
        *dir = w;
        *weight = dsc(alb, (e + 2.0) / (e + 1.0) * cn);
        return 1;
    }
    double ior = s->diel[mi];
    double ri = front ? 1.0 / ior : ior;
    double ci = fmin(-un, 1.0), si = sqrt(fmax(1 - ci * ci, 0.0));
    double r0 = (1 - ri) / (1 + ri);
    r0 *= r0;
    double R = r0 + (1 - r0) * pow(1 - ci, 5);
    if (ri * si > 1.0 || R > u1) {
        *dir = refl;
    } else {
        d3 perp = dsc(dadd(u, dsc(n, ci)), ri);
        d3 par = dsc(n, -sqrt(fabs(1 - ddot(perp, perp))));
        *dir = dnorm(dadd(perp, par));
    }
    *weight = D(1, 1, 1);
    return 1;
}

/* brute-force closest hit in double (eager MT, t in (tmin, tmax)) */
static int hit64(const scene64 *s, d3 o, d3 d, double tmin, double tmax, double *t_out, uint64_t *prim, d3 *nrm) {
    int found = 0;
    double best = tmax;
    for (uint64_t i = 0; i < s->n; ++i) {
        const float *t = s->v + 9 * i;
        d3 v0 = D(t[0], t[1], t[2]), e1 = dsub(D(t[3], t[4], t[5]), v0), e2 = dsub(D(t[6], t[7], t[8]), v0);
        d3 p = dcross(d, e2);
        double det = ddot(e1, p);
        if (fabs(det) < 1e-14) continue;
        double inv = 1.0 / det;
        d3 tv = dsub(o, v0);
        double u = ddot(tv, p) * inv;
        if (u < 0 || u > 1) continue;
        d3 q = dcross(tv, e1);
        double vv = ddot(d, q) * inv;
        if (vv < 0 || u + vv > 1) continue;
        double tt = ddot(e2, q) * inv;
        if (tt > tmin && tt < best) { best = tt; found = 1; *prim = i; *nrm = dnorm(dcross(e1, e2)); }
    }
    /* spheres: the nearer root in (tmin, best), else the farther one; the
     * outward normal (p - c) / r turns inward for a negative radius */
    for (uint64_t i = 0; i < s->ns; ++i) {
        const float *sp = s->sph + 4 * i;
        d3 c = D(sp[0], sp[1], sp[2]);
        double r = sp[3], ar = fabs(r);
        d3 oc = dsub(o, c);
        double a = ddot(d, d), hb = ddot(oc, d), cc = ddot(oc, oc) - ar * ar;
        double disc = hb * hb - a * cc;
        if (disc < 0) continue;
        double sq = sqrt(disc), t0 = (-hb - sq) / a, t1 = (-hb + sq) / a, tt;
        if (t0 > tmin && t0 < best) tt = t0;
        else if (t1 > tmin && t1 < best) tt = t1;
        else continue;
        best = tt; found = 1; *prim = SPH_BIT | i;
        d3 nn = dnorm(dsub(dadd(o, dsc(d, tt)), c));
        *nrm = r < 0 ? dsc(nn, -1) : nn;
    }
    *t_out = best;
    return found;
}

/* GGX in double, from the definitions: D(h) = a^2 / (pi ((n.h)^2 (a^2 - 1) + 1)^2),
 * Lambda(w) = (sqrt(1 + a^2 tan^2) - 1) / 2, G2 = 1 / (1 + L(o) + L(i)),
 * F = F0 + (1 - F0)(1 - o.h)^5, f = F D G2 / (4 cos_o cos_i); directions from
 * the visible normals (Heitz 2018) with weight F G2 / G1(o). */
static double ggx64_lambda(double a2, double c) {
    double t2 = fmax(1 - c * c, 0.0) / (c * c);
    return 0.5 * (sqrt(1 + a2 * t2) - 1);
}
static void basis64(d3 n, d3 *t, d3 *b) {
    d3 h = fabs(n.x) > 0.5 ? D(0, 1, 0) : D(1, 0, 0);
    *t = dnorm(dcross(h, n));
    *b = dcross(n, *t);
}
static d3 ggx64_f(const float *g, d3 n, d3 wo, d3 wi) {
    double co = ddot(n, wo), ci = ddot(n, wi);
    if (co <= 0 || ci <= 0) return D(0, 0, 0);
    double a2 = (double)g[3] * g[3];
    d3 h = dnorm(dadd(wo, wi));
    double ch = ddot(n, h), dd = ch * ch * (a2 - 1) + 1;
    double Dh = a2 / (M_PI * dd * dd);
    double G2 = 1.0 / (1 + ggx64_lambda(a2, co) + ggx64_lambda(a2, ci));
    double x = 1 - fmin(fmax(ddot(wo, h), 0.0), 1.0), x5 = x * x * x * x * x;
    double k = Dh * G2 / (4 * co * ci);
    return D((g[0] + (1 - g[0]) * x5) * k, (g[1] + (1 - g[1]) * x5) * k, (g[2] + (1 - g[2]) * x5) * k);
}
static int ggx64_sample(const float *g, d3 n, d3 wo_w, double u1, double u2, d3 *dir, d3 *weight) {
    d3 t, b;
    basis64(n, &t, &b);
    d3 wo = D(ddot(wo_w, t), ddot(wo_w, b), ddot(wo_w, n));
    if (wo.z <= 0) return 0;
    double a = g[3], a2 = a * a;
    d3 vh = dnorm(D(a * wo.x, a * wo.y, wo.z));
    double l2 = vh.x * vh.x + vh.y * vh.y;
    d3 T1 = l2 > 0 ? dsc(D(-vh.y, vh.x, 0), 1 / sqrt(l2)) : D(1, 0, 0);
    d3 T2 = dcross(vh, T1);
    double r = sqrt(u1), ph = 2 * M_PI * u2, t1 = r * cos(ph), t2 = r * sin(ph), sh = 0.5 * (1 + vh.z);
    t2 = (1 - sh) * sqrt(fmax(1 - t1 * t1, 0.0)) + sh * t2;
    d3 nh = dadd(dadd(dsc(T1, t1), dsc(T2, t2)), dsc(vh, sqrt(fmax(0.0, 1 - t1 * t1 - t2 * t2))));
    d3 h = dnorm(D(a * nh.x, a * nh.y, fmax(nh.z, 0.0)));
    double oh = ddot(wo, h);
    d3 wi = dsub(dsc(h, 2 * oh), wo);
    if (wi.z <= 0 || oh <= 0) return 0;
    double lo = ggx64_lambda(a2, wo.z), li = ggx64_lambda(a2, wi.z);
    double x = 1 - fmin(oh, 1.0), x5 = x * x * x * x * x, k = (1 + lo) / (1 + lo + li);
    *weight = D((g[0] + (1 - g[0]) * x5) * k, (g[1] + (1 - g[1]) * x5) * k, (g[2] + (1 - g[2]) * x5) * k);
    *dir = dnorm(dadd(dadd(dsc(t, wi.x), dsc(b, wi.y)), dsc(n, wi.z)));
    return 1;
}

static void render_path64(const scene64 *s, uint64_t path_id, pcg_t *rng, double *out) {
    const orc_render_settings *st = &s->st;
    const float *k = s->cam;
    uint64_t pixel = path_id / st->spp, px = pixel % st->width, py = pixel / st->width;
    double pu = pcg_d(rng), pv = pcg_d(rng);
    double fw = st->width, fh = st->height, aspect = fw / fh;
    double sx = ((px + pu) / fw) * 2.0 - 1.0, sy = 1.0 - ((py + pv) / fh) * 2.0;
    d3 origin = D(k[0], k[1], k[2]);
    d3 plane = dadd(dadd(D(k[3], k[4], k[5]), dsc(D(k[6], k[7], k[8]), sx * aspect * k[12])), dsc(D(k[9], k[10], k[11]), sy * k[12]));
    d3 o = origin, d = dnorm(dsub(dadd(origin, dsc(plane, k[13])), origin));
    d3 T = D(1, 1, 1);
    int alg = st->algorithm;
    uint32_t depth = alg == 1 || alg == 2 ? 1u : st->max_depth;
    out[0] = out[1] = out[2] = 0;
    double tmin = 0.001;
    for (uint32_t b = 0; b < depth; ++b) {
        double t;
        uint64_t prim;
        d3 nrm;
        if (!hit64(s, o, d, tmin, 1e300, &t, &prim, &nrm)) {
            if (alg == 0 || alg == 3) {
                d3 u = dnorm(d);
                double a = 0.5 * (u.y + 1.0);
                d3 sk = dadd(dsc(D(1, 1, 1), 1.0 - a), dsc(D(0.5, 0.7, 1.0), a));
                d3 c = dmul(T, sk);
                out[0] += c.x; out[1] += c.y; out[2] += c.z;
            }
            return;
        }
        int front = ddot(d, nrm) < 0;
        d3 n = front ? nrm : dsc(nrm, -1);
        d3 p = dadd(o, dsc(d, t));
        uint32_t surf = (prim & SPH_BIT) ? s->ssurf[prim & ~SPH_BIT] : s->surf[prim], kind = surf >> 28, mi = surf & IDX_MASK;
        if (alg == 1) { out[0] += 0.5 * (n.x + 1); out[1] += 0.5 * (n.y + 1); out[2] += 0.5 * (n.z + 1); return; }
        if (alg == 2) {
            double u1 = pcg_d(rng), u2 = pcg_d(rng);
            double z = 1 - 2 * u2, r = sqrt(fmax(0.0, 1 - z * z)), th = 2 * M_PI * u1;
            d3 dir = D(r * cos(th), r * sin(th), z);
            if (ddot(dir, n) < 0) dir = dsc(dir, -1);
            double tt; uint64_t pp; d3 nn;
            double v = hit64(s, p, dir, 0.001, 4.0, &tt, &pp, &nn) ? 0.08 : 1.0;
            out[0] += v; out[1] += v; out[2] += v;
            return;
        }
        if (kind == MAT_EMIT) {
            if (front && (alg == 0 || b == 0)) {
                const float *e = s->emit + 3 * mi;
                out[0] += T.x * e[0]; out[1] += T.y * e[1]; out[2] += T.z * e[2];
            }
            return;
        }
        if (kind == MAT_METAL || kind == MAT_DIEL) {
            double u1 = pcg_d(rng), u2 = pcg_d(rng);
            d3 nd, wgt;
            if (!scatter64(s, kind, mi, d, n, front, u1, u2, &nd, &wgt)) return;
            T = dmul(T, wgt);
            if (!roulette64(rng, b, &T)) return;
            o = p; d = nd;
            continue;
        }
        const int ggx = kind == MAT_GGX;
        const float *gm = ggx ? s->st.ggx + 4 * (size_t)mi : NULL;
        d3 alb = ggx ? D(0, 0, 0) : D(s->lamb[3 * mi], s->lamb[3 * mi + 1], s->lamb[3 * mi + 2]);
        d3 wo = dnorm(dsc(d, -1));
        if (alg == 3 && s->nl > 0) {
            /* light chosen by power, point uniform on the triangle or the sphere */
            double xi = pcg_d(rng) * s->total, accw = 0;
            uint32_t l = 0;
            while (l + 1 < s->nl && accw + s->lw[l] < xi) accw += s->lw[l++];
            d3 ln, lp;
            double area;
            uint32_t lsurf;
            if (s->lsph[l]) {
                const float *sp = s->sph + 4 * (uint64_t)s->ltri[l];
                double r = fabs(sp[3]), z = 1 - 2 * pcg_d(rng), th = 2 * M_PI * pcg_d(rng);
                double rr = sqrt(fmax(0.0, 1 - z * z));
                ln = D(rr * cos(th), rr * sin(th), z);
                lp = dadd(D(sp[0], sp[1], sp[2]), dsc(ln, r));
                area = 4 * M_PI * r * r;
                lsurf = s->ssurf[s->ltri[l]];
            } else {
                const float *lt = s->v + 9 * (uint64_t)s->ltri[l];
                d3 p0 = D(lt[0], lt[1], lt[2]), e1 = dsub(D(lt[3], lt[4], lt[5]), p0), e2 = dsub(D(lt[6], lt[7], lt[8]), p0);
                d3 av = dcross(e1, e2);
                double area2 = sqrt(ddot(av, av));
                ln = dsc(av, 1.0 / area2);
                double su = sqrt(pcg_d(rng)), b1 = su * (1 - pcg_d(rng)), b2 = su - b1;
                lp = dadd(dadd(p0, dsc(e1, b1)), dsc(e2, b2));
                area = 0.5 * area2;
                lsurf = s->surf[s->ltri[l]];
            }
            d3 tl = dsub(lp, p);
            double d2 = ddot(tl, tl), dist = sqrt(d2);
            d3 dir = dsc(tl, 1.0 / dist);
            double sc = ddot(n, dir), lc = -ddot(ln, dir);
            if (sc > 0 && lc > 0 && dist - 0.002 > 0.001) {
                double tt; uint64_t pp; d3 nn;
                if (!hit64(s, p, dir, 0.001, dist - 0.002, &tt, &pp, &nn)) {
                    const float *e = s->emit + 3 * (lsurf & IDX_MASK);
                    double pdf_l = (s->lw[l] / s->total) * d2 / (lc * area);
                    d3 fv = ggx ? ggx64_f(gm, n, wo, dir) : dsc(alb, 1.0 / M_PI);
                    double f = sc / pdf_l;
                    out[0] += T.x * fv.x * e[0] * f; out[1] += T.y * fv.y * e[1] * f; out[2] += T.z * fv.z * e[2] * f;
                }
            }
        }
        if (ggx) {
            double u1 = pcg_d(rng), u2 = pcg_d(rng);
            d3 nd, wgt;
            if (!ggx64_sample(gm, n, wo, u1, u2, &nd, &wgt)) return;
            T = dmul(T, wgt);
            if (!roulette64(rng, b, &T)) return;
            o = p; d = nd;
            continue;
        }
        /* cosine-weighted bounce, throughput *= albedo */
        double u1 = pcg_d(rng), u2 = pcg_d(rng);
        double z = 1 - 2 * u2, r = sqrt(fmax(0.0, 1 - z * z)), th = 2 * M_PI * u1;
        d3 sd = dadd(n, D(r * cos(th), r * sin(th), z));
        if (ddot(sd, sd) < 1e-20) sd = n;
        T = dmul(T, alb);
        if (!roulette64(rng, b, &T)) return;
        o = p; d = dnorm(sd);
    }
}

typedef struct { const scene64 *s; uint64_t first, last; double *px; uint64_t seed; } job64;

static void *worker64(void *p) {
    job64 *j = (job64 *)p;
    const orc_render_settings *st = &j->s->st;
    for (uint64_t pix = j->first; pix < j->last; ++pix) {
        double sum[3] = {0, 0, 0};
        for (uint32_t k = 0; k < st->spp; ++k) {
            uint64_t path = pix * st->spp + k;
            pcg_t rng = {0, ((path ^ j->seed) << 1u) | 1u};
            pcg_next(&rng); rng.state += 0x853c49e6748fea9bULL + j->seed; pcg_next(&rng);
            double o[3];
            render_path64(j->s, path, &rng, o);
            sum[0] += o[0]; sum[1] += o[1]; sum[2] += o[2];
        }
        for (int c = 0; c < 3; ++c) j->px[3 * pix + c] = sum[c] / st->spp;
    }
    return NULL;
}

/* AO rays of a closest-hit launch (the renderer's AO ray, bajo's
 * _make_ao_ray, rt/gpu/path_shading.mojo:650-665, per sample): ray
 * j = r * samples + k uses stream (seed, path j, stage 1); a primary ray
 * without a hit yields origin 0, direction +z, t_max -1.  Field-major in and
 * out (8 planes), hits AoS (8 floats). */
void orc_ao_rays(const float *rays, const float *hits, uint64_t n, uint32_t samples, uint64_t seed, float *out) {
    const uint64_t m = n * samples;
    for (uint64_t j = 0; j < m; ++j) {
        uint64_t r = j / samples;
        const float *h = hits + 8 * r;
        uint32_t prim;
        memcpy(&prim, h + 2, 4);
        orc_v3 o = orc_v(rays[r], rays[n + r], rays[2 * n + r]);
        orc_v3 d = orc_v(rays[4 * n + r], rays[5 * n + r], rays[6 * n + r]);
        float tmax = rays[7 * n + r];
        orc_v3 po = orc_v(0, 0, 0), dir = orc_v(0, 0, 1);
        float thi = -1.0f;
        if (prim != ORC_EMPTY && h[7] < tmax) {
            orc_v3 hn = orc_v(h[4], h[5], h[6]);
            orc_v3 nn = orc_dot(d, hn) < 0.0f ? hn : orc_v(-hn.x, -hn.y, -hn.z);
            po = orc_add(o, orc_scale(d, h[7]));
            orc_rng ar;
            rng_init(&ar, seed, (uint32_t)j, 1u);
            float as_, ac_;
            orc_sincos_turns(rng_next(&ar), &as_, &ac_);
            float z = 1.0f - 2.0f * rng_next(&ar);
            float rad = sqrtf(1.0f - z * z);
            dir = orc_v(rad * ac_, rad * as_, z);
            if (orc_dot(dir, nn) < 0.0f) dir = orc_v(-dir.x, -dir.y, -dir.z);
            thi = 4.0f;
        }
        float f[8] = {po.x, po.y, po.z, 0.001f, dir.x, dir.y, dir.z, thi};
        for (int k = 0; k < 8; ++k) out[(uint64_t)k * m + j] = f[k];
    }
}

/* FP64 estimate of the PATH (0), NORMALS (1), AO (2) or NEE (3) image; scenes with metal or dielectric
 * surfaces are estimated with PATH (its NEE handles Lambertian surfaces only). */
int orc_render_fp64(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *lambertian,
                    const float *metal, const float *dielectric, const float *emissive,
                    const orc_render_settings *st, const float *cam20, double *pixels, int threads) {
    return orc_render_fp64_mixed(vertices, n, surfaces, NULL, 0, NULL, lambertian, metal, dielectric, emissive, st,
                                 cam20, pixels, threads);
}

int orc_render_fp64_mixed(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *spheres,
                          uint64_t ns, const uint32_t *sphere_surfaces, const float *lambertian, const float *metal,
                          const float *dielectric, const float *emissive, const orc_render_settings *st,
                          const float *cam20, double *pixels, int threads) {
    scene64 s = {vertices, n, surfaces, lambertian, metal, dielectric, emissive, NULL, NULL, 0, 0.0, *st, cam20,
                 spheres, ns, sphere_surfaces, NULL};
    s.lw = (double *)malloc((n + ns ? n + ns : 1) * sizeof(double));
    s.ltri = (uint32_t *)malloc((n + ns ? n + ns : 1) * sizeof(uint32_t));
    s.lsph = (uint8_t *)calloc(n + ns ? n + ns : 1, 1);
    for (uint64_t i = 0; i < n; ++i) {
        if ((surfaces[i] >> 28) != MAT_EMIT) continue;
        const float *t = vertices + 9 * i;
        d3 p0 = D(t[0], t[1], t[2]);
        d3 cr = dcross(dsub(D(t[3], t[4], t[5]), p0), dsub(D(t[6], t[7], t[8]), p0));
        const float *e = emissive + 3 * (surfaces[i] & IDX_MASK);
        double w = 0.5 * sqrt(ddot(cr, cr)) * fmax((e[0] + (double)e[1] + e[2]) / 3.0, 0.0);
        if (w > 0) { s.lw[s.nl] = w; s.ltri[s.nl] = (uint32_t)i; s.total += w; s.nl++; }
    }
    for (uint64_t i = 0; i < ns; ++i) {
        if ((sphere_surfaces[i] >> 28) != MAT_EMIT) continue;
        double r = fabs(spheres[4 * i + 3]);
        const float *e = emissive + 3 * (sphere_surfaces[i] & IDX_MASK);
        double w = 4 * M_PI * r * r * fmax((e[0] + (double)e[1] + e[2]) / 3.0, 0.0);
        if (w > 0) { s.lw[s.nl] = w; s.ltri[s.nl] = (uint32_t)i; s.lsph[s.nl] = 1; s.total += w; s.nl++; }
    }
    uint64_t pixels_n = (uint64_t)st->width * st->height;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    job64 jobs[256];
    for (int t = 0; t < threads; ++t) {
        job64 j = {&s, pixels_n * (uint64_t)t / (uint64_t)threads, pixels_n * (uint64_t)(t + 1) / (uint64_t)threads, pixels, st->seed};
        jobs[t] = j;
        pthread_create(&tid[t], NULL, worker64, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    free(s.lw); free(s.ltri); free(s.lsph);
    return 0;
}
