/*
 * CPU oracle -- test infrastructure only.
 *
 * Plain-C restatement of bajo's scalar math that the BVH hot path depends on.
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load this code; the product library never links it.
 *
 * Every helper follows the reference operation-for-operation (same FMA
 * placement, same select-based min/max, same epsilons) so that results are
 * bit-identical to a faithful FP32 implementation.  Build with
 * -ffp-contract=off so the compiler adds no fused operations of its own.
 */
#ifndef ORC_CORE_H
#define ORC_CORE_H

#include <float.h>
#include <math.h>
#include <stdint.h>

typedef struct { float x, y, z; } orc_v3;
typedef struct { orc_v3 lo, hi; } orc_aabb;

#define ORC_FLT_MAX 3.40282347e+38f
#define ORC_EMPTY 0xFFFFFFFFu
#define ORC_SENTINEL 0xFFFFFFFFu
#define ORC_LEAF_FLAG 0x80000000u
#define ORC_INDEX_MASK 0x7FFFFFFFu

/* core/utils.mojo:60-72 -- select-based max/min (not IEEE fmin/fmax). */
static inline float orc_fmax(float a, float b) { return a > b ? a : b; }
static inline float orc_fmin(float a, float b) { return a < b ? a : b; }

static inline orc_v3 orc_v(float x, float y, float z) { orc_v3 r = {x, y, z}; return r; }
static inline orc_v3 orc_add(orc_v3 a, orc_v3 b) { return orc_v(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline orc_v3 orc_sub(orc_v3 a, orc_v3 b) { return orc_v(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline orc_v3 orc_mul(orc_v3 a, orc_v3 b) { return orc_v(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline orc_v3 orc_scale(orc_v3 a, float s) { return orc_v(a.x * s, a.y * s, a.z * s); }
/* core/vec.mojo vmin/vmax: component-wise select min/max. */
static inline orc_v3 orc_vmin(orc_v3 a, orc_v3 b) { return orc_v(orc_fmin(a.x, b.x), orc_fmin(a.y, b.y), orc_fmin(a.z, b.z)); }
static inline orc_v3 orc_vmax(orc_v3 a, orc_v3 b) { return orc_v(orc_fmax(a.x, b.x), orc_fmax(a.y, b.y), orc_fmax(a.z, b.z)); }

/* core/vec.mojo:435-440 -- dot = fma(ax,bx, fma(ay,by, az*bz)). */
static inline float orc_dot(orc_v3 a, orc_v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

/* core/vec.mojo:517-526 -- cross with one FMA per component. */
static inline orc_v3 orc_cross(orc_v3 a, orc_v3 b) {
    return orc_v(fmaf(a.y, b.z, -(a.z * b.y)),
                 fmaf(a.z, b.x, -(a.x * b.z)),
                 fmaf(a.x, b.y, -(a.y * b.x)));
}

/* core/vec.mojo:529-540 -- normalize with a 1e-20 length threshold. */
static inline orc_v3 orc_normalize(orc_v3 v) {
    float l = sqrtf(orc_dot(v, v));
    float inv = l > 1.0e-20f ? 1.0f / l : 0.0f;
    return orc_scale(v, inv);
}

/* core/vec.mojo:373-395 -- safe_inv with eps 1e-20 (zero when tiny). */
static inline orc_v3 orc_safe_inv(orc_v3 v) {
    const float e = 1.0e-20f;
    return orc_v(fabsf(v.x) > e ? 1.0f / v.x : 0.0f,
                 fabsf(v.y) > e ? 1.0f / v.y : 0.0f,
                 fabsf(v.z) > e ? 1.0f / v.z : 0.0f);
}

/* core/ray.mojo:86-112 -- reciprocal direction, |d| <= 1e-9 -> +-1e9. */
static inline orc_v3 orc_rcp_dir(orc_v3 d) {
    const float e = 1.0e-9f, large = 1.0f / 1.0e-9f;
    return orc_v(fabsf(d.x) > e ? 1.0f / d.x : (d.x < 0.0f ? -large : large),
                 fabsf(d.y) > e ? 1.0f / d.y : (d.y < 0.0f ? -large : large),
                 fabsf(d.z) > e ? 1.0f / d.z : (d.z < 0.0f ? -large : large));
}

/* core/aabb.mojo:28-35 -- the empty box is (+FLT_MAX, -FLT_MAX). */
static inline orc_aabb orc_aabb_invalid(void) {
    orc_aabb b = {{ORC_FLT_MAX, ORC_FLT_MAX, ORC_FLT_MAX}, {-ORC_FLT_MAX, -ORC_FLT_MAX, -ORC_FLT_MAX}};
    return b;
}
static inline orc_aabb orc_aabb_merge(orc_aabb a, orc_aabb b) {
    orc_aabb r = {orc_vmin(a.lo, b.lo), orc_vmax(a.hi, b.hi)};
    return r;
}
static inline void orc_aabb_grow_pt(orc_aabb *a, orc_v3 p) {
    a->lo = orc_vmin(a->lo, p);
    a->hi = orc_vmax(a->hi, p);
}
/* core/aabb.mojo:53-54 -- centroid = (min + max) * 0.5. */
static inline orc_v3 orc_aabb_centroid(orc_aabb a) { return orc_scale(orc_add(a.lo, a.hi), 0.5f); }
static inline orc_aabb orc_aabb_load6(const float *p) {
    orc_aabb b = {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}};
    return b;
}
static inline void orc_aabb_store6(float *p, orc_aabb b) {
    p[0] = b.lo.x; p[1] = b.lo.y; p[2] = b.lo.z;
    p[3] = b.hi.x; p[4] = b.hi.y; p[5] = b.hi.z;
}

/* core/morton.mojo:48-58 -- spread 10 bits to every third bit. */
static inline uint32_t orc_expand_bits_3d(uint32_t n) {
    n = (n ^ (n << 16)) & 0xFF0000FFu;
    n = (n ^ (n << 8)) & 0x0300F00Fu;
    n = (n ^ (n << 4)) & 0x030C30C3u;
    n = (n ^ (n << 2)) & 0x09249249u;
    return n;
}
/* core/morton.mojo:61-89 -- truncating cast of v*1024 clamped to [0,1023]. */
static inline uint32_t orc_quantize_axis(float v) {
    float s = v * 1024.0f;
    uint32_t u;
    if (!(s > 0.0f)) u = 0;              /* NaN and negatives saturate to 0 */
    else if (s >= 4294967296.0f) u = 0xFFFFFFFFu;
    else u = (uint32_t)s;
    return u > 1023u ? 1023u : u;
}
static inline uint32_t orc_morton3(float x, float y, float z) {
    return (orc_expand_bits_3d(orc_quantize_axis(z)) << 2) |
           (orc_expand_bits_3d(orc_quantize_axis(y)) << 1) |
           orc_expand_bits_3d(orc_quantize_axis(x));
}
/* core/morton.mojo:6-31 -- 64-bit (code, sorted index) key and its common prefix. */
static inline uint64_t orc_morton_key(uint32_t code, uint32_t idx) { return ((uint64_t)code << 32) | idx; }
static inline int orc_morton_common_prefix(uint32_t ca, uint32_t ia, uint32_t cb, uint32_t ib) {
    uint64_t d = orc_morton_key(ca, ia) ^ orc_morton_key(cb, ib);
    return d == 0 ? 64 : __builtin_clzll(d);
}

/* core/intersect.mojo:193-215 -- slab test with precomputed reciprocal;
 * the near distance is clamped at 0 and the far distance at t_max. */
static inline int orc_ray_aabb(orc_v3 o, orc_v3 rcp, orc_aabb b, float t_max, float *t_out) {
    orc_v3 t0 = orc_mul(orc_sub(b.lo, o), rcp);
    orc_v3 t1 = orc_mul(orc_sub(b.hi, o), rcp);
    orc_v3 tn = orc_vmin(t0, t1), tf = orc_vmax(t0, t1);
    float tmin = orc_fmax(orc_fmax(tn.x, tn.y), orc_fmax(tn.z, 0.0f));
    float tmax = orc_fmin(orc_fmin(tf.x, tf.y), orc_fmin(tf.z, t_max));
    *t_out = tmin;
    return tmin <= tmax;
}

/* core/intersect.mojo:415-459 -- eager Moller-Trumbore (EPSILON 1e-8). */
static inline int orc_ray_tri_eager(orc_v3 o, orc_v3 d, orc_v3 v0, orc_v3 e1, orc_v3 e2,
                                    float t_max, float t_min, float *t, float *u, float *v) {
    orc_v3 p = orc_cross(d, e2);
    float det = orc_dot(e1, p);
    int det_ok = det > 1e-8f || det < -1e-8f;
    float inv_det = 1.0f / det;
    orc_v3 tv = orc_sub(o, v0);
    float uu = orc_dot(tv, p) * inv_det;
    orc_v3 q = orc_cross(tv, e1);
    float vv = orc_dot(d, q) * inv_det;
    float tt = orc_dot(e2, q) * inv_det;
    int mask = det_ok && uu >= 0.0f && uu <= 1.0f && vv >= 0.0f && (uu + vv) <= 1.0f && tt > t_min && tt < t_max;
    *t = mask ? tt : ORC_FLT_MAX;
    *u = uu;
    *v = vv;
    return mask;
}

/* core/intersect.mojo:462-508 -- division-free (determinant-scaled) MT. */
typedef struct { int mask; float abs_det, t_scaled, u_scaled, v_scaled; } orc_scaled_hit;
static inline orc_scaled_hit orc_ray_tri_scaled(orc_v3 o, orc_v3 d, orc_v3 v0, orc_v3 e1, orc_v3 e2,
                                               float t_max, float t_min) {
    orc_scaled_hit h;
    orc_v3 p = orc_cross(d, e2);
    float det = orc_dot(e1, p);
    int pos = det >= 0.0f;
    float abs_det = pos ? det : -det;
    orc_v3 tv = orc_sub(o, v0);
    float u_num = orc_dot(tv, p);
    float us = pos ? u_num : -u_num;
    orc_v3 q = orc_cross(tv, e1);
    float v_num = orc_dot(d, q);
    float t_num = orc_dot(e2, q);
    float vs = pos ? v_num : -v_num;
    float ts = pos ? t_num : -t_num;
    h.mask = abs_det > 1e-8f && us >= 0.0f && us <= abs_det && vs >= 0.0f && (us + vs) <= abs_det &&
             ts > t_min * abs_det && ts < t_max * abs_det;
    h.abs_det = abs_det;
    h.t_scaled = ts;
    h.u_scaled = us;
    h.v_scaled = vs;
    return h;
}

#endif
