/* CPU oracle -- test infrastructure only: types shared between oracle files. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include "orc_core.h"

typedef struct { float u, v; uint32_t prim, inst; orc_v3 n; float t; } orc_hit;

typedef struct {
    const float *vertices; /* f32[9n], original order */
    const uint32_t *leaf_ids;
    const uint32_t *node_meta;
    const float *node_bounds;
    uint32_t root;
    /* 8-wide layout (orc_wide.c): when non-NULL, queries traverse it */
    const void *wide;
    uint64_t wide_count;
    uint64_t prim_count;
    /* sphere BVH: when non-NULL, leaves are spheres f32[4n] (center, radius) */
    const float *spheres;
} orc_bvh;

/* bajo's intersect_ray_sphere (core/intersect.mojo:309-350): the nearer root
 * inside (t_min, t_max), else the farther one. */
static inline int orc_sphere_t(const float *sp, orc_v3 o, orc_v3 d, float t_min, float t_max, float *t) {
    orc_v3 oc = orc_sub(o, orc_v(sp[0], sp[1], sp[2]));
    float a = orc_dot(d, d), half_b = orc_dot(oc, d), c = orc_dot(oc, oc) - sp[3] * sp[3];
    float det = half_b * half_b - a * c;
    int ok = det >= 0.0f;
    float sq = sqrtf(orc_fmax(det, 0.0f));
    float inv_a = 1.0f / a;
    float t0 = (-half_b - sq) * inv_a, t1 = (-half_b + sq) * inv_a;
    int near = ok && t0 > t_min && t0 < t_max;
    int far = ok && !near && t1 > t_min && t1 < t_max;
    *t = near ? t0 : t1;
    return near || far;
}

/* The sphere hit record of bvh/gpu/sphere_bvh.mojo:464-476: u = v = 0,
 * normal = normalize(o + t d - center). */
static inline void orc_sphere_record(const float *sp, uint32_t prim, orc_v3 o, orc_v3 d, float t, orc_hit *hit) {
    orc_v3 p = orc_add(o, orc_scale(d, t));
    hit->t = t;
    hit->u = 0.0f;
    hit->v = 0.0f;
    hit->prim = prim;
    hit->inst = ORC_EMPTY;
    hit->n = orc_normalize(orc_sub(p, orc_v(sp[0], sp[1], sp[2])));
}

int orc_trace_wide_query(const orc_bvh *b, int any, orc_v3 o, orc_v3 d, float t_min, float t_max, orc_hit *out,
                         uint32_t *stats);

orc_hit orc_trace_closest(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max, uint32_t *stats);
orc_hit orc_trace_any(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max);

#endif
