/*
 * CPU oracle -- test infrastructure only (see orc_core.h).
 *
 * Instances under a top-level hierarchy, restating bajo's TLAS
 * (bvh/gpu/tlas.mojo:54-233, bvh/tlas_common.mojo:8-45) with bajo's Affine3
 * (core/transform.mojo: point, vector, inverse, normal_from_inverse) and
 * AABB.apply_transform (core/aabb.mojo:101-137):
 *   - instance world box = the BLAS root box through the transform;
 *   - the top level is a binary layout over those boxes (built by
 *     orc_build_lbvh_bounds or the H-PLOC restatement), walked near-first;
 *   - at an instance leaf the ray is mapped by the inverse transform (t is
 *     unchanged) and the BLAS traversed with t_max = the closest hit so far;
 *     a closer BLAS hit is promoted with inst = the instance; the winning
 *     normal goes through the inverse transpose and is normalized;
 *   - any hit: bajo's shadow hit carrying the instance.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "orc_internal.h"

/* Affine3.inverse: adjugate / det, translation -inv(M) t; 0 when |det| <= 1e-8. */
int orc_affine_inverse(const float *a, float *o) {
    float m00 = a[0], m01 = a[1], m02 = a[2], tx = a[3];
    float m10 = a[4], m11 = a[5], m12 = a[6], ty = a[7];
    float m20 = a[8], m21 = a[9], m22 = a[10], tz = a[11];
    float i00 = m11 * m22 - m12 * m21, i01 = m02 * m21 - m01 * m22, i02 = m01 * m12 - m02 * m11;
    float i10 = m12 * m20 - m10 * m22, i11 = m00 * m22 - m02 * m20, i12 = m02 * m10 - m00 * m12;
    float i20 = m10 * m21 - m11 * m20, i21 = m01 * m20 - m00 * m21, i22 = m00 * m11 - m01 * m10;
    float det = m00 * i00 + m01 * i10 + m02 * i20;
    int ok = fabsf(det) > 1.0e-8f;
    float r = 1.0f / (ok ? det : 1.0f);
    i00 *= r; i01 *= r; i02 *= r;
    i10 *= r; i11 *= r; i12 *= r;
    i20 *= r; i21 *= r; i22 *= r;
    float v[12] = {i00, i01, i02, -(i00 * tx + i01 * ty + i02 * tz),
                   i10, i11, i12, -(i10 * tx + i11 * ty + i12 * tz),
                   i20, i21, i22, -(i20 * tx + i21 * ty + i22 * tz)};
    memcpy(o, v, sizeof(v));
    return ok;
}

/* AABB.apply_transform: translation plus, per column, min / max of the column
 * scaled by the box's low and high coordinate. */
void orc_transform_box(const float *m, const float *b, float *out) {
    float lo[3] = {m[3], m[7], m[11]}, hi[3] = {m[3], m[7], m[11]};
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            float col = m[4 * r + c];
            float a = col * b[c], z = col * b[3 + c];
            lo[r] = lo[r] + orc_fmin(a, z);
            hi[r] = hi[r] + orc_fmax(a, z);
        }
    for (int k = 0; k < 3; ++k) { out[k] = lo[k]; out[3 + k] = hi[k]; }
}

static orc_v3 xf_point(const float *m, orc_v3 p) {
    return orc_v(fmaf(m[0], p.x, fmaf(m[1], p.y, fmaf(m[2], p.z, m[3]))),
                 fmaf(m[4], p.x, fmaf(m[5], p.y, fmaf(m[6], p.z, m[7]))),
                 fmaf(m[8], p.x, fmaf(m[9], p.y, fmaf(m[10], p.z, m[11]))));
}
static orc_v3 xf_vector(const float *m, orc_v3 v) {
    return orc_v(fmaf(m[0], v.x, fmaf(m[1], v.y, m[2] * v.z)), fmaf(m[4], v.x, fmaf(m[5], v.y, m[6] * v.z)),
                 fmaf(m[8], v.x, fmaf(m[9], v.y, m[10] * v.z)));
}
static orc_v3 normal_from_inverse(const float *m, orc_v3 n) {
    return orc_normalize(orc_v(fmaf(m[0], n.x, fmaf(m[4], n.y, m[8] * n.z)),
                               fmaf(m[1], n.x, fmaf(m[5], n.y, m[9] * n.z)),
                               fmaf(m[2], n.x, fmaf(m[6], n.y, m[10] * n.z))));
}

typedef struct {
    const orc_bvh *blas;
    const float *inv;
    const uint32_t *inst_blas;
    uint32_t n_inst;
    const uint32_t *top_ids, *top_meta;
    const float *top_bounds;
    uint32_t top_root;
} tlas_t;

/* visit one instance leaf; returns 1 once an any-hit query is answered */
static int visit(const tlas_t *t, uint32_t pos, int any, orc_v3 o, orc_v3 d, float tmin, float tmax, orc_hit *hit,
                 uint32_t *occ_inst) {
    uint32_t inst = t->top_ids[pos];
    if (inst >= t->n_inst) return 0;
    const orc_bvh *b = t->blas + t->inst_blas[inst];
    if (b->prim_count == 0) return 0;
    const float *m = t->inv + 12 * (size_t)inst;
    orc_v3 lo = xf_point(m, o), ld = xf_vector(m, d);
    if (any) {
        orc_hit lh = orc_trace_any(b, lo, ld, tmin, tmax);
        if (lh.prim == ORC_EMPTY && lh.t == 0.0f) { *occ_inst = inst; return 1; }
        return 0;
    }
    orc_hit lh = orc_trace_closest(b, lo, ld, tmin, hit->t, NULL);
    if (lh.prim != ORC_EMPTY && lh.t < hit->t) { *hit = lh; hit->inst = inst; }
    return 0;
}

static orc_hit tlas_query(const tlas_t *t, int any, orc_v3 o, orc_v3 d, float tmin, float tmax) {
    orc_hit hit = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, tmax};
    uint32_t occ = ORC_EMPTY;
    int done = 0;
    if (t->top_root & ORC_LEAF_FLAG) {
        done = visit(t, t->top_root & ORC_INDEX_MASK, any, o, d, tmin, tmax, &hit, &occ);
    } else if (t->top_root != ORC_SENTINEL) {
        orc_v3 rcp = orc_rcp_dir(d);
        uint32_t stack[128];
        int sp = 0;
        uint32_t cur = t->top_root;
        for (;;) {
            const uint32_t *mm = t->top_meta + 4 * (size_t)cur;
            const float *nb = t->top_bounds + 12 * (size_t)cur;
            float node_tmax = any ? tmax : hit.t;
            float ct[2];
            int hm[2];
            hm[0] = orc_ray_aabb(o, rcp, orc_aabb_load6(nb), node_tmax, &ct[0]);
            hm[1] = orc_ray_aabb(o, rcp, orc_aabb_load6(nb + 6), node_tmax, &ct[1]);
            int valid[2] = {0, 0};
            for (int lane = 0; lane < 2 && !done; ++lane) {
                uint32_t c = mm[1 + lane];
                if (!hm[lane]) continue;
                if (c & ORC_LEAF_FLAG) done = visit(t, c & ORC_INDEX_MASK, any, o, d, tmin, tmax, &hit, &occ);
                else valid[lane] = 1;
            }
            if (done) break;
            int near = -1;
            float near_t = ORC_FLT_MAX;
            if (valid[0]) { near = 0; near_t = ct[0]; }
            if (valid[1] && (near == -1 || ct[1] < near_t)) { near = 1; near_t = ct[1]; }
            float bound = any ? tmax : hit.t;
            for (int lane = 0; lane < 2; ++lane)
                if (valid[lane] && near != lane && ct[lane] <= bound) {
                    if (sp >= 128) return hit;
                    stack[sp++] = mm[1 + lane];
                }
            if (near != -1 && near_t <= bound) { cur = mm[1 + near]; continue; }
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    if (any) {
        orc_hit r = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, 0.0f};
        if (done) { r.inst = occ; return r; }
        orc_hit miss = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, tmax};
        return miss;
    }
    if (hit.prim != ORC_EMPTY) hit.n = normal_from_inverse(t->inv + 12 * (size_t)hit.inst, hit.n);
    return hit;
}

/* R field-major rays through the instances; mode 0 closest, 1 any hit.
 * blas: n_blas BVH descriptions; inv f32[12 I] (world -> BLAS); top: binary
 * layout over the instance boxes (leaf ids = instance indices). */
void orc_tlas_trace(const orc_bvh *blas, uint32_t n_blas, const float *inv, const uint32_t *inst_blas,
                    uint32_t n_inst, const uint32_t *top_ids, const uint32_t *top_meta, const float *top_bounds,
                    uint32_t top_root, const float *rays, uint64_t R, int mode, float *hits) {
    (void)n_blas;
    tlas_t t = {blas, inv, inst_blas, n_inst, top_ids, top_meta, top_bounds, top_root};
    for (uint64_t r = 0; r < R; ++r) {
        orc_v3 o = orc_v(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
        orc_v3 d = orc_v(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
        orc_hit h = tlas_query(&t, mode, o, d, rays[3 * R + r], rays[7 * R + r]);
        float *p = hits + 8 * r;
        p[0] = h.u; p[1] = h.v;
        memcpy(p + 2, &h.prim, 4);
        memcpy(p + 3, &h.inst, 4);
        p[4] = h.n.x; p[5] = h.n.y; p[6] = h.n.z; p[7] = h.t;
    }
}
