/*
 * CPU oracle -- test infrastructure only (see orc_core.h).
 *
 * Restates bajo's ray queries on the CPU:
 *   - Camera basis and raster rays   bvh/camera.mojo:86-154
 *   - packed ray / hit ABI           bvh/gpu/ray_launch.mojo:28-54, bvh/types.mojo:16-115
 *   - near-first BVH2 closest hit    bvh/gpu/trace.mojo:341-428
 *   - unordered any hit              bvh/gpu/trace.mojo:465-491
 *   - leaf test: scaled MT, t = t_scaled * (1/|det|), normal = normalize(e1 x e2)
 *                                    bvh/gpu/triangle_bvh.mojo:906-931, bvh/cpu/triangle_bvh.mojo:92-146
 *   - brute force                    test/bvh/fixtures.mojo:179-209
 */
#include <stdlib.h>
#include <string.h>

#include "orc_core.h"
#include "oracle.h"
#include "orc_internal.h"

/* bvh/camera.mojo:86-106 (defocus angle 0). */
void orc_camera_make(const float *origin3, const float *target3, const float *up3, float fov_scale,
                     float focus_dist, float *out20) {
    orc_v3 o = orc_v(origin3[0], origin3[1], origin3[2]);
    orc_v3 t = orc_v(target3[0], target3[1], target3[2]);
    orc_v3 up = orc_v(up3[0], up3[1], up3[2]);
    orc_v3 fwd = orc_normalize(orc_sub(t, o));
    orc_v3 right = orc_normalize(orc_cross(fwd, up));
    orc_v3 cup = orc_normalize(orc_cross(right, fwd));
    float v[20] = {o.x, o.y, o.z, fwd.x, fwd.y, fwd.z, right.x, right.y, right.z,
                   cup.x, cup.y, cup.z, fov_scale, focus_dist, 0, 0, 0, 0, 0, 0};
    memcpy(out20, v, sizeof(v));
}

/* bvh/camera.mojo:137-154 -- one raster ray, t in [0, FLT_MAX]. */
static void raster_ray(const float *cam, uint32_t px, uint32_t py, uint32_t width, float inv_height,
                       orc_v3 *o, orc_v3 *d) {
    float sx = (2.0f * ((float)px + 0.5f) - (float)width) * inv_height;
    float sy = 1.0f - 2.0f * ((float)py + 0.5f) * inv_height;
    orc_v3 fwd = orc_v(cam[3], cam[4], cam[5]);
    orc_v3 right = orc_v(cam[6], cam[7], cam[8]);
    orc_v3 up = orc_v(cam[9], cam[10], cam[11]);
    float fov = cam[12];
    orc_v3 dir = orc_add(orc_add(fwd, orc_scale(right, sx * fov)), orc_scale(up, sy * fov));
    *o = orc_v(cam[0], cam[1], cam[2]);
    *d = orc_normalize(dir);
}

/* bvh/gpu/camera_launch.mojo:34-56: ray r -> view r/(w*h), pixel (r%w, r/w),
 * written in the field-major packed-ray layout (plane k at k*R). */
void orc_camera_rays(const float *cams, uint64_t R, uint32_t w, uint32_t h, float *rays) {
    float inv_height = 1.0f / (float)h;
    uint64_t ppv = (uint64_t)w * h;
    for (uint64_t r = 0; r < R; ++r) {
        uint64_t view = r / ppv, local = r - view * ppv;
        orc_v3 o, d;
        raster_ray(cams + 20 * view, (uint32_t)(local % w), (uint32_t)(local / w), w, inv_height, &o, &d);
        rays[0 * R + r] = o.x; rays[1 * R + r] = o.y; rays[2 * R + r] = o.z; rays[3 * R + r] = 0.0f;
        rays[4 * R + r] = d.x; rays[5 * R + r] = d.y; rays[6 * R + r] = d.z; rays[7 * R + r] = ORC_FLT_MAX;
    }
}

static orc_hit hit_miss(float t) { orc_hit h = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, t}; return h; }

static void store_hit(float *hits, uint64_t r, orc_hit h) {
    float *p = hits + 8 * r;
    p[0] = h.u; p[1] = h.v;
    memcpy(p + 2, &h.prim, 4);
    memcpy(p + 3, &h.inst, 4);
    p[4] = h.n.x; p[5] = h.n.y; p[6] = h.n.z; p[7] = h.t;
}

/* One scaled-MT leaf test that updates the closest hit. */
static int leaf_closest(const orc_bvh *b, uint32_t sorted_leaf, orc_v3 o, orc_v3 d, float t_min, orc_hit *hit) {
    uint32_t prim = b->leaf_ids[sorted_leaf];
    if (b->spheres) {
        const float *sp = b->spheres + 4 * (uint64_t)prim;
        float t;
        if (!orc_sphere_t(sp, o, d, t_min, hit->t, &t)) return 0;
        orc_sphere_record(sp, prim, o, d, t, hit);
        return 1;
    }
    const float *t = b->vertices + 9 * (uint64_t)prim;
    orc_v3 v0 = orc_v(t[0], t[1], t[2]);
    orc_v3 e1 = orc_sub(orc_v(t[3], t[4], t[5]), v0);
    orc_v3 e2 = orc_sub(orc_v(t[6], t[7], t[8]), v0);
    orc_scaled_hit s = orc_ray_tri_scaled(o, d, v0, e1, e2, hit->t, t_min);
    if (!s.mask) return 0;
    float inv = 1.0f / s.abs_det;
    hit->t = s.t_scaled * inv;
    hit->u = s.u_scaled * inv;
    hit->v = s.v_scaled * inv;
    hit->prim = prim;
    hit->inst = ORC_EMPTY;
    hit->n = orc_normalize(orc_cross(e1, e2));
    return 1;
}

static int leaf_any(const orc_bvh *b, uint32_t sorted_leaf, orc_v3 o, orc_v3 d, float t_min, float t_max) {
    uint32_t prim = b->leaf_ids[sorted_leaf];
    if (b->spheres) {
        float t;
        return orc_sphere_t(b->spheres + 4 * (uint64_t)prim, o, d, t_min, t_max, &t);
    }
    const float *t = b->vertices + 9 * (uint64_t)prim;
    orc_v3 v0 = orc_v(t[0], t[1], t[2]);
    orc_v3 e1 = orc_sub(orc_v(t[3], t[4], t[5]), v0);
    orc_v3 e2 = orc_sub(orc_v(t[6], t[7], t[8]), v0);
    return orc_ray_tri_scaled(o, d, v0, e1, e2, t_max, t_min).mask;
}

#define ORC_STACK 128

orc_hit orc_trace_closest(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max, uint32_t *stats) {
    if (b->wide) {
        orc_hit h;
        orc_trace_wide_query(b, 0, o, d, t_min, t_max, &h, NULL);
        return h;
    }
    orc_hit hit = hit_miss(t_max);
    if (b->root == ORC_SENTINEL) return hit;
    if (b->root & ORC_LEAF_FLAG) {
        leaf_closest(b, b->root & ORC_INDEX_MASK, o, d, t_min, &hit);
        if (stats) stats[3]++;
        return hit;
    }
    orc_v3 rcp = orc_rcp_dir(d);
    uint32_t stack[ORC_STACK];
    int sp = 0;
    uint32_t cur = b->root;
    for (;;) {
        if (stats) stats[0]++;
        const uint32_t *m = b->node_meta + 4 * (uint64_t)cur;
        const float *nb = b->node_bounds + 12 * (uint64_t)cur;
        float ct[2];
        int hm[2];
        hm[0] = orc_ray_aabb(o, rcp, orc_aabb_load6(nb), hit.t, &ct[0]);
        hm[1] = orc_ray_aabb(o, rcp, orc_aabb_load6(nb + 6), hit.t, &ct[1]);
        int valid[2] = {0, 0};
        uint32_t data[2] = {0, 0};
        for (int lane = 0; lane < 2; ++lane) {
            uint32_t c = m[1 + lane];
            if (!hm[lane]) continue;
            if (c & ORC_LEAF_FLAG) {
                if (stats) stats[3]++;
                leaf_closest(b, c & ORC_INDEX_MASK, o, d, t_min, &hit);
            } else {
                valid[lane] = 1;
                data[lane] = c;
                if (stats) stats[1]++;
            }
        }
        int near = -1;
        float near_t = ORC_FLT_MAX;
        if (valid[0]) { near = 0; near_t = ct[0]; }
        if (valid[1] && (near == -1 || ct[1] < near_t)) { near = 1; near_t = ct[1]; }
        for (int lane = 0; lane < 2; ++lane)
            if (valid[lane] && near != lane && ct[lane] <= hit.t) {
                if (sp >= ORC_STACK) return hit_miss(-1.0f); /* cannot happen: depth <= 64 */
                stack[sp++] = data[lane];
                if (stats && (uint32_t)sp > stats[4]) stats[4] = (uint32_t)sp;
            }
        if (near != -1 && near_t <= hit.t) { cur = data[near]; continue; }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return hit;
}

orc_hit orc_trace_any(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max) {
    if (b->wide) {
        orc_hit h;
        orc_trace_wide_query(b, 1, o, d, t_min, t_max, &h, NULL);
        return h;
    }
    orc_hit occluded = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, 0.0f};
    if (b->root == ORC_SENTINEL) return hit_miss(t_max);
    if (b->root & ORC_LEAF_FLAG) return leaf_any(b, b->root & ORC_INDEX_MASK, o, d, t_min, t_max) ? occluded : hit_miss(t_max);
    orc_v3 rcp = orc_rcp_dir(d);
    uint32_t stack[ORC_STACK];
    int sp = 0;
    uint32_t cur = b->root;
    for (;;) {
        const uint32_t *m = b->node_meta + 4 * (uint64_t)cur;
        const float *nb = b->node_bounds + 12 * (uint64_t)cur;
        for (int lane = 0; lane < 2; ++lane) {
            float tt;
            if (!orc_ray_aabb(o, rcp, orc_aabb_load6(nb + 6 * lane), t_max, &tt)) continue;
            uint32_t c = m[1 + lane];
            if (c & ORC_LEAF_FLAG) {
                if (leaf_any(b, c & ORC_INDEX_MASK, o, d, t_min, t_max)) return occluded;
            } else {
                if (sp >= ORC_STACK) return hit_miss(-1.0f);
                stack[sp++] = c;
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return hit_miss(t_max);
}

static void trace_bvh2(const orc_bvh *b, const float *rays, uint64_t R, int mode, float *hits, uint32_t *stats);

/* Trace R field-major rays against the Karras-layout BVH; mode 0 closest
 * hit, 1 any hit.  stats (optional) u32[5R]: node visits, internal child
 * hits, (unused), leaf tests, max stack depth. */
void orc_trace_bvh2(const float *vertices, const uint32_t *leaf_ids, const uint32_t *node_meta,
                    const float *node_bounds, uint32_t root, const float *rays, uint64_t R, int mode,
                    float *hits, uint32_t *stats) {
    orc_bvh b = {vertices, leaf_ids, node_meta, node_bounds, root, NULL, 0, 0, NULL};
    trace_bvh2(&b, rays, R, mode, hits, stats);
}

/* The same near-first / unordered queries over a sphere BVH (bajo's
 * GpuSphereBvh, bvh/gpu/sphere_bvh.mojo:372-478): spheres f32[4n] in input
 * order, the binary layout of the box build over their bounds. */
void orc_trace_bvh2_spheres(const float *spheres, const uint32_t *leaf_ids, const uint32_t *node_meta,
                            const float *node_bounds, uint32_t root, const float *rays, uint64_t R, int mode,
                            float *hits, uint32_t *stats) {
    orc_bvh b = {NULL, leaf_ids, node_meta, node_bounds, root, NULL, 0, 0, spheres};
    trace_bvh2(&b, rays, R, mode, hits, stats);
}

/* Closest hit over every sphere in index order (strict '<': the lowest
 * index wins exact ties), t_min as the ray carries it. */
void orc_trace_brute_spheres(const float *spheres, uint64_t n, const float *rays, uint64_t R, float *hits) {
    for (uint64_t r = 0; r < R; ++r) {
        orc_v3 o = orc_v(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
        orc_v3 d = orc_v(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
        orc_hit hit = hit_miss(rays[7 * R + r]);
        for (uint64_t i = 0; i < n; ++i) {
            float t;
            if (orc_sphere_t(spheres + 4 * i, o, d, rays[3 * R + r], hit.t, &t))
                orc_sphere_record(spheres + 4 * i, (uint32_t)i, o, d, t, &hit);
        }
        store_hit(hits, r, hit);
    }
}

static void trace_bvh2(const orc_bvh *bp, const float *rays, uint64_t R, int mode, float *hits, uint32_t *stats) {
    orc_bvh b = *bp;
    for (uint64_t r = 0; r < R; ++r) {
        orc_v3 o = orc_v(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
        orc_v3 d = orc_v(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
        float t_min = rays[3 * R + r], t_max = rays[7 * R + r];
        uint32_t *st = NULL;
        if (stats) { st = stats + 5 * r; memset(st, 0, 5 * sizeof(uint32_t)); }
        orc_hit h = mode == 1 ? orc_trace_any(&b, o, d, t_min, t_max) : orc_trace_closest(&b, o, d, t_min, t_max, st);
        store_hit(hits, r, h);
    }
}

/* test/bvh/fixtures.mojo:179-209: eager MT over every triangle, strict
 * '<' so the lowest index wins exact ties.  t_min is the fixture's 1e-4
 * default unless the ray carries a larger one. */
void orc_trace_brute(const float *vertices, uint64_t n, const float *rays, uint64_t R, float *hits) {
    for (uint64_t r = 0; r < R; ++r) {
        orc_v3 o = orc_v(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
        orc_v3 d = orc_v(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
        float t_min = rays[3 * R + r] > 1e-4f ? rays[3 * R + r] : 1e-4f;
        orc_hit hit = hit_miss(rays[7 * R + r]);
        for (uint64_t i = 0; i < n; ++i) {
            const float *t = vertices + 9 * i;
            orc_v3 v0 = orc_v(t[0], t[1], t[2]);
            orc_v3 e1 = orc_sub(orc_v(t[3], t[4], t[5]), v0);
            orc_v3 e2 = orc_sub(orc_v(t[6], t[7], t[8]), v0);
            float tt, uu, vv;
            if (orc_ray_tri_eager(o, d, v0, e1, e2, ORC_FLT_MAX, t_min, &tt, &uu, &vv) && tt < hit.t) {
                hit.t = tt; hit.u = uu; hit.v = vv; hit.prim = (uint32_t)i; hit.inst = ORC_EMPTY;
                hit.n = orc_normalize(orc_cross(e1, e2));
            }
        }
        store_hit(hits, r, hit);
    }
}

/* Known-answer hooks for the scalar intersectors (test/core/test_intersect.mojo). */
int orc_intersect_aabb(const float *o3, const float *rcp3, const float *box6, float t_max, float *t_out) {
    return orc_ray_aabb(orc_v(o3[0], o3[1], o3[2]), orc_v(rcp3[0], rcp3[1], rcp3[2]), orc_aabb_load6(box6), t_max, t_out);
}
int orc_intersect_tri_eager(const float *o3, const float *d3, const float *v0, const float *e1, const float *e2,
                            float t_max, float t_min, float *tuv3) {
    return orc_ray_tri_eager(orc_v(o3[0], o3[1], o3[2]), orc_v(d3[0], d3[1], d3[2]), orc_v(v0[0], v0[1], v0[2]),
                             orc_v(e1[0], e1[1], e1[2]), orc_v(e2[0], e2[1], e2[2]), t_max, t_min, &tuv3[0], &tuv3[1], &tuv3[2]);
}
int orc_intersect_tri_scaled(const float *o3, const float *d3, const float *v0, const float *e1, const float *e2,
                             float t_max, float t_min, float *dtuv4) {
    orc_scaled_hit s = orc_ray_tri_scaled(orc_v(o3[0], o3[1], o3[2]), orc_v(d3[0], d3[1], d3[2]), orc_v(v0[0], v0[1], v0[2]),
                                          orc_v(e1[0], e1[1], e1[2]), orc_v(e2[0], e2[1], e2[2]), t_max, t_min);
    dtuv4[0] = s.abs_det; dtuv4[1] = s.t_scaled; dtuv4[2] = s.u_scaled; dtuv4[3] = s.v_scaled;
    return s.mask;
}
void orc_rcp_direction(const float *d3, float *out3) {
    orc_v3 r = orc_rcp_dir(orc_v(d3[0], d3[1], d3[2]));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

/* Multi-threaded wrapper used only for the timed CPU baseline: the rays are
 * split into contiguous ranges, one per thread (bajo parallelizes its CPU
 * tracer over ray chunks the same way, rt/cpu/wavefront/__init__.mojo:129-158). */
#include <pthread.h>

typedef struct {
    const float *vertices; const uint32_t *leaf_ids; const uint32_t *node_meta; const float *node_bounds;
    uint32_t root; const float *rays; uint64_t R, first, last; int mode; float *hits;
} trace_job;

static void *trace_worker(void *p) {
    trace_job *j = (trace_job *)p;
    orc_bvh b = {j->vertices, j->leaf_ids, j->node_meta, j->node_bounds, j->root, NULL, 0, 0, NULL};
    uint64_t R = j->R;
    for (uint64_t r = j->first; r < j->last; ++r) {
        orc_v3 o = orc_v(j->rays[0 * R + r], j->rays[1 * R + r], j->rays[2 * R + r]);
        orc_v3 d = orc_v(j->rays[4 * R + r], j->rays[5 * R + r], j->rays[6 * R + r]);
        float t_min = j->rays[3 * R + r], t_max = j->rays[7 * R + r];
        orc_hit h = j->mode == 1 ? orc_trace_any(&b, o, d, t_min, t_max) : orc_trace_closest(&b, o, d, t_min, t_max, NULL);
        store_hit(j->hits, r, h);
    }
    return NULL;
}

void orc_trace_bvh2_mt(const float *vertices, const uint32_t *leaf_ids, const uint32_t *node_meta,
                       const float *node_bounds, uint32_t root, const float *rays, uint64_t R, int mode,
                       float *hits, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    trace_job jobs[256];
    for (int t = 0; t < threads; ++t) {
        trace_job j = {vertices, leaf_ids, node_meta, node_bounds, root, rays, R,
                       R * (uint64_t)t / (uint64_t)threads, R * (uint64_t)(t + 1) / (uint64_t)threads, mode, hits};
        jobs[t] = j;
        pthread_create(&tid[t], NULL, trace_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}
