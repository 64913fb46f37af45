/*
 * CPU oracle -- test infrastructure only (see orc_core.h).
 *
 * Restates the binary -> 8-wide collapse with quantized child boxes and the
 * traversal of the resulting nodes:
 *   - collapse rule: open the child of largest surface area until a node has
 *     8 children, first among children with more triangles than the leaf
 *     size, then among any inner child; subtrees of at most the leaf size
 *     become leaf children       bvh/gpu/builder/wide_collapse.mojo:148-643
 *                                (fat leaves <= min(L, 4), largest-area opening)
 *   - power-of-two per-axis grid, floor / ceil quantization of child boxes
 *                                bvh/gpu/compressed_bounds_bvh.mojo:43-76, 84-326
 *   - near-first stack traversal, closest hit pruned by hit.t, any hit exits
 *     on the first occluder       bvh/gpu/triangle_bvh.mojo:1010-1104,
 *                                bvh/gpu/compressed_bounds_bvh.mojo:469-576
 *
 * The 96-byte node record is this library's own (origin f32[3], biased
 * exponents u8[3], child count u8, lo_x..hi_z u8[8] each, child u32[8]); a
 * leaf child holds (count-1) << 28 | first sorted triangle with bit 31 set.
 * Numbering is breadth-first with children in slot order; the device build
 * numbers nodes within a level in the order waves reserve them, so tests
 * compare the two hierarchies in a canonical depth-first order.
 */
#include <stdlib.h>
#include <string.h>

#include "orc_core.h"
#include "oracle.h"
#include "orc_internal.h"

typedef struct {
    float ox, oy, oz;
    uint8_t ex, ey, ez, n;
    uint8_t lo[3][8], hi[3][8];
    uint32_t child[8];
} orc_wide_node;

typedef struct { uint32_t ref, count; float b[6]; } wchild;

#define ORC_WIDE_START_MASK 0x0FFFFFFFu

static float bits_to_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t float_to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float half_area(const float *b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dx * dz + dy * dz;
}

static void load_children(const uint32_t *meta, const float *nb, const uint32_t *counts, uint32_t node,
                          wchild *l, wchild *r) {
    memcpy(l->b, nb + 12 * (uint64_t)node, 6 * sizeof(float));
    memcpy(r->b, nb + 12 * (uint64_t)node + 6, 6 * sizeof(float));
    l->ref = meta[4 * (uint64_t)node + 1];
    r->ref = meta[4 * (uint64_t)node + 2];
    l->count = (l->ref & ORC_LEAF_FLAG) ? 1u : counts[l->ref];
    r->count = (r->ref & ORC_LEAF_FLAG) ? 1u : counts[r->ref];
}

/* Smallest biased exponent e (1..253) with o + 254 * 2^(e-127) >= hi. */
uint32_t orc_grid_exponent(float o, float hi) {
    float extent = hi - o;
    uint32_t e = 1;
    if (extent > 0.0f) {
        uint32_t bits = float_to_bits(extent / 254.0f);
        e = (bits >> 23) & 0xFFu;
        if (bits & 0x7FFFFFu) e += 1;
        if (e < 1) e = 1;
    }
    while (e < 253) {
        float s = bits_to_float(e << 23);
        if (o + 254.0f * s >= hi) break;
        ++e;
    }
    return e;
}

/* Grid cells [qlo, qhi] of [lo, hi] on the grid o + q * 2^(e-127), widened
 * by one cell on each side. */
void orc_quantize(float o, uint32_t e, float lo, float hi, uint8_t *qlo, uint8_t *qhi) {
    float s = bits_to_float(e << 23);
    float inv = bits_to_float((254u - e) << 23);
    float fl = floorf((lo - o) * inv);
    fl = fl < 0.0f ? 0.0f : (fl > 255.0f ? 255.0f : fl);
    while (fl > 0.0f && o + fl * s > lo) fl -= 1.0f;
    float fh = ceilf((hi - o) * inv);
    fh = fh < 0.0f ? 0.0f : (fh > 255.0f ? 255.0f : fh);
    while (fh < 255.0f && o + fh * s < hi) fh += 1.0f;
    fl = fl > 0.0f ? fl - 1.0f : 0.0f;
    fh = fh < 255.0f ? fh + 1.0f : 255.0f;
    *qlo = (uint8_t)fl;
    *qhi = (uint8_t)fh;
}

static void empty_slots(orc_wide_node *w, int from) {
    for (int i = from; i < 8; ++i) {
        for (int a = 0; a < 3; ++a) { w->lo[a][i] = 255; w->hi[a][i] = 0; }
        w->child[i] = 0;
    }
}

uint64_t orc_collapse_wide(const uint32_t *node_meta, const float *node_bounds, const uint32_t *leaf_counts,
                           uint64_t n, uint32_t root, const float *scene_bounds, uint32_t leaf_size, void *out_nodes,
                           uint64_t capacity) {
    return orc_collapse_wide_n(node_meta, node_bounds, leaf_counts, n, root, scene_bounds, leaf_size, 8, out_nodes,
                               capacity);
}

/* The collapse at width W (4 or 8): nodes open children until they hold W
 * (bajo's W parameter, bvh/gpu/builder/wide_collapse.mojo:148-643), in the
 * same node record with W live slots. */
uint64_t orc_collapse_wide_n(const uint32_t *node_meta, const float *node_bounds, const uint32_t *leaf_counts,
                             uint64_t n, uint32_t root, const float *scene_bounds, uint32_t leaf_size, uint32_t width,
                             void *out_nodes, uint64_t capacity) {
    orc_wide_node *out = (orc_wide_node *)out_nodes;
    if (n == 0 || capacity == 0 || leaf_size < 1 || leaf_size > 8 || (width != 4 && width != 8)) return 0;
    if (n == 1) {
        orc_wide_node w;
        w.ox = scene_bounds[0]; w.oy = scene_bounds[1]; w.oz = scene_bounds[2];
        w.ex = (uint8_t)orc_grid_exponent(scene_bounds[0], scene_bounds[3]);
        w.ey = (uint8_t)orc_grid_exponent(scene_bounds[1], scene_bounds[4]);
        w.ez = (uint8_t)orc_grid_exponent(scene_bounds[2], scene_bounds[5]);
        w.n = 1;
        empty_slots(&w, 1);
        uint32_t e[3] = {w.ex, w.ey, w.ez};
        for (int a = 0; a < 3; ++a)
            orc_quantize(scene_bounds[a], e[a], scene_bounds[a], scene_bounds[a + 3], &w.lo[a][0], &w.hi[a][0]);
        w.child[0] = ORC_LEAF_FLAG;
        out[0] = w;
        return 1;
    }
    uint32_t *src = (uint32_t *)malloc(capacity * sizeof(uint32_t));
    src[0] = root; /* binary root: an inner node when n > 1 */
    uint64_t count = 1;
    for (uint64_t w = 0; w < count; ++w) {
        wchild ch[8];
        int nc = 2;
        load_children(node_meta, node_bounds, leaf_counts, src[w], &ch[0], &ch[1]);
        for (int phase = 0; phase < 2; ++phase) {
            while (nc < (int)width) {
                int pick = -1;
                float best = -1.0f;
                for (int i = 0; i < nc; ++i) {
                    int inner = !(ch[i].ref & ORC_LEAF_FLAG);
                    int eligible = inner && (phase == 0 ? ch[i].count > leaf_size : ch[i].count >= 2);
                    float a = half_area(ch[i].b);
                    if (eligible && a > best) { best = a; pick = i; }
                }
                if (pick < 0) break;
                wchild l, r;
                load_children(node_meta, node_bounds, leaf_counts, ch[pick].ref, &l, &r);
                ch[pick] = l;
                ch[nc++] = r;
            }
        }
        uint64_t k = 0;
        for (int i = 0; i < nc; ++i)
            if (!(ch[i].ref & ORC_LEAF_FLAG) && ch[i].count > leaf_size) ++k;
        if (count + k > capacity) { free(src); return 0; }
        float box[6] = {ORC_FLT_MAX, ORC_FLT_MAX, ORC_FLT_MAX, -ORC_FLT_MAX, -ORC_FLT_MAX, -ORC_FLT_MAX};
        for (int i = 0; i < nc; ++i)
            for (int a = 0; a < 3; ++a) {
                box[a] = orc_fmin(box[a], ch[i].b[a]);
                box[a + 3] = orc_fmax(box[a + 3], ch[i].b[a + 3]);
            }
        orc_wide_node o;
        o.ox = box[0]; o.oy = box[1]; o.oz = box[2];
        uint32_t e[3];
        for (int a = 0; a < 3; ++a) e[a] = orc_grid_exponent(box[a], box[a + 3]);
        o.ex = (uint8_t)e[0]; o.ey = (uint8_t)e[1]; o.ez = (uint8_t)e[2];
        o.n = (uint8_t)nc;
        for (int i = 0; i < nc; ++i) {
            for (int a = 0; a < 3; ++a) orc_quantize(box[a], e[a], ch[i].b[a], ch[i].b[a + 3], &o.lo[a][i], &o.hi[a][i]);
            uint32_t ref = ch[i].ref;
            if (ref & ORC_LEAF_FLAG) {
                o.child[i] = ORC_LEAF_FLAG | (ref & ORC_INDEX_MASK);
            } else if (ch[i].count <= leaf_size) {
                uint32_t fence = node_meta[4 * (uint64_t)ref + 3];
                o.child[i] = ORC_LEAF_FLAG | ((ch[i].count - 1u) << 28) | (fence + 1u - ch[i].count);
            } else {
                o.child[i] = (uint32_t)count;
                src[count++] = ref;
            }
        }
        empty_slots(&o, nc);
        out[w] = o;
    }
    free(src);
    return count;
}

/* ---- traversal ---------------------------------------------------------- */

static int wide_leaf_closest(const orc_bvh *b, uint32_t s, orc_v3 o, orc_v3 d, float t_min, orc_hit *hit) {
    uint32_t prim = b->leaf_ids[s];
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
    orc_scaled_hit sh = orc_ray_tri_scaled(o, d, v0, e1, e2, hit->t, t_min);
    if (!sh.mask) return 0;
    float inv = 1.0f / sh.abs_det;
    hit->t = sh.t_scaled * inv;
    hit->u = sh.u_scaled * inv;
    hit->v = sh.v_scaled * inv;
    hit->prim = prim;
    hit->inst = ORC_EMPTY;
    hit->n = orc_normalize(orc_cross(e1, e2));
    return 1;
}

static int wide_leaf_any(const orc_bvh *b, uint32_t s, orc_v3 o, orc_v3 d, float t_min, float t_max) {
    uint32_t prim = b->leaf_ids[s];
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

#define ORC_WIDE_STACK 256

/* stats (may be NULL): [0] node visits, [1] inner children entered, [3]
 * triangle tests, [4] deepest stack.  Returns 0, or 1 on stack overflow. */
int orc_trace_wide_query(const orc_bvh *b, int any, orc_v3 o, orc_v3 d, float t_min, float t_max, orc_hit *out,
                         uint32_t *stats) {
    const orc_wide_node *nodes = (const orc_wide_node *)b->wide;
    orc_hit hit = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, t_max};
    orc_v3 rcp = orc_rcp_dir(d);
    int pos[3] = {rcp.x >= 0.0f, rcp.y >= 0.0f, rcp.z >= 0.0f};
    float ro[3] = {o.x, o.y, o.z}, rr[3] = {rcp.x, rcp.y, rcp.z};
    uint32_t stack[ORC_WIDE_STACK];
    int sp = 0, occluded = 0, overflow = 0;
    uint32_t visits = 0, inner = 0, leaves = 0, max_sp = 0;
    uint32_t cur = b->wide_count ? 0u : ORC_SENTINEL;
    while (cur != ORC_SENTINEL) {
        const orc_wide_node *w = nodes + cur;
        visits++;
        float org[3] = {w->ox, w->oy, w->oz};
        uint8_t ex[3] = {w->ex, w->ey, w->ez};
        float A[3], B[3];
        for (int a = 0; a < 3; ++a) {
            A[a] = (org[a] - ro[a]) * rr[a];
            B[a] = bits_to_float((uint32_t)ex[a] << 23) * rr[a];
        }
        float node_tmax = any ? t_max : hit.t;
        float th[8];
        uint32_t inner_mask = 0, leaf_mask = 0;
        for (int i = 0; i < 8; ++i) {
            float tn[3], tf[3];
            for (int a = 0; a < 3; ++a) {
                float qn = (float)(pos[a] ? w->lo[a][i] : w->hi[a][i]);
                float qf = (float)(pos[a] ? w->hi[a][i] : w->lo[a][i]);
                tn[a] = fmaf(qn, B[a], A[a]);
                tf[a] = fmaf(qf, B[a], A[a]);
            }
            float tmin = fmaxf(fmaxf(tn[0], tn[1]), fmaxf(tn[2], 0.0f));
            float tmax = fminf(fminf(tf[0], tf[1]), fminf(tf[2], node_tmax));
            th[i] = tmin;
            if (i < w->n && tmin <= tmax) {
                if (w->child[i] & ORC_LEAF_FLAG) leaf_mask |= 1u << i;
                else inner_mask |= 1u << i;
            }
        }
        for (int i = 0; i < 8 && !occluded; ++i) {
            if (!((leaf_mask >> i) & 1u)) continue;
            uint32_t ref = w->child[i];
            uint32_t s = ref & ORC_WIDE_START_MASK, e = s + ((ref >> 28) & 7u) + 1u;
            if (e > b->prim_count) e = (uint32_t)b->prim_count;
            for (; s < e; ++s) {
                leaves++;
                if (any) {
                    if (wide_leaf_any(b, s, o, d, t_min, t_max)) { occluded = 1; break; }
                } else {
                    wide_leaf_closest(b, s, o, d, t_min, &hit);
                }
            }
        }
        if (occluded) break;
        if (!any)
            for (int i = 0; i < 8; ++i)
                if (th[i] > hit.t) inner_mask &= ~(1u << i);
        inner += (uint32_t)__builtin_popcount(inner_mask);
        if (inner_mask) {
            while (inner_mask & (inner_mask - 1u)) {
                int far = 0;
                float far_t = -1.0f;
                for (int i = 0; i < 8; ++i)
                    if (((inner_mask >> i) & 1u) && th[i] >= far_t) { far_t = th[i]; far = i; }
                inner_mask &= ~(1u << far);
                if (sp >= ORC_WIDE_STACK) { overflow = 1; break; }
                stack[sp++] = w->child[far];
            }
            if (overflow) break;
            if ((uint32_t)sp > max_sp) max_sp = (uint32_t)sp;
            cur = w->child[__builtin_ctz(inner_mask)];
            continue;
        }
        cur = sp == 0 ? ORC_SENTINEL : stack[--sp];
    }
    if (any) {
        orc_hit sh = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, 0.0f};
        orc_hit ms = {0, 0, ORC_EMPTY, ORC_EMPTY, {0, 0, 0}, t_max};
        hit = occluded ? sh : ms;
    }
    if (stats) { stats[0] = visits; stats[1] = inner; stats[2] = 0; stats[3] = leaves; stats[4] = max_sp; }
    *out = hit;
    return overflow;
}

/* Field-major rays f32[8R] -> AoS hits f32[8R]; stats u32[5R] or NULL.
 * Returns the number of rays whose traversal overflowed the stack. */
static uint64_t trace_wide(const orc_bvh *b, const float *rays, uint64_t R, int mode, float *hits, uint32_t *stats);

uint64_t orc_trace_wide(const float *vertices, uint64_t n, const uint32_t *leaf_ids, const void *wide,
                        uint64_t wide_count, const float *rays, uint64_t R, int mode, float *hits, uint32_t *stats) {
    orc_bvh b;
    memset(&b, 0, sizeof(b));
    b.vertices = vertices;
    b.leaf_ids = leaf_ids;
    b.wide = wide;
    b.wide_count = wide_count;
    b.prim_count = n;
    return trace_wide(&b, rays, R, mode, hits, stats);
}

/* The 8-wide queries over a sphere BVH (spheres f32[4n], input order). */
uint64_t orc_trace_wide_spheres(const float *spheres, uint64_t n, const uint32_t *leaf_ids, const void *wide,
                                uint64_t wide_count, const float *rays, uint64_t R, int mode, float *hits,
                                uint32_t *stats) {
    orc_bvh b;
    memset(&b, 0, sizeof(b));
    b.spheres = spheres;
    b.leaf_ids = leaf_ids;
    b.wide = wide;
    b.wide_count = wide_count;
    b.prim_count = n;
    return trace_wide(&b, rays, R, mode, hits, stats);
}

static uint64_t trace_wide(const orc_bvh *bp, const float *rays, uint64_t R, int mode, float *hits, uint32_t *stats) {
    orc_bvh b = *bp;
    uint64_t overflows = 0;
    for (uint64_t r = 0; r < R; ++r) {
        orc_v3 o = orc_v(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
        orc_v3 d = orc_v(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
        orc_hit h;
        overflows += (uint64_t)orc_trace_wide_query(&b, mode == 1, o, d, rays[3 * R + r], rays[7 * R + r], &h,
                                                    stats ? stats + 5 * r : NULL);
        float *p = hits + 8 * r;
        p[0] = h.u; p[1] = h.v;
        memcpy(p + 2, &h.prim, 4);
        memcpy(p + 3, &h.inst, 4);
        p[4] = h.n.x; p[5] = h.n.y; p[6] = h.n.z; p[7] = h.t;
    }
    return overflows;
}
