/*
 * CPU oracle -- test infrastructure only (see orc_core.h).
 *
 * Restates bajo's LBVH build on the CPU:
 *   - triangle bounds            bvh/gpu/triangle_bvh.mojo:680-727
 *   - scene/centroid bounds      bvh/gpu/builder/binary_layout.mojo:128-204
 *   - Morton codes               bvh/gpu/builder/lbvh.mojo:30-62, core/morton.mojo:61-89
 *   - stable radix sort          bvh/cpu/builder/lbvh.mojo:36-70 (4 x 8-bit, u32 code)
 *                                sort/gpu/radix_sort.mojo:510-597 (8 x 8-bit, u64 key)
 *   - Karras topology            bvh/gpu/builder/lbvh.mojo:164-329
 *   - bottom-up refit            bvh/gpu/builder/lbvh.mojo:116-161
 *   - top-down CPU LBVH          bvh/cpu/builder/lbvh.mojo:144-163, 270-315
 *
 * Layouts follow bvh/constants.mojo:38-45: node_meta u32[4] = parent, left,
 * right, fence; node_bounds f32[12] = left child box then right child box;
 * child references carry bit 31 for leaves (bvh/tagged_ref.mojo:3-32).
 */
#include <stdlib.h>
#include <string.h>

#include "orc_core.h"
#include "oracle.h"

uint32_t orc_expand_bits(uint32_t n) { return orc_expand_bits_3d(n); }
uint32_t orc_morton(float x, float y, float z) { return orc_morton3(x, y, z); }
int orc_common_prefix(uint32_t ca, uint32_t ia, uint32_t cb, uint32_t ib) {
    return orc_morton_common_prefix(ca, ia, cb, ib);
}

/* bvh/gpu/triangle_bvh.mojo:680-727: per-triangle box = vmin(vmin(v0,v1),v2). */
void orc_triangle_bounds(const float *v, uint64_t n, float *bounds6) {
    for (uint64_t i = 0; i < n; ++i) {
        const float *t = v + 9 * i;
        orc_v3 a = orc_v(t[0], t[1], t[2]), b = orc_v(t[3], t[4], t[5]), c = orc_v(t[6], t[7], t[8]);
        orc_aabb box = {orc_vmin(orc_vmin(a, b), c), orc_vmax(orc_vmax(a, b), c)};
        orc_aabb_store6(bounds6 + 6 * i, box);
    }
}

/* binary_layout.mojo:128-164: root box (grow by leaf box) and centroid box
 * (grow by leaf centroid); out12 = root box then centroid box. */
void orc_scene_bounds(const float *bounds6, uint64_t n, float *out12) {
    orc_aabb root = orc_aabb_invalid(), cent = orc_aabb_invalid();
    for (uint64_t i = 0; i < n; ++i) {
        orc_aabb b = orc_aabb_load6(bounds6 + 6 * i);
        root = orc_aabb_merge(root, b);
        orc_aabb_grow_pt(&cent, orc_aabb_centroid(b));
    }
    orc_aabb_store6(out12, root);
    orc_aabb_store6(out12 + 6, cent);
}

/* bvh/gpu/builder/lbvh.mojo:46-58: c = (centroid - cmin) * safe_inv(extent). */
void orc_morton_codes(const float *bounds6, uint64_t n, const float *centroid_box6, uint32_t *codes) {
    orc_aabb cb = orc_aabb_load6(centroid_box6);
    orc_v3 inv = orc_safe_inv(orc_sub(cb.hi, cb.lo));
    for (uint64_t i = 0; i < n; ++i) {
        orc_aabb b = orc_aabb_load6(bounds6 + 6 * i);
        orc_v3 c = orc_mul(orc_sub(orc_aabb_centroid(b), cb.lo), inv);
        codes[i] = orc_morton3(c.x, c.y, c.z);
    }
}

/* Stable LSD radix sort of (u64 key, u32 value) pairs over 8-bit digits
 * [begin_bit, end_bit).  Restates the contract of sort/gpu/radix_sort.mojo:
 * 510-597 and sort/gpu/onesweep.mojo:454-551 (stable, ascending). */
int orc_radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, uint64_t n, int begin_bit, int end_bit) {
    if (n == 0) return 0;
    uint64_t *k2 = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *v2 = vals ? (uint32_t *)malloc(n * sizeof(uint32_t)) : NULL;
    if (!k2 || (vals && !v2)) { free(k2); free(v2); return -1; }
    uint64_t *ks = keys, *kd = k2;
    uint32_t *vs = vals, *vd = v2;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        uint64_t count[257];
        memset(count, 0, sizeof(count));
        for (uint64_t i = 0; i < n; ++i) count[((ks[i] >> shift) & 0xFF) + 1]++;
        for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t dst = count[(ks[i] >> shift) & 0xFF]++;
            kd[dst] = ks[i];
            if (vs) vd[dst] = vs[i];
        }
        uint64_t *tk = ks; ks = kd; kd = tk;
        uint32_t *tv = vs; vs = vd; vd = tv;
    }
    if (ks != keys) {
        memcpy(keys, ks, n * sizeof(uint64_t));
        if (vals) memcpy(vals, vs, n * sizeof(uint32_t));
    }
    free(k2);
    free(v2);
    return 0;
}

/* bvh/cpu/builder/lbvh.mojo:36-70: four stable 8-bit passes over the u32
 * Morton code carrying the item index. */
int orc_radix_sort_morton_pairs(uint32_t *codes, uint32_t *ids, uint64_t n) {
    uint64_t *k = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    if (!k) return -1;
    for (uint64_t i = 0; i < n; ++i) k[i] = codes[i];
    int rc = orc_radix_sort_pairs_u64(k, ids, n, 0, 32);
    for (uint64_t i = 0; i < n; ++i) codes[i] = (uint32_t)k[i];
    free(k);
    return rc;
}

/* bvh/gpu/builder/lbvh.mojo:221-232 -- -1 outside the segment. */
static int cp(const uint32_t *codes, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    return orc_morton_common_prefix(codes[i], (uint32_t)i, codes[j], (uint32_t)j);
}

/* bvh/gpu/builder/lbvh.mojo:164-191. */
static void find_range(const uint32_t *codes, int64_t n, int64_t i, int64_t *first, int64_t *last) {
    int d_next = cp(codes, n, i, i + 1);
    int d_prev = cp(codes, n, i, i - 1);
    int64_t d = d_next < d_prev ? -1 : 1;
    int delta_min = cp(codes, n, i, i - d);
    int64_t lmax = 2;
    while (cp(codes, n, i, i + lmax * d) > delta_min) {
        lmax <<= 1;
        if (lmax > n * 2) break;
    }
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t > 0; t >>= 1)
        if (cp(codes, n, i, i + (l + t) * d) > delta_min) l += t;
    int64_t j = i + l * d;
    *first = i < j ? i : j;
    *last = i < j ? j : i;
}

/* bvh/gpu/builder/lbvh.mojo:194-218 and bvh/cpu/builder/lbvh.mojo:144-163. */
static int64_t find_split(const uint32_t *codes, int64_t n, int64_t first, int64_t last) {
    int node_prefix = cp(codes, n, first, last);
    int64_t split = first, step = last - first;
    while (step > 1) {
        step = (step + 1) >> 1;
        int64_t ns = split + step;
        if (ns < last && cp(codes, n, first, ns) > node_prefix) split = ns;
    }
    return split;
}

int64_t orc_lbvh_find_split(const uint32_t *codes, int64_t n, int64_t first, int64_t last) {
    return find_split(codes, n, first, last);
}

/* bvh/gpu/builder/lbvh.mojo:235-329 (single segment): Karras hierarchy over
 * the sorted codes.  Writes node_meta[4(n-1)], leaf_parent[n],
 * leaf_counts[n-1]; returns the root reference. */
uint32_t orc_karras_topology(const uint32_t *sorted_codes, uint64_t n, uint32_t *node_meta,
                             uint32_t *leaf_parent, uint32_t *leaf_counts) {
    if (n == 0) return ORC_SENTINEL;
    if (n == 1) { leaf_parent[0] = ORC_SENTINEL; return ORC_LEAF_FLAG | 0u; }
    int64_t nn = (int64_t)n;
    for (int64_t i = 0; i < nn - 1; ++i) {
        int64_t first, last;
        find_range(sorted_codes, nn, i, &first, &last);
        leaf_counts[i] = (uint32_t)(last - first + 1);
        if (first == 0 && last == nn - 1) node_meta[4 * i + 0] = ORC_SENTINEL;
        int64_t split = find_split(sorted_codes, nn, first, last);
        uint32_t left, right;
        if (split == first) { left = ORC_LEAF_FLAG | (uint32_t)split; leaf_parent[split] = (uint32_t)i; }
        else { left = (uint32_t)split; node_meta[4 * split + 0] = (uint32_t)i; }
        if (split + 1 == last) { right = ORC_LEAF_FLAG | (uint32_t)(split + 1); leaf_parent[split + 1] = (uint32_t)i; }
        else { right = (uint32_t)(split + 1); node_meta[4 * (split + 1) + 0] = (uint32_t)i; }
        node_meta[4 * i + 1] = left;
        node_meta[4 * i + 2] = right;
        node_meta[4 * i + 3] = (uint32_t)last;
    }
    return 0u;
}

/* Post-order refit: node_bounds[i] = (box(left), box(right)); a box of an
 * internal child is merge(left box, right box), the same fold the atomic
 * bottom-up refit of bvh/gpu/builder/lbvh.mojo:116-161 performs. */
static orc_aabb refit_rec(uint32_t ref, const float *leaf_bounds, const uint32_t *leaf_ids,
                          const uint32_t *node_meta, float *node_bounds, int depth) {
    if (ref & ORC_LEAF_FLAG) return orc_aabb_load6(leaf_bounds + 6 * (uint64_t)leaf_ids[ref & ORC_INDEX_MASK]);
    (void)depth;
    orc_aabb l = refit_rec(node_meta[4 * (uint64_t)ref + 1], leaf_bounds, leaf_ids, node_meta, node_bounds, depth + 1);
    orc_aabb r = refit_rec(node_meta[4 * (uint64_t)ref + 2], leaf_bounds, leaf_ids, node_meta, node_bounds, depth + 1);
    orc_aabb_store6(node_bounds + 12 * (uint64_t)ref, l);
    orc_aabb_store6(node_bounds + 12 * (uint64_t)ref + 6, r);
    return orc_aabb_merge(l, r);
}

void orc_refit(uint32_t root, const float *leaf_bounds, const uint32_t *leaf_ids, const uint32_t *node_meta,
               float *node_bounds) {
    if (root == ORC_SENTINEL || (root & ORC_LEAF_FLAG)) return;
    refit_rec(root, leaf_bounds, leaf_ids, node_meta, node_bounds, 0);
}

/* The GPU refit's bottom-up climb (refit_lbvh_bounds_from_leaves_kernel,
 * bvh/gpu/builder/lbvh.mojo:116-161), run leaf by leaf: a leaf's box climbs
 * through leaf_parent / node parents; the first child to arrive at a node
 * stops, the second merges (left then right) and climbs on.  It restates the
 * library's corruption checks: a parent index outside the node array, a node
 * that does not list the arriving child, or a climb past 128 levels stops
 * that climb and makes the result 5 (BAJO_STATUS_REFIT_CORRUPT); 0 otherwise.
 * leaf_bounds are indexed by sorted position; node_bounds receive each
 * node's left box at +0 and right box at +6. */
uint32_t orc_refit_climb(const float *leaf_bounds, uint64_t n, const uint32_t *node_meta,
                         const uint32_t *leaf_parent, float *node_bounds) {
    if (n < 2) return 0;
    const uint64_t m = n - 1;
    uint32_t status = 0;
    uint8_t *arrived = (uint8_t *)calloc(m, 1);
    for (uint64_t s = 0; s < n; ++s) {
        orc_aabb box = orc_aabb_load6(leaf_bounds + 6 * s);
        uint32_t cur = ORC_LEAF_FLAG | (uint32_t)s;
        uint32_t parent = leaf_parent[s];
        int depth = 0;
        for (; parent != ORC_SENTINEL && depth < 128; ++depth) {
            if (parent >= m) { status = 5; break; }
            const uint32_t *meta = node_meta + 4 * (uint64_t)parent;
            const int is_left = cur == meta[1];
            if (!is_left && cur != meta[2]) { status = 5; break; }
            orc_aabb_store6(node_bounds + 12 * (uint64_t)parent + (is_left ? 0 : 6), box);
            if (arrived[parent]++ == 0) break;
            orc_aabb other = orc_aabb_load6(node_bounds + 12 * (uint64_t)parent + (is_left ? 6 : 0));
            box = is_left ? orc_aabb_merge(box, other) : orc_aabb_merge(other, box);
            cur = parent;
            parent = meta[0];
        }
        if (depth == 128) status = 5;
    }
    free(arrived);
    return status;
}

/* Full single-segment LBVH in the GPU builder's layout.  Outputs (caller
 * allocated): sorted_keys u64[n] = (code<<32)|prim, leaf_ids u32[n],
 * node_meta u32[4(n-1)], leaf_parent u32[n], leaf_counts u32[n-1],
 * node_bounds f32[12(n-1)], scene f32[12] (root box, centroid box). */
uint32_t orc_build_lbvh(const float *vertices, uint64_t n, uint64_t *sorted_keys, uint32_t *leaf_ids,
                        uint32_t *node_meta, uint32_t *leaf_parent, uint32_t *leaf_counts, float *node_bounds,
                        float *scene12) {
    float *bounds = (float *)malloc((n ? n : 1) * 6 * sizeof(float));
    orc_triangle_bounds(vertices, n, bounds);
    uint32_t root = orc_build_lbvh_bounds(bounds, n, sorted_keys, leaf_ids, node_meta, leaf_parent, leaf_counts,
                                          node_bounds, scene12);
    free(bounds);
    return root;
}

/* The same build over caller leaf boxes f32[6n] (lo.xyz, hi.xyz), bajo's
 * GpuBinaryBoundsBvh input (bvh/gpu/builder/binary_layout.mojo:331-428). */
uint32_t orc_build_lbvh_bounds(const float *bounds, uint64_t n, uint64_t *sorted_keys, uint32_t *leaf_ids,
                               uint32_t *node_meta, uint32_t *leaf_parent, uint32_t *leaf_counts, float *node_bounds,
                               float *scene12) {
    uint32_t *codes = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    orc_scene_bounds(bounds, n, scene12);
    orc_morton_codes(bounds, n, scene12 + 6, codes);
    for (uint64_t i = 0; i < n; ++i) { sorted_keys[i] = ((uint64_t)codes[i] << 32) | (uint64_t)i; leaf_ids[i] = (uint32_t)i; }
    orc_radix_sort_pairs_u64(sorted_keys, leaf_ids, n, 0, 64);
    for (uint64_t i = 0; i < n; ++i) codes[i] = (uint32_t)(sorted_keys[i] >> 32);
    uint32_t root = orc_karras_topology(codes, n, node_meta, leaf_parent, leaf_counts);
    orc_refit(root, bounds, leaf_ids, node_meta, node_bounds);
    free(codes);
    return root;
}

/* ---- top-down CPU LBVH (bvh/cpu/builder/lbvh.mojo:219-315) --------------
 * Nodes: box f32[6], then u32 a, u32 count: count == 0 -> internal with
 * children a, a+1; count > 0 -> leaf over sorted items [a, a+count). */
typedef struct {
    const uint32_t *codes;
    const uint32_t *ids;
    const float *bounds;
    int64_t n;
    int leaf_size;
    float *boxes;
    uint32_t *first;
    uint32_t *count;
    uint32_t used;
} td_state;

static orc_aabb td_rec(td_state *s, uint32_t node, int64_t first, int64_t count) {
    if (count <= s->leaf_size) {
        orc_aabb box = orc_aabb_invalid();
        for (int64_t i = 0; i < count; ++i)
            box = orc_aabb_merge(box, orc_aabb_load6(s->bounds + 6 * (uint64_t)s->ids[first + i]));
        orc_aabb_store6(s->boxes + 6 * (uint64_t)node, box);
        s->first[node] = (uint32_t)first;
        s->count[node] = (uint32_t)count;
        return box;
    }
    int64_t last = first + count - 1;
    int64_t split = find_split(s->codes, s->n, first, last);
    int64_t lc = split - first + 1;
    uint32_t left = s->used;
    s->used += 2;
    orc_aabb lb = td_rec(s, left, first, lc);
    orc_aabb rb = td_rec(s, left + 1, split + 1, count - lc);
    orc_aabb box = orc_aabb_merge(lb, rb);
    orc_aabb_store6(s->boxes + 6 * (uint64_t)node, box);
    s->first[node] = left;
    s->count[node] = 0;
    return box;
}

/* Returns the number of nodes written (root = node 0); node arrays need
 * room for 2n-1 entries. */
uint32_t orc_build_lbvh_topdown(const uint32_t *sorted_codes, const uint32_t *sorted_ids, const float *bounds6,
                                uint64_t n, int leaf_size, float *boxes6, uint32_t *first, uint32_t *count) {
    if (n == 0) return 0;
    td_state s = {sorted_codes, sorted_ids, bounds6, (int64_t)n, leaf_size < 1 ? 1 : leaf_size, boxes6, first, count, 1};
    td_rec(&s, 0, 0, (int64_t)n);
    return s.used;
}

/* Canonical topology: one (first, last, split) triple per internal node over
 * sorted positions, listed in pre-order from the root (left subtree first).
 * Both numberings above map onto it, so it is the object compared bit-exact. */
static uint64_t canon_karras_rec(uint32_t ref, const uint32_t *node_meta, const uint32_t *leaf_counts,
                                 uint32_t *out, uint64_t k, uint32_t *lo, uint32_t *hi) {
    if (ref & ORC_LEAF_FLAG) { *lo = *hi = ref & ORC_INDEX_MASK; return k; }
    uint64_t slot = k++;
    uint32_t l0, l1, r0, r1;
    k = canon_karras_rec(node_meta[4 * (uint64_t)ref + 1], node_meta, leaf_counts, out, k, &l0, &l1);
    k = canon_karras_rec(node_meta[4 * (uint64_t)ref + 2], node_meta, leaf_counts, out, k, &r0, &r1);
    out[3 * slot + 0] = l0;
    out[3 * slot + 1] = r1;
    out[3 * slot + 2] = l1;
    *lo = l0;
    *hi = r1;
    return k;
}

uint64_t orc_canonical_karras(uint32_t root, const uint32_t *node_meta, const uint32_t *leaf_counts, uint32_t *out) {
    uint32_t lo, hi;
    if (root == ORC_SENTINEL) return 0;
    return canon_karras_rec(root, node_meta, leaf_counts, out, 0, &lo, &hi);
}

static uint64_t canon_td_rec(uint32_t node, const uint32_t *first, const uint32_t *count, uint32_t *out,
                             uint64_t k, uint32_t *lo, uint32_t *hi) {
    if (count[node] > 0) { *lo = first[node]; *hi = first[node] + count[node] - 1; return k; }
    uint64_t slot = k++;
    uint32_t l0, l1, r0, r1;
    k = canon_td_rec(first[node], first, count, out, k, &l0, &l1);
    k = canon_td_rec(first[node] + 1, first, count, out, k, &r0, &r1);
    out[3 * slot + 0] = l0;
    out[3 * slot + 1] = r1;
    out[3 * slot + 2] = l1;
    *lo = l0;
    *hi = r1;
    return k;
}

uint64_t orc_canonical_topdown(const uint32_t *first, const uint32_t *count, uint32_t *out) {
    uint32_t lo, hi;
    return canon_td_rec(0, first, count, out, 0, &lo, &hi);
}
