/*
 * CPU oracle -- test infrastructure only (see orc_core.h).
 *
 * Restates bajo's deterministic H-PLOC topology (bvh/cpu/builder/hploc.mojo:
 * 855-917, build_hploc_topology):
 *   - the LBVH split of the sorted Morton range guides a recursion
 *     (bvh/cpu/builder/hploc.mojo:602-663, split = bvh/cpu/builder/lbvh.mojo:
 *     144-163, the Karras split over (code, sorted position));
 *   - a leaf range yields one cluster; an inner range concatenates the
 *     cluster lists of its two halves (left then right) and, when it holds
 *     more than the merging threshold (1 at the root), runs merge rounds
 *     until it does not (:573-599);
 *   - one merge round (:450-570): every cluster finds the neighbour within
 *     the search radius whose merged box has the smallest surface area,
 *     trying right then left at radius 1, 2, ... with a strict '<'
 *     (:351-377); mutual pairs (i < j) merge into a new node numbered in
 *     cluster order; the list is compacted in order (the merged node takes
 *     the lower position); a round without a mutual pair merges the first
 *     two clusters (:470-500);
 *   - surface area = 2 (dx dy + dx dz + dy dz) of the merged box (:130-142).
 * Nodes 0..n-1 are the leaves (by leaf id), inner nodes are numbered from n
 * in creation order; leaf bounds are indexed by leaf id.
 */
#include <stdlib.h>
#include <string.h>

#include "orc_core.h"
#include "oracle.h"

typedef struct {
    const uint32_t *codes, *ids;
    uint64_t n;
    int radius, threshold;
    uint32_t *parent, *left, *right, *leaf_id, *leaf_count;
    float *box; /* 6 per node */
    uint32_t *clusters;
    uint32_t next;
    uint64_t stats[6]; /* guide nodes, merge calls, merge rounds, hierarchical rounds, final rounds, max clusters */
} hp_state;

static int hp_prefix(const hp_state *s, int64_t i, int64_t j) {
    if (j < 0 || j >= (int64_t)s->n) return -1;
    return orc_morton_common_prefix(s->codes[i], (uint32_t)i, s->codes[j], (uint32_t)j);
}

static int64_t hp_split(const hp_state *s, int64_t first, int64_t last) {
    int node_prefix = hp_prefix(s, first, last);
    int64_t split = first, step = last - first;
    while (step > 1) {
        step = (step + 1) >> 1;
        int64_t ns = split + step;
        if (ns < last && hp_prefix(s, first, ns) > node_prefix) split = ns;
    }
    return split;
}

static float hp_merged_area(const float *a, const float *b) {
    float dx = orc_fmax(a[3], b[3]) - orc_fmin(a[0], b[0]);
    float dy = orc_fmax(a[4], b[4]) - orc_fmin(a[1], b[1]);
    float dz = orc_fmax(a[5], b[5]) - orc_fmin(a[2], b[2]);
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

static void hp_write_node(hp_state *s, uint32_t idx, uint32_t a, uint32_t b) {
    float *o = s->box + 6 * (uint64_t)idx;
    const float *x = s->box + 6 * (uint64_t)a, *y = s->box + 6 * (uint64_t)b;
    for (int k = 0; k < 3; ++k) {
        o[k] = orc_fmin(x[k], y[k]);
        o[k + 3] = orc_fmax(x[k + 3], y[k + 3]);
    }
    s->parent[idx] = ORC_SENTINEL;
    s->left[idx] = a;
    s->right[idx] = b;
    s->leaf_id[idx] = ORC_SENTINEL;
    s->leaf_count[idx] = s->leaf_count[a] + s->leaf_count[b];
    s->parent[a] = idx;
    s->parent[b] = idx;
}

/* One merge round over the cluster list at clusters[first .. first+count). */
static int hp_round(hp_state *s, int64_t first, int count) {
    uint32_t *cl = s->clusters + first;
    int nearest[64];
    for (int p = 0; p < count; ++p) {
        float best_area = ORC_FLT_MAX;
        int best = -1;
        const float *bp = s->box + 6 * (uint64_t)cl[p];
        for (int r = 1; r <= s->radius; ++r) {
            int right = p + r;
            if (right < count) {
                float a = hp_merged_area(bp, s->box + 6 * (uint64_t)cl[right]);
                if (a < best_area) { best_area = a; best = right; }
            }
            int left = p - r;
            if (left >= 0) {
                float a = hp_merged_area(bp, s->box + 6 * (uint64_t)cl[left]);
                if (a < best_area) { best_area = a; best = left; }
            }
        }
        nearest[p] = best;
    }
    int merges = 0;
    for (int p = 0; p < count; ++p)
        if (nearest[p] >= 0 && nearest[nearest[p]] == p && p < nearest[p]) merges++;
    if (merges == 0) {
        uint32_t idx = s->next++;
        hp_write_node(s, idx, cl[0], cl[1]);
        cl[0] = idx;
        for (int p = 2; p < count; ++p) cl[p - 1] = cl[p];
        return count - 1;
    }
    uint32_t base = s->next;
    s->next += (uint32_t)merges;
    int rank = 0, out = 0;
    uint32_t tmp[64];
    for (int p = 0; p < count; ++p) {
        int nb = nearest[p];
        int mutual = nb >= 0 && nearest[nb] == p;
        if (mutual && p < nb) {
            uint32_t idx = base + (uint32_t)rank++;
            hp_write_node(s, idx, cl[p], cl[nb]);
            tmp[out++] = idx;
        } else if (!mutual) {
            tmp[out++] = cl[p];
        }
    }
    memcpy(cl, tmp, (size_t)out * sizeof(uint32_t));
    return out;
}

static int hp_reduce(hp_state *s, int64_t first, int count, int threshold, int final) {
    if (count <= threshold) return count;
    s->stats[1]++;
    if ((uint64_t)count > s->stats[5]) s->stats[5] = (uint64_t)count;
    while (count > threshold) {
        count = hp_round(s, first, count);
        s->stats[2]++;
        s->stats[final ? 4 : 3]++;
    }
    return count;
}

static int hp_range(hp_state *s, int64_t first, int64_t last) {
    if (first == last) {
        s->clusters[first] = s->ids[first];
        return 1;
    }
    s->stats[0]++;
    int64_t split = hp_split(s, first, last);
    int lc = hp_range(s, first, split);
    int rc = hp_range(s, split + 1, last);
    for (int i = 0; i < rc; ++i) s->clusters[first + lc + i] = s->clusters[split + 1 + i];
    int final = first == 0 && last == (int64_t)s->n - 1;
    return hp_reduce(s, first, lc + rc, final ? 1 : s->threshold, final);
}

/* leaf_bounds6: f32[6n] by leaf id.  Outputs sized 2n-1 (bounds 6(2n-1)).
 * Returns the root node index, or ORC_SENTINEL for invalid parameters. */
uint32_t orc_hploc_topology(const float *leaf_bounds6, const uint32_t *sorted_codes, const uint32_t *sorted_ids,
                            uint64_t n, int search_radius, int merging_threshold, uint32_t *parent, uint32_t *left,
                            uint32_t *right, uint32_t *leaf_id, uint32_t *leaf_count, float *bounds6,
                            uint64_t *stats6) {
    if (n == 0 || search_radius < 1 || merging_threshold < 1 || merging_threshold > 32) return ORC_SENTINEL;
    hp_state s;
    memset(&s, 0, sizeof(s));
    s.codes = sorted_codes; s.ids = sorted_ids; s.n = n;
    s.radius = search_radius; s.threshold = merging_threshold;
    s.parent = parent; s.left = left; s.right = right; s.leaf_id = leaf_id; s.leaf_count = leaf_count;
    s.box = bounds6;
    for (uint64_t i = 0; i < n; ++i) {
        parent[i] = ORC_SENTINEL; left[i] = ORC_SENTINEL; right[i] = ORC_SENTINEL;
        leaf_id[i] = (uint32_t)i; leaf_count[i] = 1;
        memcpy(bounds6 + 6 * i, leaf_bounds6 + 6 * i, 6 * sizeof(float));
    }
    s.clusters = (uint32_t *)malloc(n * sizeof(uint32_t));
    s.next = (uint32_t)n;
    int roots = hp_range(&s, 0, (int64_t)n - 1);
    uint32_t root = s.clusters[0];
    free(s.clusters);
    if (stats6) memcpy(stats6, s.stats, sizeof(s.stats));
    if (roots != 1 || s.next != 2 * n - 1) return ORC_SENTINEL;
    return root;
}

/* Preorder serialization of a tree, for topology comparisons independent of
 * node numbering: an inner node emits 0xFFFFFFFF, a leaf its leaf id.
 * Topology form (nodes with left/right/leaf_id arrays, leaf when left is the
 * sentinel).  Returns the number of entries written (2n-1). */
uint64_t orc_preorder_topology(uint32_t root, const uint32_t *left, const uint32_t *right, const uint32_t *leaf_id,
                               uint64_t node_count, uint32_t *out) {
    uint32_t *stack = (uint32_t *)malloc((node_count + 1) * sizeof(uint32_t));
    uint64_t sp = 0, k = 0;
    stack[sp++] = root;
    while (sp) {
        uint32_t x = stack[--sp];
        if (left[x] == ORC_SENTINEL) { out[k++] = leaf_id[x]; continue; }
        out[k++] = ORC_SENTINEL;
        stack[sp++] = right[x];
        stack[sp++] = left[x];
    }
    free(stack);
    return k;
}

/* The same serialization of a binary layout (node_meta stride 4, children
 * tagged with bit 31 for leaves whose index is a position into leaf_ids). */
uint64_t orc_preorder_binary(uint32_t root, const uint32_t *node_meta, const uint32_t *leaf_ids, uint64_t n,
                             uint32_t *out) {
    uint32_t *stack = (uint32_t *)malloc((2 * n + 1) * sizeof(uint32_t));
    uint64_t sp = 0, k = 0;
    stack[sp++] = root;
    while (sp) {
        uint32_t x = stack[--sp];
        if (x & ORC_LEAF_FLAG) { out[k++] = leaf_ids[x & ORC_INDEX_MASK]; continue; }
        if (k >= 2 * n) break;
        out[k++] = ORC_SENTINEL;
        stack[sp++] = node_meta[4 * (uint64_t)x + 2];
        stack[sp++] = node_meta[4 * (uint64_t)x + 1];
    }
    free(stack);
    return k;
}

/* The topology as a binary layout with leaves renumbered depth-first (the
 * device builder's output form): inner topology node t becomes binary node
 * t - n, node_meta = parent, left, right, fence (last depth-first leaf
 * position), node_bounds = left then right child box, leaf_ids by position,
 * leaf_counts per node.  Returns the binary root (tagged when n == 1). */
typedef struct {
    const uint32_t *left, *right, *leaf_id, *leaf_count, *numbering;
    const float *box;
    uint64_t n;
    uint32_t *meta, *leaf_ids, *counts, *leaf_parent;
    float *nb;
    uint32_t next_pos;
} hb_state;

static uint32_t hb_ref(hb_state *s, uint32_t t, uint32_t parent, uint32_t *last_pos) {
    if (s->left[t] == ORC_SENTINEL) {
        uint32_t pos = s->next_pos++;
        s->leaf_ids[pos] = s->leaf_id[t];
        s->leaf_parent[pos] = parent;
        *last_pos = pos;
        return ORC_LEAF_FLAG | pos;
    }
    uint32_t b = s->numbering ? s->numbering[t] : t - (uint32_t)s->n;
    uint32_t *m = s->meta + 4 * (uint64_t)b;
    m[0] = parent;
    uint32_t lp = 0, rp = 0;
    m[1] = hb_ref(s, s->left[t], b, &lp);
    m[2] = hb_ref(s, s->right[t], b, &rp);
    m[3] = rp;
    memcpy(s->nb + 12 * (uint64_t)b, s->box + 6 * (uint64_t)s->left[t], 6 * sizeof(float));
    memcpy(s->nb + 12 * (uint64_t)b + 6, s->box + 6 * (uint64_t)s->right[t], 6 * sizeof(float));
    s->counts[b] = s->leaf_count[t];
    *last_pos = rp;
    return b;
}

uint32_t orc_hploc_binary(uint32_t root, const uint32_t *left, const uint32_t *right, const uint32_t *leaf_id,
                          const uint32_t *leaf_count, const float *bounds6, uint64_t n, const uint32_t *numbering,
                          uint32_t *node_meta, float *node_bounds, uint32_t *leaf_ids, uint32_t *leaf_counts,
                          uint32_t *leaf_parent) {
    hb_state s = {left, right, leaf_id, leaf_count, numbering, bounds6, n, node_meta, leaf_ids, leaf_counts,
                  leaf_parent, node_bounds, 0};
    uint32_t last = 0;
    return hb_ref(&s, root, ORC_SENTINEL, &last);
}

/* The device builder's inner-node numbering (a counter-free scheme with no
 * counterpart in bajo): every cluster is named by the lowest sorted position
 * among its leaves; a merge retires the larger of its two names and the new
 * node takes index (retired - 1).  Position 0 is never retired, so the n-1
 * inner nodes receive 0..n-2.  Inner topology nodes are created after their
 * children (indices n.. in creation order), so one ascending pass suffices.
 * sorted_ids: leaf id at each sorted position.  out: 2n-1 entries. */
void orc_hploc_lowest_leaf_numbering(const uint32_t *left, const uint32_t *right, const uint32_t *sorted_ids,
                                     uint64_t n, uint32_t *out) {
    uint32_t *low = (uint32_t *)malloc((2 * n - 1) * sizeof(uint32_t));
    for (uint64_t p = 0; p < n; ++p) low[sorted_ids[p]] = (uint32_t)p;
    for (uint64_t t = 0; t < n; ++t) out[t] = ORC_SENTINEL;
    for (uint64_t t = n; t < 2 * n - 1; ++t) {
        uint32_t a = low[left[t]], b = low[right[t]];
        low[t] = a < b ? a : b;
        out[t] = (a > b ? a : b) - 1u;
    }
    free(low);
}
