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
} orc_bvh;

int orc_trace_wide_query(const orc_bvh *b, int any, orc_v3 o, orc_v3 d, float t_min, float t_max, orc_hit *out,
                         uint32_t *stats);

orc_hit orc_trace_closest(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max, uint32_t *stats);
orc_hit orc_trace_any(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max);

#endif
