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
} orc_bvh;

orc_hit orc_trace_closest(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max, uint32_t *stats);
orc_hit orc_trace_any(const orc_bvh *b, orc_v3 o, orc_v3 d, float t_min, float t_max);

#endif
