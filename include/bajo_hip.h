/*
 * bajo_hip.h -- C ABI of the MI355X (gfx950) BVH / ray-tracing hot path.
 *
 * Each entry point replaces one function of bajo's Mojo GPU API (cited as
 * path:line into jgsimard/bajo).  All buffers are plain device pointers owned
 * by the caller unless returned through an opaque handle, which the library
 * owns until the matching *_free call.  Work is enqueued on the context's
 * stream and is asynchronous unless stated otherwise; functions that hand
 * back host data synchronize that stream first.
 *
 * Every function returns a bajo status: 0 = OK, negative = argument or HIP
 * failure (bajo_last_error() has the message), positive = a device status
 * word reported by a kernel (bajo raises these at its sync points, e.g.
 * bvh/gpu/builder/segmented_build.mojo:53-66).
 */
#ifndef BAJO_HIP_H
#define BAJO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BAJO_OK 0
#define BAJO_ERR_INVALID_ARGUMENT (-1)
#define BAJO_ERR_HIP (-2)
#define BAJO_ERR_OUT_OF_MEMORY (-3)
#define BAJO_ERR_UNSUPPORTED (-4)
#define BAJO_STATUS_STACK_OVERFLOW 1 /* a traversal needed more than 64 stack entries */
#define BAJO_STATUS_BAD_NODE 2       /* a traversal met a node index outside the hierarchy */

/* Trace modes, bvh/constants.mojo:9-13 (TRACE.CLOSEST_HIT / ANY_HIT). */
#define BAJO_TRACE_CLOSEST_HIT 0
#define BAJO_TRACE_ANY_HIT 1

/* Build methods, bvh/gpu GpuBvhBuildMethod (LBVH only in this build). */
#define BAJO_BUILD_LBVH 0

typedef struct bajo_ctx bajo_ctx;
typedef struct bajo_bvh bajo_bvh;

/* Per-stage build timings in nanoseconds measured with HIP events on the
 * context stream (no host synchronization between stages).  Mirrors
 * GpuBuildTimings, bvh/gpu/utils.mojo:27-46. */
typedef struct {
  uint64_t bounds_ns;   /* triangle + scene/centroid bounds           */
  uint64_t morton_ns;   /* Morton keys (and the fused key histogram)  */
  uint64_t sort_ns;     /* onesweep radix sort of the 64-bit keys     */
  uint64_t topology_ns; /* Karras hierarchy                           */
  uint64_t refit_ns;    /* bottom-up refit + sorted triangle packing  */
  uint64_t total_ns;
} bajo_build_timings;

/* Shape of a built BVH (host copy of its sizes and device addresses). */
typedef struct {
  uint64_t prim_count;
  uint64_t node_count;       /* internal nodes = prim_count - 1         */
  uint32_t root;             /* tagged reference (bit 31 = leaf)        */
  uint32_t reserved;
  const void* nodes;         /* 64 B per node: bounds[12] f32, meta[4] u32 (parent,left,right,fence) */
  const void* triangles;     /* 48 B per sorted leaf: v0.xyz,prim | e1.xyz,0 | e2.xyz,0 */
  const void* sorted_keys;   /* u64[prim_count]: (morton<<32)|prim       */
  const void* leaf_parent;   /* u32[prim_count]                          */
  const void* leaf_counts;   /* u32[node_count]                          */
  const void* scene_bounds;  /* f32[12]: root box, centroid box          */
} bajo_bvh_info;

/* ---- context ------------------------------------------------------------
 * Replaces the MAX DeviceContext the Mojo API threads through every call
 * (e.g. bvh/gpu/triangle_bvh.mojo:617 `ctx: DeviceContext`).  `stream` may be
 * NULL for the device's null stream. */
int32_t bajo_ctx_create(int32_t device, void* stream, bajo_ctx** out);
int32_t bajo_ctx_destroy(bajo_ctx* ctx);
int32_t bajo_ctx_synchronize(bajo_ctx* ctx);
const char* bajo_last_error(void);
const char* bajo_status_string(int32_t status);
int32_t bajo_device_name(int32_t device, char* out, int32_t cap);

/* ---- sort: bajo/sort/gpu ------------------------------------------------
 * Stable ascending LSD radix sort (8-bit digits) of device buffers, result
 * in place.  `workspace` may be NULL (the context's scratch arena is used);
 * otherwise it must hold bajo_sort_workspace_bytes(n, key_bytes, has_values).
 */
uint64_t bajo_sort_workspace_bytes(uint64_t n, int32_t key_bytes, int32_t has_values);

/* device_radix_sort_pairs[u64,u32], sort/gpu/radix_sort.mojo:510-597. */
int32_t bajo_device_radix_sort_pairs_u64_u32(bajo_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n,
                                             void* workspace, uint64_t workspace_bytes);
/* device_radix_sort_pairs[u32,u32], sort/gpu/radix_sort.mojo:600-609. */
int32_t bajo_device_radix_sort_pairs_u32_u32(bajo_ctx* ctx, uint32_t* keys, uint32_t* values, uint64_t n,
                                             void* workspace, uint64_t workspace_bytes);
/* device_radix_sort_keys[u64 / u32], sort/gpu/radix_sort.mojo:397-471. */
int32_t bajo_device_radix_sort_keys_u64(bajo_ctx* ctx, uint64_t* keys, uint64_t n, void* workspace,
                                        uint64_t workspace_bytes);
int32_t bajo_device_radix_sort_keys_u32(bajo_ctx* ctx, uint32_t* keys, uint64_t n, void* workspace,
                                        uint64_t workspace_bytes);
/* onesweep_radix_sort_keys[u64], sort/gpu/onesweep.mojo:362-451, with the
 * sorted bit range [begin_bit, end_bit) (multiples of 8) made explicit so
 * callers whose low bits are already ordered skip those passes. */
int32_t bajo_onesweep_radix_sort_keys_u64(bajo_ctx* ctx, uint64_t* keys, uint64_t n, int32_t begin_bit,
                                          int32_t end_bit, void* workspace, uint64_t workspace_bytes);
/* onesweep_radix_sort_pairs[u64,u32], sort/gpu/onesweep.mojo:454-572. */
int32_t bajo_onesweep_radix_sort_pairs_u64_u32(bajo_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n,
                                               int32_t begin_bit, int32_t end_bit, void* workspace,
                                               uint64_t workspace_bytes);

/* ---- BVH build: bajo/bvh/gpu -------------------------------------------
 * build_triangle_bvh[frame, W, L, LBVH](ctx, vertices f32[9N]),
 * bvh/gpu/triangle_bvh.mojo:617-630 (and build_triangle_bvh_measured,
 * :633-646, when `timings` is non-NULL).  `vertices` is a device buffer of
 * n triangles (v0, v1, v2 as 9 consecutive floats).  The returned handle
 * owns the hierarchy (a binary LBVH whose nodes hold both child boxes). */
int32_t bajo_build_triangle_bvh(bajo_ctx* ctx, const float* vertices, uint64_t n, int32_t method,
                                bajo_build_timings* timings, bajo_bvh** out);
int32_t bajo_bvh_free(bajo_bvh* bvh);
int32_t bajo_bvh_get_info(const bajo_bvh* bvh, bajo_bvh_info* out);
/* Copies the binary LBVH into bajo's GpuBinaryBoundsBvh layout
 * (bvh/gpu/builder/binary_layout.mojo:303-330): any pointer may be NULL.
 * sorted_keys u64[n], leaf_ids u32[n], node_meta u32[4(n-1)] (parent, left,
 * right, fence), node_bounds f32[12(n-1)], leaf_parent u32[n],
 * leaf_counts u32[n-1], scene_bounds f32[12].  Host pointers; synchronizes. */
int32_t bajo_bvh_copy_binary(bajo_ctx* ctx, const bajo_bvh* bvh, uint64_t* sorted_keys, uint32_t* leaf_ids,
                             uint32_t* node_meta, float* node_bounds, uint32_t* leaf_parent,
                             uint32_t* leaf_counts, float* scene_bounds);

/* ---- traversal ----------------------------------------------------------
 * GpuTriangleBvh.launch_rays[mode], bvh/gpu/triangle_bvh.mojo:587-614:
 * rays f32[8R] field-major (plane k = origin.xyz, t_min, dir.xyz, t_max at
 * k*R, bvh/gpu/ray_launch.mojo:28-46), hits f32[8R] AoS (u, v, prim, inst,
 * normal.xyz, t; bvh/types.mojo:16-115).  `stats` may be NULL, else u32[5R]
 * (node visits, internal child hits, unused, leaf tests, max stack depth)
 * like GpuTraversalStats, bvh/gpu/trace.mojo:62-93. */
int32_t bajo_trace_rays(bajo_ctx* ctx, const bajo_bvh* bvh, int32_t mode, const float* rays, float* hits,
                        uint64_t ray_count, uint32_t* stats);
/* GpuTriangleBvh.launch_camera, bvh/gpu/triangle_bvh.mojo:520-548: camera
 * f32[20 * views] (bvh/camera.mojo:19-30), one raster ray per pixel. */
int32_t bajo_trace_camera(bajo_ctx* ctx, const bajo_bvh* bvh, const float* camera, float* hits,
                          uint64_t ray_count, uint32_t width, uint32_t height, uint32_t* stats);
/* Camera(origin, target, up, fov_scale, focus_dist), bvh/camera.mojo:86-106,
 * flattened to 20 floats (bvh/camera.mojo:210-231).  Host-side helper. */
int32_t bajo_camera_make(const float* origin3, const float* target3, const float* up3, float fov_scale,
                         float focus_dist, float* out20);

/* ---- synthetic scenes (host) -------------------------------------------
 * Seeded PCG32 (O'Neill 2014) generators for the benchmark workloads.  All
 * write n triangles as f32[9n] into a host buffer. */
int32_t bajo_scene_soup(uint64_t initstate, uint64_t initseq, uint64_t n, float edge_min, float edge_max,
                        float* out_vertices);
int32_t bajo_scene_icosphere(uint32_t frequency, uint64_t n, uint64_t noise_seed, float amplitude,
                             float noise_frequency, float* out_vertices);

#ifdef __cplusplus
}
#endif
#endif
