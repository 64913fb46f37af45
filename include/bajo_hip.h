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
/* Context status words (raised by kernels into the context's device status
 * word, reported at its sync points, like bajo's HPLOC_STATUS / collapse
 * TIMEOUT words, bvh/gpu/builder/hploc_layout.mojo:9-11,
 * wide_collapse.mojo:34-38,236-254): */
#define BAJO_STATUS_SORT_LOOKBACK 4  /* a sort tile gave up waiting for its predecessor (spin cap) */
#define BAJO_STATUS_REFIT_CORRUPT 5  /* the refit climb met an inconsistent hierarchy or passed the depth bound */
#define BAJO_STATUS_HPLOC_FAILED 6   /* H-PLOC made no progress, met a corrupt cluster list or ended without one root */
#define BAJO_STATUS_COLLAPSE_FAILED 7 /* the wide collapse ran out of nodes, met a bad root or passed its level bound */
/* WAVE_STATUS words of the renderer, rt/wavefront_contract.mojo:25-29, offset by 10 */
#define BAJO_STATUS_WAVE_BASE 10
#define BAJO_STATUS_PATH_OVERFLOW 11
#define BAJO_STATUS_SHADOW_OVERFLOW 13

/* Trace modes, bvh/constants.mojo:9-13 (TRACE.CLOSEST_HIT / ANY_HIT). */
#define BAJO_TRACE_CLOSEST_HIT 0
#define BAJO_TRACE_ANY_HIT 1

/* Build methods, bvh/gpu GpuBvhBuildMethod (bvh/gpu/builder/__init__.mojo). */
#define BAJO_BUILD_LBVH 0
/* H-PLOC (bvh/gpu/builder/hploc_multi_wave.mojo:187-777): LBVH-guided
 * bottom-up clustering, search radius 8, merging threshold 16 (or 32,
 * bajo_build_options.merging_threshold), with the topology of bajo's CPU
 * definition (bvh/cpu/builder/hploc.mojo:855-917) at that threshold.
 * Leaves are then renumbered depth-first, so in the exported binary layout
 * every subtree covers a contiguous range of positions and `fence` is the
 * last one (bajo leaves FENCE unset for H-PLOC).  The root stays on the
 * device (traversals read it there; bajo_bvh_get_info reads it back) and a
 * failed clustering raises BAJO_STATUS_HPLOC_FAILED at the next sync point. */
#define BAJO_BUILD_HPLOC 1

/* Node widths (bajo's W parameter, bvh/gpu/triangle_bvh.mojo:617): 2 keeps
 * the binary LBVH; 8 collapses it into 8-wide nodes with 8-bit quantized
 * child boxes (bajo's wide collapse + compressed BVH8 layout,
 * bvh/gpu/builder/wide_collapse.mojo:148-643,
 * bvh/gpu/compressed_bounds_bvh.mojo:84-369). */
#define BAJO_WIDTH_BINARY 2
#define BAJO_WIDTH_WIDE8 8
/* 4-wide nodes (bajo's render_gpu default W = 4, bvh/gpu/builder/
 * wide_collapse.mojo:148-643): the same collapse opening up to 4 children,
 * stored in the same quantized node record with 4 live slots. */
#define BAJO_WIDTH_WIDE4 4

/* Build options (bajo's compile-time W, L and method parameters). */
typedef struct {
  int32_t method;     /* BAJO_BUILD_LBVH or BAJO_BUILD_HPLOC               */
  int32_t node_width; /* BAJO_WIDTH_BINARY, BAJO_WIDTH_WIDE4 or _WIDE8      */
  int32_t leaf_size;  /* width 8: most triangles per leaf child, 1..8; 0 = 4 */
  int32_t merging_threshold; /* H-PLOC: 0 or 16 = bajo's CPU definition (bvh/cpu/builder/
                                hploc.mojo:855-917), 32 = its GPU build on wave64 (WARP_SIZE / 2,
                                bvh/gpu/builder/hploc_layout.mojo:6); LBVH ignores it */
} bajo_build_options;

typedef struct bajo_ctx bajo_ctx;
typedef struct bajo_bvh bajo_bvh;

/* Per-stage build timings in nanoseconds measured with HIP events on the
 * context stream (no host synchronization between stages).  Mirrors
 * GpuBuildTimings, bvh/gpu/utils.mojo:27-46. */
typedef struct {
  uint64_t bounds_ns;   /* triangle + scene/centroid bounds           */
  uint64_t morton_ns;   /* Morton keys (and the fused key histogram)  */
  uint64_t sort_ns;     /* onesweep radix sort of the 64-bit keys     */
  uint64_t topology_ns; /* Karras hierarchy (LBVH; H-PLOC: 0)         */
  uint64_t refit_ns;    /* LBVH: bottom-up refit + triangle packing;
                           H-PLOC: guide, clustering, leaf reorder      */
  uint64_t collapse_ns; /* binary -> 8-wide collapse (width 8 only)   */
  uint64_t total_ns;
} bajo_build_timings;

/* Shape of a built BVH (host copy of its sizes and device addresses).  For a
 * segment view of a BLAS set the addresses point at the segment's first
 * records inside the set's shared arrays, whose references are numbered
 * across the whole set; `root` is the segment's own (0 or leaf 0), as in the
 * export of bajo_bvh_copy_binary. */
typedef struct {
  uint64_t prim_count;
  uint64_t node_count;       /* internal nodes = prim_count - 1         */
  uint32_t root;             /* tagged reference (bit 31 = leaf)        */
  uint32_t node_width;       /* 2 or 8                                  */
  const void* nodes;         /* 64 B per node: bounds[12] f32, meta[4] u32 (parent,left,right,fence) */
  const void* triangles;     /* 48 B per sorted leaf: v0.xyz,prim | e1.xyz,0 | e2.xyz,0 */
  const void* sorted_keys;   /* u64[prim_count]: (morton<<32)|prim       */
  const void* leaf_parent;   /* u32[prim_count]                          */
  const void* leaf_counts;   /* u32[node_count]                          */
  const void* scene_bounds;  /* f32[12]: root box, centroid box          */
  uint64_t wide_node_count;  /* width 8: 96-byte wide nodes, root = node 0 */
  const void* wide_nodes;    /* width 8: origin f32[3], exponents u8[3], child count u8,
                                lo_x,lo_y,lo_z,hi_x,hi_y,hi_z u8[8] each, child u32[8] */
} bajo_bvh_info;

/* ---- context ------------------------------------------------------------
 * Replaces the MAX DeviceContext the Mojo API threads through every call
 * (e.g. bvh/gpu/triangle_bvh.mojo:617 `ctx: DeviceContext`).  `stream` may be
 * NULL for the device's null stream. */
int32_t bajo_ctx_create(int32_t device, void* stream, bajo_ctx** out);
int32_t bajo_ctx_destroy(bajo_ctx* ctx);
/* Synchronizes the context stream; returns the context status word (a
 * positive BAJO_STATUS_*) if a kernel raised one since the last check, and
 * clears it. */
int32_t bajo_ctx_synchronize(bajo_ctx* ctx);
/* Synchronizes and reads the context status word into *status (0 = none
 * raised), then clears it.  Returns BAJO_OK unless the read itself fails. */
int32_t bajo_ctx_status(bajo_ctx* ctx, uint32_t* status);
/* Spin cap of the sort's decoupled look-back wait (0 = the default 2^24):
 * a tile that waits longer gives up and raises BAJO_STATUS_SORT_LOOKBACK
 * instead of waiting forever on a corrupted look-back table.  bajo's wide
 * collapse has the same kind of fail-safe (a 10M-spin cap,
 * bvh/gpu/builder/wide_collapse.mojo:236-254); a tiny cap lets tests force
 * the status path from the host. */
int32_t bajo_ctx_set_spin_limit(bajo_ctx* ctx, uint32_t spin_limit);
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
/* No counterpart in bajo.  The same sort as
 * bajo_onesweep_radix_sort_pairs_u64_u32 for callers that cannot allocate
 * device memory (code under HIP graph capture, tight submission loops,
 * harnesses that forbid hipMalloc): its workspace is a fixed 16 MiB device
 * array inside the library (one per GPU, held until the process ends), so
 * the call allocates nothing.  n is at most 2^20 pairs (larger n returns
 * BAJO_ERR_INVALID_ARGUMENT).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other streams, other threads) race. */
int32_t bajo_onesweep_radix_sort_pairs_u64_u32_noalloc(bajo_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n,
                                                       int32_t begin_bit, int32_t end_bit);

/* ---- BVH build: bajo/bvh/gpu -------------------------------------------
 * build_triangle_bvh[frame, W, L, LBVH](ctx, vertices f32[9N]),
 * bvh/gpu/triangle_bvh.mojo:617-630 (and build_triangle_bvh_measured,
 * :633-646, when `timings` is non-NULL).  `vertices` is a device buffer of
 * n triangles (v0, v1, v2 as 9 consecutive floats).  The returned handle
 * owns the hierarchy (a binary LBVH whose nodes hold both child boxes). */
int32_t bajo_build_triangle_bvh(bajo_ctx* ctx, const float* vertices, uint64_t n, int32_t method,
                                bajo_build_timings* timings, bajo_bvh** out);
/* build_triangle_bvh[frame, W, L, method] with W and L chosen at run time.
 * With node_width 8 the LBVH is collapsed top-down into 8-wide nodes: a node
 * opens the child of largest surface area until it has 8 children, subtrees
 * of at most leaf_size triangles become leaf children (ranges of the sorted
 * triangle records), child boxes are stored on a per-axis power-of-two grid
 * and padded by one grid step.  Every build is enqueue-only: nothing is read
 * back and the host never waits (without `timings`), so a rebuild can be
 * captured into a HIP graph once its buffers are allocated.  The collapse
 * runs a depth-bounded number of level launches (a launch past the last
 * level does nothing) and a single-block kernel for any deeper levels; the
 * wide node count is read back at the first bajo_bvh_get_info.  Failures
 * raise BAJO_STATUS_* into the context status word.  `options` NULL =
 * binary LBVH.  n < 2^28 for widths 4 and 8. */
int32_t bajo_build_triangle_bvh_ex(bajo_ctx* ctx, const float* vertices, uint64_t n,
                                   const bajo_build_options* options, bajo_build_timings* timings,
                                   bajo_bvh** out);
/* GpuBinaryBoundsBvh(ctx, leaf_bounds, payloads, ws) + build_binary_bvh[LBVH |
 * HPLOC] (bvh/gpu/builder/binary_layout.mojo:303-448,
 * bvh/gpu/builder/binary_builder.mojo:21-45) for one segment: the same
 * builders over caller boxes f32[6N] (lo.xyz, hi.xyz per leaf; centroid =
 * (lo + hi) / 2) instead of triangles.  bajo_bvh_copy_binary exports the
 * layout; leaf_ids index the input boxes (bajo's payloads are looked up
 * through them).  Options as for triangles (node width 8 collapses the
 * boxes hierarchy too).  Ray launches reject a boxes hierarchy. */
int32_t bajo_build_binary_bounds(bajo_ctx* ctx, const float* leaf_bounds, uint64_t n,
                                 const bajo_build_options* options, bajo_build_timings* timings, bajo_bvh** out);
/* build_sphere_bvh (bvh/gpu/sphere_bvh.mojo:320-371): spheres f32[4N]
 * (center.xyz, radius; 16-byte aligned) -> the box builder over Sphere.bounds
 * (LBVH or H-PLOC, node width 2 or 8), then each sorted leaf's sphere packed
 * into its leaf record.  Ray launches trace it with bajo's ray-sphere test
 * (core/intersect.mojo:309-350; hit u = v = 0, normal = normalize(p - c));
 * it enters a TLAS like a triangle BVH.  A non-NULL *out (a sphere BVH) is
 * rebuilt in place.  Uses the context scratch (24 B per sphere) while
 * building. */
int32_t bajo_build_sphere_bvh(bajo_ctx* ctx, const float* spheres, uint64_t n, const bajo_build_options* options,
                              bajo_build_timings* timings, bajo_bvh** out);
int32_t bajo_bvh_free(bajo_bvh* bvh);
int32_t bajo_bvh_get_info(const bajo_bvh* bvh, bajo_bvh_info* out);
/* Copies the binary LBVH into bajo's GpuBinaryBoundsBvh layout
 * (bvh/gpu/builder/binary_layout.mojo:303-330): any pointer may be NULL.
 * sorted_keys u64[n], leaf_ids u32[n], node_meta u32[4(n-1)] (parent, left,
 * right, fence), node_bounds f32[12(n-1)], leaf_parent u32[n],
 * leaf_counts u32[n-1], scene_bounds f32[12].  Host pointers; synchronizes.
 * A segment view exports its segment renumbered from 0, equal to the
 * standalone build of its mesh. */
int32_t bajo_bvh_copy_binary(bajo_ctx* ctx, const bajo_bvh* bvh, uint64_t* sorted_keys, uint32_t* leaf_ids,
                             uint32_t* node_meta, float* node_bounds, uint32_t* leaf_parent,
                             uint32_t* leaf_counts, float* scene_bounds);
/* Copies the 8-wide nodes (96 bytes each, layout in bajo_bvh_info) into a
 * host buffer of `capacity` nodes; *count receives the node count.  Node
 * numbering depends on the order in which waves reserve nodes, so only the
 * hierarchy it describes is reproducible (as for bajo's wide collapse).  A
 * segment view copies the nodes reachable from its root in depth-first order,
 * leaf ranges relative to its first triangle. */
int32_t bajo_bvh_copy_wide(bajo_ctx* ctx, const bajo_bvh* bvh, void* nodes, uint64_t capacity, uint64_t* count);

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
/* Ambient-occlusion rays of a closest-hit launch: `samples` rays per
 * primary ray, each bajo's _make_ao_ray (rt/gpu/path_shading.mojo:650-665):
 * origin at the hit, direction uniform on the hemisphere of the normal
 * oriented against the ray (Philox stream (seed, path r*samples+k, stage 1)),
 * t in [0.001, 4.0].  rays / hits as bajo_trace_rays; ao_rays f32[8 * R *
 * samples] field-major, ready for bajo_trace_rays(ANY_HIT).  Rays of primary
 * rays that hit nothing get t_max = -1.  bajo builds this ray inside its AO
 * render kernel; here it is a launch of its own so the AO rays of any ray set
 * (the H-PLOC + 8-wide random-ray workload) can be traced. */
int32_t bajo_ao_rays(bajo_ctx* ctx, const float* rays, const float* hits, uint64_t ray_count, uint32_t samples,
                     uint64_t seed, float* ao_rays);
/* Camera(origin, target, up, fov_scale, focus_dist), bvh/camera.mojo:86-106,
 * flattened to 20 floats (bvh/camera.mojo:210-231).  Host-side helper. */
int32_t bajo_camera_make(const float* origin3, const float* target3, const float* up3, float fov_scale,
                         float focus_dist, float* out20);

/* ---- BLAS sets: bajo/bvh/gpu/triangle_bvh.mojo:366-491, blas_storage.mojo ----
 * build_triangle_blas_set (triangle_bvh.mojo:469-491): the meshes of a set
 * built in one segmented pass (every stage one launch over all segments).
 * vertices f32[9N] hold the meshes back to back; segment_offsets (host)
 * u64[S + 1] hold each mesh's first triangle (offsets[0] = 0, increasing,
 * offsets[S] = N).  Each segment's hierarchy equals the standalone build of
 * its mesh; hit prim ids are indices inside the mesh.  Method LBVH or
 * H-PLOC (bajo's default for sets), node width 2 or 8.  bajo_blas_set_get returns segment k as a bajo_bvh view
 * (bajo's BlasDesc, blas_desc.mojo) that traces, exports and enters a TLAS
 * like any BVH; it shares the set's storage, cannot be rebuilt or freed on
 * its own, and lives until bajo_blas_set_free.  A non-NULL *out is rebuilt
 * in place (same segment count; its views stay valid and see the new build). */
typedef struct bajo_blas_set bajo_blas_set;
int32_t bajo_build_triangle_blas_set(bajo_ctx* ctx, const float* vertices, const uint64_t* segment_offsets,
                                     uint32_t segment_count, const bajo_build_options* options,
                                     bajo_build_timings* timings, bajo_blas_set** out);
int32_t bajo_blas_set_count(const bajo_blas_set* set, uint32_t* count);
int32_t bajo_blas_set_get(const bajo_blas_set* set, uint32_t index, bajo_bvh** view);
int32_t bajo_blas_set_free(bajo_blas_set* set);
/* Batched camera launch over a BLAS set (the batched many-environment trace:
 * one launch for every segment instead of a per-segment launch loop; bajo
 * has no counterpart, its closest code is launch_camera per BLAS,
 * bvh/gpu/triangle_bvh.mojo:520-548): camera f32[20 S], view e traced against
 * segment e, width x height raster rays per view, hits f32[8 S w h] view-major
 * (prim = index inside the segment's mesh); `stats` NULL or u32[5 S w h]. */
int32_t bajo_blas_set_trace_camera(bajo_ctx* ctx, const bajo_blas_set* set, const float* cameras, float* hits,
                                   uint32_t width, uint32_t height, uint32_t* stats);

/* ---- instances: bajo/bvh/gpu/tlas.mojo -----------------------------------
 * build_triangle_tlas (bvh/gpu/tlas.mojo:618-645): a top-level hierarchy
 * over instances of triangle BVHs.  transforms f32[12 I] are row-major 3x4
 * affines (m00 m01 m02 tx / m10 m11 m12 ty / m20 m21 m22 tz) from BLAS space
 * to world space (bajo's Instance and Affine3, bvh/types.mojo:153-190,
 * core/transform.mojo); blas_indices u32[I] select from blases[].  An
 * instance's world box is its BLAS root box through the transform
 * (AABB.apply_transform, core/aabb.mojo:101-137); the top level is
 * bajo_build_binary_bounds over those boxes (options: method; node width 2).
 * Host arrays in, device state out; the BLASes must outlive the TLAS. */
typedef struct bajo_tlas bajo_tlas;
int32_t bajo_tlas_create(bajo_ctx* ctx, const bajo_bvh* const* blases, uint32_t blas_count, const float* transforms,
                         const uint32_t* blas_indices, uint32_t instance_count, const bajo_build_options* options,
                         bajo_tlas** out);
/* Rays through the instances (_trace_tlas_ray, bvh/gpu/tlas.mojo:176-233):
 * at each instance leaf the ray is mapped into BLAS space by the inverse
 * transform (t unchanged) and the BLAS, binary or 8-wide, is traversed; a
 * closest hit reports inst = the instance index and the normal through the
 * inverse transpose; an any hit is bajo's shadow hit with inst set.  rays /
 * hits as bajo_trace_rays. */
int32_t bajo_tlas_trace_rays(bajo_ctx* ctx, const bajo_tlas* tlas, int32_t mode, const float* rays, float* hits,
                             uint64_t ray_count);
/* GpuTriangleTlas.launch_camera (bvh/gpu/tlas.mojo:499-548): pinhole rays as
 * bajo_trace_camera, closest hit. */
int32_t bajo_tlas_trace_camera(bajo_ctx* ctx, const bajo_tlas* tlas, const float* camera, float* hits,
                               uint64_t ray_count, uint32_t width, uint32_t height);
/* The top-level boxes hierarchy (owned by the TLAS; bajo_bvh_copy_binary
 * exports it, its leaf_ids are instance indices) and the instance world
 * boxes f32[6 I] on the host. */
int32_t bajo_tlas_get_top(const bajo_tlas* tlas, bajo_bvh** top);
int32_t bajo_tlas_copy_instance_bounds(const bajo_tlas* tlas, float* out6);
/* Traversal status word of the last launch (BAJO_STATUS_*), synchronizing. */
int32_t bajo_tlas_status(bajo_ctx* ctx, const bajo_tlas* tlas, uint32_t* status);
int32_t bajo_tlas_free(bajo_tlas* tlas);

/* ---- wavefront renderer: bajo/rt/gpu ------------------------------------
 * RENDER enum, rt/types.mojo:51-57. */
#define BAJO_RENDER_PATH 0
#define BAJO_RENDER_NORMALS 1
#define BAJO_RENDER_AO 2
#define BAJO_RENDER_NEE 3
#define BAJO_RENDER_MIS 4
/* MAT enum (rt/types.mojo:24-29); a surface id is kind << 28 | index. */
#define BAJO_MAT_LAMBERTIAN 0
#define BAJO_MAT_METAL 1
#define BAJO_MAT_DIELECTRIC 2
#define BAJO_MAT_EMISSIVE 3
/* A GGX microfacet conductor (no bajo counterpart; bajo's MAT stops at
 * EMISSIVE): Trowbridge-Reitz distribution of width alpha = roughness in
 * (0, 1], Smith height-correlated masking-shadowing, Schlick Fresnel with
 * F0 = the specular color; sampled from the visible normals (Heitz 2018).
 * It takes direct light (NEE / MIS) like Lambertian surfaces. */
#define BAJO_MAT_GGX 4

typedef struct bajo_scene bajo_scene;

/* RenderSettings (rt/types.mojo:324-357) plus the launch schedule.  None of
 * the schedule knobs changes an image (every path draws its own random
 * streams and owns its sample slot; the tail keeps the wavefront's order):
 *   path_capacity: paths in flight per chunk, the path arena of
 *     GpuRtRenderTarget (rt/gpu/resources.mojo:56-99); 0 = the whole frame;
 *   extend_waves: the occupancy the extension kernel of bounces >= 1 is
 *     compiled for (waves per SIMD: 4, 6 or 8; 0 = 6) with its LDS traversal
 *     stack sized to match (32, 24 or 16 entries), which trades VGPRs and
 *     on-chip stack per lane against resident waves; used by PATH, NEE and
 *     MIS over 4- and 8-wide triangle and batch scenes (others ignore it);
 *   tail_paths: live paths at or below which the tail kernel finishes the
 *     chunk's remaining bounces in one launch (0 = the tail grid's resident
 *     threads);
 *   shadow_walk: the 4- and 8-wide any-hit walk of the shadow-ray kernel:
 *     1 = one node per step with one leaf record loaded ahead, 2 = two nodes
 *     per step (the stack top with the current node), 3 = one node with leaf
 *     records loaded four at a time, 4 = both; 0 = the library's default;
 *   sparse_shadow_rays: shadow rays per bounce at or below which each ray
 *     gets a whole wave (up to eight nodes and all their leaf records per
 *     step; 4- and 8-wide triangle and batch scenes); 0 = 8192, UINT32_MAX =
 *     never;
 *   shadow_group: lanes per shadow ray above that threshold: 1 = one lane
 *     per ray (shadow_walk applies), 8, 16 or 32 = that many lanes per ray
 *     (one node per 8 lanes per step); 0 = 16 lanes up to 2^20 rays and one
 *     lane per ray beyond;
 *   tail_group: lanes per path in the tail kernel (4- and 8-wide triangle
 *     and batch scenes): 1 = one lane per path, 16, 32 or 64 = that many
 *     (extension rays by a group walk equal to the one-lane walk, shadow
 *     rays by group any hit; one node per 8 lanes per step); 0 = 32 up to
 *     2^15 live paths, one beyond. */
typedef struct {
  uint32_t image_width, image_height, samples_per_pixel, max_depth;
  uint64_t rng_seed;
  int32_t algorithm;
  uint32_t path_capacity;
  uint32_t extend_waves;
  uint32_t tail_paths;
  uint32_t shadow_walk;
  uint32_t sparse_shadow_rays;
  uint32_t shadow_group;
  uint32_t tail_group;
} bajo_render_settings;

/* Rays traced by one render (read back from device counters). */
typedef struct {
  uint64_t extension_rays;      /* closest-hit rays, all bounces        */
  uint64_t shadow_rays;         /* any-hit rays (NEE / AO)              */
  uint64_t incoherent_rays;     /* extension + shadow rays of bounce >= 1 */
  uint64_t rays_per_bounce[16]; /* extension rays per bounce            */
  uint32_t status;              /* WAVE_STATUS (0 = ok)                 */
  uint32_t traversal_status;    /* BAJO_STATUS_* of the traversals      */
  uint64_t extension_node_visits, extension_leaf_tests; /* traversal work */
  uint64_t shadow_node_visits, shadow_leaf_tests;
  float total_ms;               /* HIP events on the context stream      */
  float extend_ms;              /* sum over extend launches              */
  float shadow_ms;              /* sum over shadow launches              */
  uint32_t extend_launches, shadow_launches;
  float extend_bounce_ms[16];   /* extend launch time per bounce (first chunk) */
  float shadow_bounce_ms[16];   /* shadow launch time per bounce (first chunk) */
  uint32_t max_extension_visits[16]; /* most nodes one extension ray visited, per bounce */
  uint32_t max_shadow_visits[16];    /* most nodes one shadow ray visited, per bounce    */
} bajo_render_stats;

/* SceneData for triangle scenes (rt/types.mojo:382-526): host arrays of
 * vertices f32[9N], per-triangle surface ids u32[N], Lambertian albedos
 * f32[3L] and emissive radiances f32[3E].  Uploads them, builds the LBVH
 * and the Walker alias light table (rt/types.mojo:207-247, weight = area x
 * mean radiance). */
int32_t bajo_scene_create(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                          const float* lambertian, uint32_t n_lambertian, const float* emissive,
                          uint32_t n_emissive, bajo_scene** out);
/* bajo_scene_create with explicit build options (NULL = binary LBVH); the
 * renderer traverses whichever layout the options select. */
int32_t bajo_scene_create_ex(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                             const float* lambertian, uint32_t n_lambertian, const float* emissive,
                             uint32_t n_emissive, const bajo_build_options* options, bajo_scene** out);
/* The material arrays of SceneData (rt/types.mojo:382-526), flattened as
 * bajo's GPU shading view does (rt/gpu/path_shading.mojo:87-120): Lambertian
 * albedo f32[3], metal albedo + fuzz f32[4], dielectric refraction index
 * f32[1], emissive radiance f32[3] per material, and this library's GGX
 * conductor f32[4] (BAJO_MAT_GGX).  Surface ids select them by kind (MAT: 0
 * Lambertian, 1 metal, 2 dielectric, 3 emissive, 4 GGX). */
typedef struct bajo_materials {
  const float* lambertian;
  uint32_t n_lambertian;
  const float* metal;
  uint32_t n_metal;
  const float* dielectric;
  uint32_t n_dielectric;
  const float* emissive;
  uint32_t n_emissive;
  const float* ggx; /* f32[4] per GGX material: F0.rgb, roughness; may be NULL with n_ggx 0 */
  uint32_t n_ggx;
} bajo_materials;
/* bajo_scene_create_ex with every material kind: metal (bajo's Phong-lobe
 * metal, mirror at fuzz <= 1e-4) and dielectric (Schlick reflection,
 * refraction) surfaces are shaded as in rt/gpu/path_shading.mojo:669-783
 * and take direct light like Lambertian ones where the BSDF has a density
 * (rough metal). */
int32_t bajo_scene_create_materials(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                                    const bajo_materials* materials, const bajo_build_options* options,
                                    bajo_scene** out);
/* Rebuilds the scene's BVH from its device-resident vertices. */
/* A mixed scene (GpuRtMixedScene / SceneData with spheres, rt/gpu/mixed_path.mojo,
 * rt/types.mojo:382-530): triangles as above (n may be 0) plus spheres
 * f32[4S] (center, signed radius: a negative radius turns the normal
 * inward) with one surface id each.  Emissive spheres are lights (weight
 * 4 pi r^2 * mean radiance, after the triangle lights).  Closest hits test
 * the spheres first, then the triangles (bajo's scene trace order).  The
 * sphere BVH is the LBVH at the scene's node width and leaf size. */
int32_t bajo_scene_create_mixed(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                                const float* spheres, uint64_t n_spheres, const uint32_t* sphere_surfaces,
                                const bajo_materials* materials, const bajo_build_options* options,
                                bajo_scene** out);
/* A batch of environments (the batched many-environment render; bajo has no
 * counterpart, its closest code is the segmented BLAS build,
 * bvh/gpu/builder/segmented_build.mojo:89-162, and render_gpu per scene,
 * rt/gpu/render.mojo:322-420): environment e is the triangles
 * [offsets[e], offsets[e+1]) of `vertices` (host f32[9N], surfaces u32[N]
 * sharing `materials`), with its own lights (its emissive triangles, its own
 * alias table).  All environments are built in one segmented pass
 * (options: method LBVH or H-PLOC, width 2 or 8).  bajo_render of a batch renders
 * every environment through the same camera and settings into pixels
 * f32[E h w 3] (environment-major) in one wavefront; environment e's image
 * equals a render of its scene alone with seed rng_seed + e.  A render
 * target renders a batch only if made by bajo_render_target_create_envs with
 * the same environment count. */
int32_t bajo_scene_create_batch(bajo_ctx* ctx, const float* vertices, const uint64_t* offsets, uint32_t env_count,
                                const uint32_t* surfaces, const bajo_materials* materials,
                                const bajo_build_options* options, bajo_scene** out);
int32_t bajo_scene_rebuild(bajo_ctx* ctx, bajo_scene* scene, bajo_build_timings* timings);
int32_t bajo_scene_free(bajo_scene* scene);
int32_t bajo_scene_get_bvh(const bajo_scene* scene, bajo_bvh** out);
/* Host copy of the light table: fields f32[14 * count] (p0, p1, p2, radius,
 * emission, alias probability), kinds u32[count] (alias << 4 | prim kind). */
int32_t bajo_scene_light_table(const bajo_scene* scene, float* fields, uint32_t* kinds, uint32_t* count,
                               float* total_weight);
/* render_gpu[ALGORITHM] (rt/gpu/render.mojo:322-420) into device pixels
 * f32[3 * width * height]; `camera` is a host f32[20] record.  Enqueued on
 * the context stream; with `stats` it synchronizes and returns the ray
 * counts and any queue status (bajo raises those in download_gpu_pixels,
 * rt/gpu/resources.mojo:268-285). */
int32_t bajo_render(bajo_ctx* ctx, bajo_scene* scene, const bajo_render_settings* settings, const float* camera,
                    float* pixels, bajo_render_stats* stats);
/* Persistent render target (GpuRtRenderTarget, update_gpu_camera and the
 * enqueue-only submission of rt/gpu/resources.mojo:38-155 and
 * rt/gpu/render.mojo:61-235): the target keeps its path queues, sample
 * planes and the frame's camera in device memory across frames and renders
 * into the caller's pixel buffer f32[3 w h] (which must outlive it).
 * `settings` are fixed at creation.  bajo_render_target_set_camera writes the
 * camera on the stream (a kernel, so it orders with renders and replays).
 * bajo_render_enqueue enqueues one render of `scene` and returns without
 * synchronizing.  bajo_render_target_capture records that enqueue into a
 * HIP graph (allocations happen first; the graph holds kernel nodes only) and
 * bajo_render_target_replay launches it on the context stream: a replay
 * reads the camera buffer when it runs, so camera updates between replays
 * take effect; a new scene, scene rebuild to a larger size or new settings
 * need a new capture.  Targets render concurrently with each other. */
typedef struct bajo_render_target bajo_render_target;
int32_t bajo_render_target_create(bajo_ctx* ctx, const bajo_render_settings* settings, const float* camera,
                                  float* pixels, bajo_render_target** out);
int32_t bajo_render_target_set_camera(bajo_ctx* ctx, bajo_render_target* target, const float* camera);
/* bajo_render with one camera per environment of a batch: cameras f32[20 E]
 * (camera_count = E; 1 = one camera for all, as bajo_render).  Environment
 * e's image equals a render of its scene alone through camera e with seed
 * rng_seed + e. */
int32_t bajo_render_cameras(bajo_ctx* ctx, bajo_scene* scene, const bajo_render_settings* settings,
                            const float* cameras, uint32_t camera_count, float* pixels, bajo_render_stats* stats);
/* A render target whose pixels (f32[E h w 3]) hold `environments` images:
 * it renders batches of that many environments (bajo_scene_create_batch). */
int32_t bajo_render_target_create_envs(bajo_ctx* ctx, const bajo_render_settings* settings, uint32_t environments,
                                       const float* camera, float* pixels, bajo_render_target** out);
int32_t bajo_render_enqueue(bajo_ctx* ctx, bajo_render_target* target, bajo_scene* scene);
int32_t bajo_render_target_capture(bajo_ctx* ctx, bajo_render_target* target, bajo_scene* scene);
/* Capture flags. */
#define BAJO_CAPTURE_REBUILD 1 /* record the scene's rebuild before its render */
/* bajo_render_target_capture with flags (bajo has no counterpart; its closest
 * code is the segmented build, bvh/gpu/builder/segmented_build.mojo:89-162,
 * and the enqueue-only render, rt/gpu/render.mojo:61-235).  With
 * BAJO_CAPTURE_REBUILD the graph holds bajo_scene_rebuild from the scene's
 * device vertices (bajo_scene_vertices) and then the render: for a batch,
 * the segmented build of every environment and the batched render, so one
 * bajo_render_target_replay rebuilds and renders the whole batch with no
 * host work.  One eager rebuild runs first (it sizes every buffer).  The
 * light tables are not re-derived: between replays move only non-emissive
 * triangles. */
int32_t bajo_render_target_capture_ex(bajo_ctx* ctx, bajo_render_target* target, bajo_scene* scene, uint32_t flags);
/* Device address and count of the scene's triangles f32[9 n] (every
 * environment of a batch, in order): the input of bajo_scene_rebuild, which
 * callers may rewrite on the context stream between rebuilds. */
int32_t bajo_scene_vertices(const bajo_scene* scene, float** vertices, uint64_t* count);
/* Enqueues a copy of device f32[9 n] into the scene's triangles (n = the
 * scene's count); the next bajo_scene_rebuild or captured rebuild uses them. */
int32_t bajo_scene_set_vertices(bajo_ctx* ctx, bajo_scene* scene, const float* vertices, uint64_t n);
int32_t bajo_render_target_replay(bajo_ctx* ctx, bajo_render_target* target);
int32_t bajo_render_target_free(bajo_render_target* target);
/* Host view of one stage's random stream (tests and tools): `count` floats
 * of the Philox stream keyed by (seed, path_id, stage). */
int32_t bajo_rng_stage_floats(uint64_t seed, uint32_t path_id, uint32_t stage, uint32_t count, float* out);

/* ---- learned path guiding (no bajo counterpart) ---------------------------
 * A small network maps a path vertex (position normalized to the scene box,
 * shading normal, outgoing direction; 12-frequency sin/cos encoding -> 216
 * features) through three hidden layers of 64 ReLU units to an 8-lobe von
 * Mises-Fisher mixture over directions (40 outputs: per lobe a mean
 * direction 3, log concentration 1, logit 1).  Weights are BF16 on the
 * matrix cores with FP32 accumulation; the FP32 master weights are trained
 * with Adam on the KL divergence to the incident radiance the path tracer
 * measures: per sample -w log q(direction), w = radiance luminance / density
 * it was drawn with, normalized by the batch's sum of w.  The guided renderer
 * (BAJO_RENDER_GUIDED) picks the guide or the BSDF with probability 1/2 at
 * every Lambertian and GGX vertex and weights by the combined density
 * (one-sample MIS), on top of NEE; bajo's BSDF sampling
 * (rt/gpu/path_shading.mojo:463-535, rt/shading.mojo:92-118) is the other
 * half and, without a guide, the whole of it (plain NEE). */
typedef struct bajo_guide bajo_guide;
#define BAJO_GUIDE_PARAMS 24808
#define BAJO_GUIDE_OUTPUTS 40
#define BAJO_GUIDE_SAMPLE_FLOATS 13 /* inputs 9, direction 3, weight 1 */
#define BAJO_RENDER_GUIDED 5
typedef struct {
  float learning_rate;      /* Adam step size; 0 = 1e-2                          */
  float beta1, beta2;       /* Adam moment decays; 0 = 0.9 / 0.999               */
  float epsilon;            /* Adam epsilon; 0 = 1e-8                            */
  uint32_t batch_size;      /* samples per Adam step (multiple of 32); 0 = 65536  */
  uint32_t steps_per_frame; /* Adam steps after a training render; 0 = 4         */
  uint64_t seed;            /* weight initialization (He-uniform, output x 0.1)  */
} bajo_guide_options;
int32_t bajo_guide_create(bajo_ctx* ctx, const bajo_guide_options* options, bajo_guide** out);
int32_t bajo_guide_free(bajo_guide* guide);
/* Host copies of the FP32 master weights f32[BAJO_GUIDE_PARAMS] (W1[64][216],
 * b1[64], W2[64][64], b2, W3[64][64], b3, W4[40][64], b4[40]; W[j][k] maps
 * input k to output j); set_params also resets the Adam state. */
int32_t bajo_guide_get_params(bajo_ctx* ctx, const bajo_guide* guide, float* params);
int32_t bajo_guide_set_params(bajo_ctx* ctx, bajo_guide* guide, const float* params);
/* Network outputs f32[40 n] for inputs f32[9 n] (device buffers). */
int32_t bajo_guide_forward(bajo_ctx* ctx, const bajo_guide* guide, const float* inputs, uint64_t n, float* outputs);
/* Loss gradient f32[BAJO_GUIDE_PARAMS] (device) of one batch of samples
 * f32[13 n] (device; n <= batch_size), without an update; `loss` (host,
 * may be NULL) receives the normalized loss (synchronizes). */
int32_t bajo_guide_gradient(bajo_ctx* ctx, bajo_guide* guide, const float* samples, uint64_t n, float* gradient,
                            float* loss);
/* Adam steps over samples f32[13 n] (device), one per batch_size samples. */
int32_t bajo_guide_train(bajo_ctx* ctx, bajo_guide* guide, const float* samples, uint64_t n);
/* Decodes outputs f32[40 n] into mixtures and, per row, draws a direction
 * from uniforms f32[3 n] (lobe, then the lobe's two) into dirs f32[3 n] and
 * its mixture density into pdfs f32[n] (device buffers). */
int32_t bajo_guide_mixture_sample(bajo_ctx* ctx, const float* outputs, uint64_t n, const float* uniforms, float* dirs,
                                  float* pdfs);
/* Host copy of the training samples the last training frame collected
 * (f32[13] each, at most `cap`); *count receives how many it collected. */
int32_t bajo_guide_training_samples(bajo_ctx* ctx, const bajo_guide* guide, float* samples, uint64_t cap,
                                    uint64_t* count);
/* Per-frame figures of a guided render (bajo_render_guided). */
typedef struct {
  uint64_t guided_vertices;  /* vertices that sampled through the guide's MIS  */
  uint64_t training_samples; /* samples collected for training (train != 0)   */
  uint32_t adam_steps;       /* Adam steps taken after the frame              */
  float loss;                /* mean normalized loss of the last step          */
  uint64_t infer_ns, train_ns, render_ns; /* device time of network inference, training, whole frame */
} bajo_guide_frame_stats;
/* One frame of BAJO_RENDER_GUIDED: NEE with guided continuation (settings'
 * algorithm is ignored).  With `train` the frame records its vertices and
 * the radiance that reached them, and Adam steps over those samples follow
 * (they change the guide for the next frame).  `guide` NULL renders plain
 * NEE.  The frame must fit one path chunk (path_capacity 0 or >= paths).
 * With `stats` it synchronizes. */
int32_t bajo_render_guided(bajo_ctx* ctx, bajo_scene* scene, bajo_guide* guide, const bajo_render_settings* settings,
                           const float* camera, float* pixels, int32_t train, bajo_guide_frame_stats* stats);

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
