/*
 * CPU oracle -- test infrastructure only.
 *
 * A plain-C restatement of bajo's CPU/GPU BVH algorithms (see the file
 * headers of orc_lbvh.c, orc_trace.c and orc_render.c for the reference
 * lines each function follows).  Loaded by the tests, the smoke check and
 * the CPU-baseline leg of the benchmark; never by the product library.
 */
#ifndef BAJO_ORACLE_H
#define BAJO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* core/morton.mojo */
uint32_t orc_expand_bits(uint32_t n);
uint32_t orc_morton(float x, float y, float z);
int orc_common_prefix(uint32_t ca, uint32_t ia, uint32_t cb, uint32_t ib);

/* LBVH build pieces */
void orc_triangle_bounds(const float *v, uint64_t n, float *bounds6);
void orc_scene_bounds(const float *bounds6, uint64_t n, float *out12);
void orc_morton_codes(const float *bounds6, uint64_t n, const float *centroid_box6, uint32_t *codes);
int orc_radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, uint64_t n, int begin_bit, int end_bit);
int orc_radix_sort_morton_pairs(uint32_t *codes, uint32_t *ids, uint64_t n);
int64_t orc_lbvh_find_split(const uint32_t *codes, int64_t n, int64_t first, int64_t last);
uint32_t orc_karras_topology(const uint32_t *sorted_codes, uint64_t n, uint32_t *node_meta, uint32_t *leaf_parent,
                             uint32_t *leaf_counts);
void orc_refit(uint32_t root, const float *leaf_bounds, const uint32_t *leaf_ids, const uint32_t *node_meta,
               float *node_bounds);
uint32_t orc_refit_climb(const float *leaf_bounds, uint64_t n, const uint32_t *node_meta,
                         const uint32_t *leaf_parent, float *node_bounds);
uint32_t orc_build_lbvh(const float *vertices, uint64_t n, uint64_t *sorted_keys, uint32_t *leaf_ids,
                        uint32_t *node_meta, uint32_t *leaf_parent, uint32_t *leaf_counts, float *node_bounds,
                        float *scene12);
uint32_t orc_build_lbvh_bounds(const float *bounds, uint64_t n, uint64_t *sorted_keys, uint32_t *leaf_ids,
                               uint32_t *node_meta, uint32_t *leaf_parent, uint32_t *leaf_counts, float *node_bounds,
                               float *scene12);
uint32_t orc_build_lbvh_topdown(const uint32_t *sorted_codes, const uint32_t *sorted_ids, const float *bounds6,
                                uint64_t n, int leaf_size, float *boxes6, uint32_t *first, uint32_t *count);
uint64_t orc_canonical_karras(uint32_t root, const uint32_t *node_meta, const uint32_t *leaf_counts, uint32_t *out);
uint64_t orc_canonical_topdown(const uint32_t *first, const uint32_t *count, uint32_t *out);

/* ray queries */
void orc_camera_make(const float *origin3, const float *target3, const float *up3, float fov_scale,
                     float focus_dist, float *out20);
void orc_camera_rays(const float *cams, uint64_t R, uint32_t w, uint32_t h, float *rays);
void orc_trace_bvh2(const float *vertices, const uint32_t *leaf_ids, const uint32_t *node_meta,
                    const float *node_bounds, uint32_t root, const float *rays, uint64_t R, int mode, float *hits,
                    uint32_t *stats);
void orc_trace_bvh2_mt(const float *vertices, const uint32_t *leaf_ids, const uint32_t *node_meta,
                       const float *node_bounds, uint32_t root, const float *rays, uint64_t R, int mode,
                       float *hits, int threads);
void orc_trace_brute(const float *vertices, uint64_t n, const float *rays, uint64_t R, float *hits);
int orc_intersect_aabb(const float *o3, const float *rcp3, const float *box6, float t_max, float *t_out);
int orc_intersect_tri_eager(const float *o3, const float *d3, const float *v0, const float *e1, const float *e2,
                            float t_max, float t_min, float *tuv3);
int orc_intersect_tri_scaled(const float *o3, const float *d3, const float *v0, const float *e1, const float *e2,
                             float t_max, float t_min, float *dtuv4);
void orc_rcp_direction(const float *d3, float *out3);

/* 8-wide collapse and traversal (orc_wide.c) */
uint32_t orc_grid_exponent(float o, float hi);
void orc_quantize(float o, uint32_t e, float lo, float hi, uint8_t *qlo, uint8_t *qhi);
uint64_t orc_collapse_wide(const uint32_t *node_meta, const float *node_bounds, const uint32_t *leaf_counts,
                           uint64_t n, uint32_t root, const float *scene_bounds, uint32_t leaf_size, void *out_nodes,
                           uint64_t capacity);
uint64_t orc_collapse_wide_n(const uint32_t *node_meta, const float *node_bounds, const uint32_t *leaf_counts,
                             uint64_t n, uint32_t root, const float *scene_bounds, uint32_t leaf_size, uint32_t width,
                             void *out_nodes, uint64_t capacity);
uint64_t orc_trace_wide(const float *vertices, uint64_t n, const uint32_t *leaf_ids, const void *wide,
                        uint64_t wide_count, const float *rays, uint64_t R, int mode, float *hits, uint32_t *stats);

/* H-PLOC topology (orc_hploc.c) */
uint32_t orc_hploc_topology(const float *leaf_bounds6, const uint32_t *sorted_codes, const uint32_t *sorted_ids,
                            uint64_t n, int search_radius, int merging_threshold, uint32_t *parent, uint32_t *left,
                            uint32_t *right, uint32_t *leaf_id, uint32_t *leaf_count, float *bounds6,
                            uint64_t *stats6);
uint64_t orc_preorder_topology(uint32_t root, const uint32_t *left, const uint32_t *right, const uint32_t *leaf_id,
                               uint64_t node_count, uint32_t *out);
uint64_t orc_preorder_binary(uint32_t root, const uint32_t *node_meta, const uint32_t *leaf_ids, uint64_t n,
                             uint32_t *out);
uint32_t orc_hploc_binary(uint32_t root, const uint32_t *left, const uint32_t *right, const uint32_t *leaf_id,
                          const uint32_t *leaf_count, const float *bounds6, uint64_t n, const uint32_t *numbering,
                          uint32_t *node_meta, float *node_bounds, uint32_t *leaf_ids, uint32_t *leaf_counts,
                          uint32_t *leaf_parent);
void orc_hploc_lowest_leaf_numbering(const uint32_t *left, const uint32_t *right, const uint32_t *sorted_ids,
                                     uint64_t n, uint32_t *out);

/* sphere BVHs (orc_trace.c, orc_wide.c) */
void orc_trace_bvh2_spheres(const float *spheres, const uint32_t *leaf_ids, const uint32_t *node_meta,
                            const float *node_bounds, uint32_t root, const float *rays, uint64_t R, int mode,
                            float *hits, uint32_t *stats);
void orc_trace_brute_spheres(const float *spheres, uint64_t n, const float *rays, uint64_t R, float *hits);
uint64_t orc_trace_wide_spheres(const float *spheres, uint64_t n, const uint32_t *leaf_ids, const void *wide,
                                uint64_t wide_count, const float *rays, uint64_t R, int mode, float *hits,
                                uint32_t *stats);

/* instances (orc_tlas.c); orc_bvh is declared in orc_internal.h */
int orc_affine_inverse(const float *a, float *o);
void orc_transform_box(const float *m, const float *b, float *out);
struct orc_bvh_s;

/* wavefront renderer (orc_render.c) */
void orc_ao_rays(const float *rays, const float *hits, uint64_t n, uint32_t samples, uint64_t seed, float *out);
typedef struct {
    uint32_t width, height, spp, max_depth;
    uint64_t seed;
    int algorithm; /* 0 PATH, 1 NORMALS, 2 AO, 3 NEE, 4 MIS */
    int node_width; /* 2 (binary LBVH) or 8 (collapsed 8-wide, orc_wide.c); 0 = 2 */
    int leaf_size;  /* width 8: triangles per leaf child, 1..8; 0 = 4 */
    int method;     /* 0 LBVH, 1 H-PLOC (orc_hploc.c, depth-first leaf order) */
    const float *ggx; /* f32[4] per GGX material (F0.rgb, roughness) or NULL */
} orc_render_settings;
void orc_eval_ggx(const float *ray_d, const float *n, const float *f0, float alpha, const float *dir, float *out);
void orc_sample_ggx(const float *ray_d, const float *n, const float *f0, float alpha, float u1, float u2, float *out);

/* mixed triangle + sphere scenes (orc_render.c); spheres f32[4S] with signed radii */
int orc_render_mixed(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *spheres, uint64_t ns,
                     const uint32_t *sphere_surfaces, const float *lambertian, const float *metal,
                     const float *dielectric, const float *emissive, const orc_render_settings *st,
                     const float *cam20, float *pixels, uint64_t *ray_counts, int threads);
int orc_render_fp64_mixed(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *spheres,
                          uint64_t ns, const uint32_t *sphere_surfaces, const float *lambertian, const float *metal,
                          const float *dielectric, const float *emissive, const orc_render_settings *st,
                          const float *cam20, double *pixels, int threads);
void orc_light_table_mixed(const float *v, uint64_t n, const uint32_t *surf, const float *sph, uint64_t ns,
                           const uint32_t *ssurf, const float *emissive, float *fields, uint32_t *kinds,
                           uint32_t *count, float *total);

void orc_philox4x32_10(const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out);
void orc_rng_stage_floats(uint64_t seed, uint32_t path_id, uint32_t stage, uint32_t count, float *out);
void orc_sincos_turns(float u, float *s, float *c);
void orc_alias_table(const float *w, uint32_t n, float total, float *prob, uint32_t *alias);
void orc_light_table(const float *v, uint64_t n, const uint32_t *surf, const float *emissive, float *fields,
                     uint32_t *kinds, uint32_t *count, float *total);
int orc_render(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *lambertian,
               const float *metal, const float *dielectric, const float *emissive, const orc_render_settings *st,
               const float *cam20, float *pixels, uint64_t *ray_counts, int threads);
float orc_pow_unit(float x, float y);
float orc_reflectance(float cosine, float ref_idx);
void orc_reflect3(const float *v, const float *n, float *out);
void orc_sample_metal(const float *ray_d, const float *n, const float *albedo, float fuzz, float ru, float rv,
                      float *out);
void orc_eval_metal(const float *ray_d, const float *n, const float *albedo, float fuzz, const float *out_dir,
                    float *out);
void orc_sample_dielectric(const float *ray_d, const float *n, float ior, int front, float ru, float *out);
int orc_render_fp64(const float *vertices, uint64_t n, const uint32_t *surfaces, const float *lambertian,
                    const float *metal, const float *dielectric, const float *emissive,
                    const orc_render_settings *st, const float *cam20, double *pixels, int threads);

#ifdef __cplusplus
}
#endif
#endif
