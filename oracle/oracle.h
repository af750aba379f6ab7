/*
 * oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle restates the reference's hot-path
 * semantics (ProjectAsura/Salty) on the CPU so that the HIP path can be checked
 * against it.  Only the tests, the smoke check and the CPU-baseline leg of
 * bench.py load it; the product library never does.
 */
#ifndef S3D_ORACLE_H
#define S3D_ORACLE_H

#include <stdint.h>
#include "s3dg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- random numbers ----------------------------------------------------- */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void     orc_pcg_stream(uint64_t seed, int32_t count, uint32_t *out);
void     orc_pcg_srandom_stream(uint64_t initstate, uint64_t initseq, int32_t count, uint32_t *out);
float    orc_pcg_f32(uint32_t x);
uint64_t orc_path_seed(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t scene);

/* --- single-function probes ---------------------------------------------- */
int   orc_triangle_hit(const float v0[3], const float v1[3], const float v2[3],
                       const s3dg_ray *ray, s3dg_hit *hit);
void  orc_tangent_space(const float n[3], float t[3], float b[3]);
void  orc_sample_lambert(float u, float v, float out[3]);
void  orc_sample_phong(float u, float v, float power, float out[3]);
int   orc_shade(const s3dg_material *mat, const float input[3], const float normal[3],
                uint64_t *pcg_state, float out_dir[3], float weight[3]);
void  orc_invert(const float m[16], float out[16]);

/* --- meshes: reference SAH tree + reference traversal -------------------- */
void *orc_mesh_create(uint32_t vertex_count, const s3dg_vertex *vertices, int32_t bvh_kind);
void  orc_mesh_destroy(void *mesh);
int   orc_mesh_order(const void *mesh, int32_t *order);
/* Tree dump.  Returns the node count; with non-null arrays fills per node:
 * kind (8,4,2 or 0=leaf), first child slot in `children`, child count, and
 * for leaves the [begin,begin+count) range of the depth-first order. */
int   orc_mesh_tree(const void *mesh, int32_t *kind, int32_t *child_begin, int32_t *child_count,
                    int32_t *children, int32_t *leaf_begin, int32_t *leaf_count,
                    float *boxes /* node count x 6 */);
/* method 0: reference traversal (exact slab tests, no t-clipping);
 * method 1: brute force over all triangles in depth-first order. */
int   orc_mesh_trace(const void *mesh, const s3dg_ray *rays, uint32_t count, int32_t method,
                     s3dg_hit *hits, int32_t threads);

/* --- scenes --------------------------------------------------------------- */
void *orc_scene_create(const s3dg_scene_desc *desc);
void  orc_scene_destroy(void *scene);
int   orc_scene_trace(const void *scene, const s3dg_ray *rays, uint32_t count,
                      s3dg_hit *hits, int32_t *shapes, int32_t threads);
void  orc_camera_ray(const void *scene, float x, float y, float org[3], float dir[3]);
/* Renders rows [row_begin,row_end) of cfg into accum (width*height*3 floats,
 * added to).  counts (optional): extension rays, shadow rays, paths. */
int   orc_render(const void *scene, const s3dg_render_config *cfg, int32_t row_begin,
                 int32_t row_end, float *accum, uint64_t counts[3], int32_t threads);
/* First-hit guide buffers (albedo, normal, depth), pass = first_pass. */
int   orc_render_guides(const void *scene, const s3dg_render_config *cfg,
                        float *albedo, float *normal, float *depth);

/* --- post ------------------------------------------------------------------ */
int   orc_tonemap(int32_t type, int32_t width, int32_t height, const float *in_rgb,
                  float *out_rgb, float *ave_lw, float *max_lw);
int   orc_capture(int32_t width, int32_t height, const float *tonemapped_rgb, uint8_t *out_rgb8);
int   orc_nlm(int32_t width, int32_t height, float coeff, int32_t search_radius,
              int32_t patch_radius, const float *src_rgba, float *dst_rgba, int32_t threads);

#ifdef __cplusplus
}
#endif

#endif
