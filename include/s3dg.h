/*
 * s3dg.h -- C ABI of the MI355X rendering hot path (BVH traversal, ray/triangle
 * intersection, wavefront path tracing, tonemapping, NLM denoising).
 *
 * Every entry point returns an int status (S3DG_OK == 0).  Plain pointers and
 * sizes only.  Device memory is owned by the context; calls on one context are
 * serialised on the context's single HIP stream.  Functions with a `_device`
 * suffix take device pointers that already live in HBM; the others take host
 * pointers and include the PCIe copies.
 *
 * The reference (ProjectAsura/Salty) has no C ABI: its hot path sits behind C++
 * abstract interfaces.  Each declaration below names the reference interface it
 * replaces (paths relative to the reference root).
 */
#ifndef S3DG_H
#define S3DG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Status codes
 * ------------------------------------------------------------------------- */
#define S3DG_OK                 0
#define S3DG_ERR_INVALID       -1   /* bad argument (null pointer, size, enum) */
#define S3DG_ERR_NOMEM         -2   /* host or device allocation failed        */
#define S3DG_ERR_DEVICE        -3   /* a HIP call failed                       */
#define S3DG_ERR_NODEVICE      -4   /* no usable gfx950 device                 */
#define S3DG_ERR_UNSUPPORTED   -5   /* request outside the supported envelope  */

/* ---------------------------------------------------------------------------
 * Scene description (plain data; copied by s3dg_scene_create)
 * ------------------------------------------------------------------------- */

/* s3d::Vertex (src/include/s3d_shape.h:27-32): 32 bytes. */
typedef struct s3dg_vertex {
    float position[3];
    float normal[3];
    float texcoord[2];
} s3dg_vertex;

/* Material kinds: s3d::Lambert, Mirror, Glass, Phong, Plastic
 * (src/src/s3d_{lambert,mirror,glass,phong,plastic}.cpp). */
enum {
    S3DG_MAT_LAMBERT = 0,
    S3DG_MAT_MIRROR  = 1,
    S3DG_MAT_GLASS   = 2,
    S3DG_MAT_PHONG   = 3,
    S3DG_MAT_PLASTIC = 4
};

/* Union of the MaterialFactory::Create* parameters
 * (src/include/s3d_materialfactory.h:22-49).  A texture wraps the material in
 * a TexturedMaterial (src/src/s3d_texturedmaterial.cpp:63-87): Shade and
 * GetBaseColor are multiplied by Texture::SampleColor(texcoord), the emission
 * is not. */
typedef struct s3dg_material {
    int32_t type;           /* S3DG_MAT_*                              */
    float   diffuse[3];     /* Lambert / Plastic diffuse colour         */
    float   specular[3];    /* Mirror / Glass / Phong / Plastic colour  */
    float   emissive[3];    /* IMaterial::GetEmissive                   */
    float   ior;            /* Glass index of refraction                */
    float   power;          /* Phong / Plastic exponent                 */
    int32_t texture;        /* 0: none; k > 0: the scene's texture k-1  */
} s3dg_material;

/* s3d::Texture (src/include/s3d_texture.h): width x height RGB texels, texel
 * (x, y) at rgb[3 * (width * y + x)] (Texture::GetColor); rows as the file
 * stores them (a BMP's first row is the image's bottom row). */
typedef struct s3dg_texture_desc {
    int32_t      width;
    int32_t      height;
    const float *rgb;
} s3dg_texture_desc;

/* Acceleration structure kind built for a mesh: BVH2::Create, BVH4::Create
 * or BVH8::Create (src/include/s3d_bvh{2,4,8}.h).  Mesh::Init always uses
 * BVH8 (src/src/s3d_mesh.cpp:495). */
enum { S3DG_BVH2 = 2, S3DG_BVH4 = 4, S3DG_BVH8 = 8 };

/* Mesh::Create(u32 vertexCount, Vertex*, IMaterial*) (src/include/s3d_mesh.h:42-47):
 * vertex_count is a multiple of 3, one Triangle per 3 vertices.  A per-triangle
 * material table (as the SMD loader produces) may replace the single material. */
typedef struct s3dg_mesh_desc {
    uint32_t           vertex_count;
    const s3dg_vertex *vertices;
    int32_t            material;           /* used when triangle_material == NULL */
    const int32_t     *triangle_material;  /* optional, vertex_count/3 entries    */
    int32_t            bvh_kind;           /* S3DG_BVH8 (default), BVH4 or BVH2   */
} s3dg_mesh_desc;

/* Top-level shapes, intersected in list order like Scene::Intersect
 * (src/include/s3d_scene.h:63-75). */
enum {
    S3DG_SHAPE_MESH     = 0,   /* Mesh in world space                          */
    S3DG_SHAPE_INSTANCE = 1,   /* Instance::Create(mesh, world) (s3d_instance.cpp:22-32) */
    S3DG_SHAPE_SPHERE   = 2    /* Sphere::Create(radius, center, material)     */
};

typedef struct s3dg_shape_desc {
    int32_t type;          /* S3DG_SHAPE_*                                   */
    int32_t mesh;          /* mesh index for MESH / INSTANCE                 */
    int32_t material;      /* SPHERE: material; INSTANCE: material override  */
                           /* when >= 0 (an extension: the reference's       */
                           /* Instance keeps the mesh's materials), else -1  */
    float   center[3];     /* sphere centre                                  */
    float   radius;        /* sphere radius                                  */
    float   world[16];     /* instance world matrix, row-vector convention,  */
                           /* translation in world[12..14] (Matrix _41.._43) */
} s3dg_shape_desc;

/* Light list entries (Scene::m_pLightList, s3d_scene.h:84-90).  SHAPE lights
 * name a top-level sphere and use Sphere::Sample.  QUAD lights are a
 * parallelogram made of triangles of a MESH shape; sampling it uniformly is an
 * extension (the reference only samples spheres). */
enum { S3DG_LIGHT_SHAPE = 0, S3DG_LIGHT_QUAD = 1 };

typedef struct s3dg_light_desc {
    int32_t type;            /* S3DG_LIGHT_*                                */
    int32_t shape;           /* top-level shape index                       */
    int32_t first_triangle;  /* QUAD: first triangle of the light in mesh   */
    int32_t triangle_count;  /* QUAD: number of triangles (2 for a quad)    */
    float   corner[3];       /* QUAD: p = corner + u*edge1 + v*edge2        */
    float   edge1[3];
    float   edge2[3];
} s3dg_light_desc;

enum { S3DG_CAMERA_PINHOLE = 0, S3DG_CAMERA_THINLENS = 1 };

/* PinholeCamera::Update / ThinLensCamera::Update (src/include/s3d_camera.h:63-90,171-202).
 * fov is used linearly by the reference: the vertical field of view is
 * 2*atan(fov/2). */
typedef struct s3dg_camera_desc {
    int32_t type;
    float   position[3];
    float   target[3];
    float   upward[3];
    float   fov;
    float   aspect;
    float   near_clip;
    float   lens_radius;   /* THINLENS; focus distance is |target-position| */
    float   reserved;
} s3dg_camera_desc;

typedef struct s3dg_scene_desc {
    uint32_t               mesh_count;
    const s3dg_mesh_desc  *meshes;
    uint32_t               shape_count;
    const s3dg_shape_desc *shapes;
    uint32_t               material_count;
    const s3dg_material   *materials;
    uint32_t               light_count;
    const s3dg_light_desc *lights;
    s3dg_camera_desc       camera;
    float                  environment[3];  /* radiance on a miss when no IBL texture */
    uint32_t               texture_count;
    const s3dg_texture_desc *textures;
    int32_t                environment_texture;  /* 0: constant environment; k > 0:  */
                           /* textures[k-1] as the IBL, Scene::SampleIBL: lat-long     */
                           /* Texture::SampleColor(dir) * 10 (src/include/s3d_scene.h:80-82, */
                           /* src/src/s3d_texture.cpp:957-972)                          */
} s3dg_scene_desc;

/* ---------------------------------------------------------------------------
 * Rays and hits
 * ------------------------------------------------------------------------- */

/* 32 bytes.  The reference hard-codes t in [F_HIT_MIN=1e-3, F_HIT_MAX=1e12]
 * (src/include/s3d_math.h:44-45); per-ray tmin/tmax generalise that. */
typedef struct s3dg_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} s3dg_ray;

/* 16 bytes: HitRecord (src/include/s3d_shape.h:38-54) with the triangle pointer
 * replaced by the triangle's index in its mesh's input order (-1 = miss, -2 =
 * sphere).  u, v are the barycentrics (beta, gamma) of Triangle::IsHit. */
typedef struct s3dg_hit {
    int32_t prim;
    float   t;
    float   u;
    float   v;
} s3dg_hit;

enum { S3DG_TRACE_CLOSEST = 0, S3DG_TRACE_ANY = 1 };

/* ---------------------------------------------------------------------------
 * Render configuration: superset of PathTracer::Config (src/include/s3d_pt.h:33-40)
 * ------------------------------------------------------------------------- */
typedef struct s3dg_render_config {
    int32_t  width;
    int32_t  height;
    int32_t  max_bounce_count;
    int32_t  spp;              /* passes of one sample per pixel               */
    uint64_t seed;             /* counter-based RNG key                        */
    int32_t  rr_start_depth;   /* dice honoured from this depth on (0 = reference) */
    int32_t  first_pass;       /* pass index of the first sample (progressive) */
    uint32_t flags;            /* S3DG_RENDER_* instrumentation bits           */
    uint32_t passes_in_flight; /* passes traced together per wavefront (0 = auto) */
} s3dg_render_config;

/* flags: time every traversal launch with HIP events on the context stream;
 * count compressed nodes visited and triangles tested by the traversals;
 * always use the per-bounce wavefront pipeline; trace each path in one
 * thread, in one launch, whatever the number of paths.  Without the last two
 * a render of one batch of up to four times the device's resident traversal
 * threads (about 1.5 M paths) is traced thread per path, a larger one by the
 * pipeline; both give the same sums and counts.  The first two flags, and a
 * render of several batches, imply the pipeline. */
#define S3DG_RENDER_TIME_KERNELS   1u
#define S3DG_RENDER_COUNT_WORK     2u
#define S3DG_RENDER_WAVEFRONT      4u
#define S3DG_RENDER_PER_PATH       8u

typedef struct s3dg_render_stats {
    uint64_t extension_rays;   /* closest-hit rays traced                      */
    uint64_t shadow_rays;      /* NEE visibility rays traced                   */
    uint64_t paths;            /* camera samples completed                     */
    double   milliseconds;     /* device time of the render                    */
    double   extend_ms;        /* TIME_KERNELS: summed closest-hit kernel time */
    double   shadow_ms;        /* TIME_KERNELS: summed shadow kernel time      */
    uint64_t extend_launches;
    uint64_t shadow_launches;
    uint64_t extend_nodes;     /* COUNT_WORK: nodes visited by extension rays  */
    uint64_t extend_tris;      /* COUNT_WORK: triangles tested                 */
    uint64_t shadow_nodes;
    uint64_t shadow_tris;
    uint64_t extend_stack_spills;   /* COUNT_WORK: traversal stack pushes beyond the */
    uint64_t shadow_stack_spills;   /* entries kept in LDS (to global memory)        */
} s3dg_render_stats;

typedef struct s3dg_scene_stats {
    uint64_t triangles;        /* unique triangles stored                      */
    uint64_t blas_nodes;       /* compressed 8-wide nodes of all meshes        */
    uint64_t tlas_nodes;
    uint64_t device_bytes;     /* node + triangle + shading data in HBM        */
    int32_t  max_depth;        /* deepest node path of the two-level tree      */
    double   build_ms;         /* host SAH build + re-layout time              */
} s3dg_scene_stats;

/* Tonemap operators (src/include/s3d_tonemapper.h:20-33). */
enum {
    S3DG_TONEMAP_REINHARD   = 0,
    S3DG_TONEMAP_UNCHARTED2 = 1,
    S3DG_TONEMAP_ACES       = 2
};

typedef struct s3dg_context s3dg_context;
typedef struct s3dg_scene   s3dg_scene;

/* ---------------------------------------------------------------------------
 * Context
 * ------------------------------------------------------------------------- */
/* No reference counterpart (the reference has no device); owns the HIP stream
 * and device allocations.  device < 0 selects the current device. */
int         s3dg_create(s3dg_context **out, int device);
int         s3dg_destroy(s3dg_context *ctx);
const char *s3dg_last_error(const s3dg_context *ctx);
int         s3dg_synchronize(s3dg_context *ctx);
int         s3dg_malloc(s3dg_context *ctx, size_t bytes, void **out_device_ptr);
int         s3dg_free(s3dg_context *ctx, void *device_ptr);
int         s3dg_memcpy_htod(s3dg_context *ctx, void *dst_device, const void *src_host, size_t bytes);
int         s3dg_memcpy_dtoh(s3dg_context *ctx, void *dst_host, const void *src_device, size_t bytes);
/* HIP stream of the context (hipStream_t), for callers that time on it. */
void       *s3dg_stream(s3dg_context *ctx);

/* ---------------------------------------------------------------------------
 * Scene: replaces Mesh::Create -> BVH8::Create (src/src/s3d_mesh.cpp:470-498,
 * src/src/s3d_bvh8.cpp:401-479), Instance::Create, Sphere::Create and the
 * Scene container (src/include/s3d_scene.h:22-109).  Builds every mesh's tree
 * with the reference binned-SAH algorithm on the host, re-lays it out into the
 * compressed 128-byte 8-wide node format, adds a top-level tree over the shapes
 * and uploads everything.
 * ------------------------------------------------------------------------- */
int s3dg_scene_create(s3dg_context *ctx, const s3dg_scene_desc *desc, s3dg_scene **out);
int s3dg_scene_destroy(s3dg_scene *scene);
int s3dg_scene_get_stats(const s3dg_scene *scene, s3dg_scene_stats *out);

/* Permutation the reference builder applied to a mesh's triangles: order[i] is
 * the input index of the triangle at depth-first position i (BVH8::Create
 * partitions its IShape** argument in place, s3d_bvh8.cpp:321-334).  `order`
 * holds vertex_count/3 entries. */
int s3dg_scene_mesh_order(const s3dg_scene *scene, uint32_t mesh, int32_t *order);

/* Instance::UpdateMatrix (src/src/s3d_instance.cpp:95-101) for `count`
 * INSTANCE shapes: worlds holds 16 floats per shape (row-vector convention,
 * like s3dg_shape_desc.world).  The top-level tree is rebuilt over the
 * unchanged mesh trees; renders after the call see the new matrices (a render
 * equals that of a scene created with them).  Scene sets made earlier keep
 * their copy. */
int s3dg_scene_update_instances(s3dg_scene *scene, uint32_t count, const uint32_t *shapes, const float *worlds);

/* Host-only: the depth-first permutation that BVH2/4/8::Create would apply to
 * the triangles of Mesh::Create(vertex_count, vertices, ...) (src/src/s3d_mesh.cpp:470-498),
 * and the number of tree nodes, without a device.  kind is S3DG_BVH2/4/8. */
int s3dg_build_order(uint32_t vertex_count, const s3dg_vertex *vertices, int32_t kind, int32_t *order,
                     uint32_t *node_count);

/* ---------------------------------------------------------------------------
 * Host-only loaders of the reference's file formats.
 * ------------------------------------------------------------------------- */

/* Images as RGB floats (free with s3dg_image_free).  BMP: LoadFromBMP
 * (src/src/s3d_texture.cpp:1029-1077): 24-bit only, texels read right after the
 * 54 header bytes with no row padding, rows kept in file order, channel =
 * byte / 255; the reference leaves the last texel unwritten, here it is 0.
 * HDR: LoadFromHDR (src/src/s3d_texture.cpp:321-350, RGBE with old and new
 * run-length lines, exposure and gamma ignored as there). */
int s3dg_image_load_bmp(const char *path, int32_t *width, int32_t *height, float **rgb);
int s3dg_image_load_hdr(const char *path, int32_t *width, int32_t *height, float **rgb);
void s3dg_image_free(float *rgb);

/* SMD mesh file: Mesh::LoadFromFile (src/src/s3d_mesh.cpp:250-465).  Header
 * "SMD\0", version 2, data header of six u32 (triangle, material, texture
 * counts and struct sizes); texture names (256 bytes each, relative to the
 * file's folder, a '\\' in a name read as '/'); per material an s32 type then its fields; 100-byte
 * triangles (3 x {position, normal, texcoord} + s32 material).  Materials map
 * to s3dg_material records, texture = ColorMap (or DiffuseMap) + 1.  Rejected as there:
 * a bad magic, version or data-header size, no triangles or no materials; in
 * addition an unknown material type or a short file is an error here. */
typedef struct s3dg_smd {
    uint32_t       vertex_count;        /* 3 per triangle                      */
    s3dg_vertex   *vertices;
    int32_t       *triangle_material;   /* MaterialId per triangle (-1: none)  */
    uint32_t       material_count;
    s3dg_material *materials;
    uint32_t       texture_count;
    char         (*texture_paths)[512]; /* folder of the file + "/" + name      */
} s3dg_smd;

int s3dg_smd_load(const char *path, s3dg_smd **out);
void s3dg_smd_free(s3dg_smd *smd);

/* ---------------------------------------------------------------------------
 * Traversal: replaces Scene::Intersect / IShape::IsHit(const RaySet&, HitRecord&)
 * (src/include/s3d_scene.h:63-75, src/include/s3d_shape.h:60-68).
 * CLOSEST returns the hit Scene::Intersect records; ANY reports whether any
 * hit exists in [tmin, tmax] (prim >= 0 or -2 if so, -1 if not).
 * shapes (optional) receives the top-level shape index of each hit.
 * ------------------------------------------------------------------------- */
int s3dg_trace(s3dg_scene *scene, const s3dg_ray *rays, uint32_t count, int32_t mode,
               s3dg_hit *hits, int32_t *shapes);
int s3dg_trace_device(s3dg_scene *scene, const s3dg_ray *d_rays, uint32_t count, int32_t mode,
                      s3dg_hit *d_hits, int32_t *d_shapes);
/* s3dg_trace_device that also adds the number of compressed nodes visited and
 * triangles tested to d_counters[0] and d_counters[1] (uint64, device), the
 * inputs of the traversal kernel's algorithmic-bytes figure.  No reference
 * counterpart (instrumentation). */
int s3dg_trace_counted_device(s3dg_scene *scene, const s3dg_ray *d_rays, uint32_t count, int32_t mode,
                              s3dg_hit *d_hits, uint64_t *d_counters);

/* ---------------------------------------------------------------------------
 * Rendering: replaces PathTracer::TracePath / Radiance / NextEventEstimation
 * (src/src/s3d_pt.cpp:182-353).  Adds cfg->spp passes of one pixel-centre
 * sample to the accumulation buffer (a SUM, like m_RenderTarget), width*height
 * RGB floats.  The buffer is read-modify-written, so progressive calls with
 * increasing first_pass continue a render.
 * ------------------------------------------------------------------------- */
int s3dg_render(s3dg_scene *scene, const s3dg_render_config *cfg, float *accum_rgb,
                s3dg_render_stats *stats);
int s3dg_render_device(s3dg_scene *scene, const s3dg_render_config *cfg, float *d_accum_rgb,
                       s3dg_render_stats *stats);

/* Guide buffers for denoisers (no reference counterpart): first-hit albedo
 * (GetBaseColor), shading normal and distance per pixel, each W*H*3 / W*H floats. */
int s3dg_render_guides_device(s3dg_scene *scene, const s3dg_render_config *cfg,
                              float *d_albedo, float *d_normal, float *d_depth);

/* ---------------------------------------------------------------------------
 * Tuning of the traversal and render kernels (no reference counterpart), per
 * context.  s3dg_autotune times trial renders of `cfg` (its samples reduced
 * only beyond 2^27 paths) on the scene for the candidate settings and keeps
 * the fastest in the context; results do not depend on the tuning, only the
 * speed does.
 * ------------------------------------------------------------------------- */
typedef struct s3dg_tuning {
    uint32_t traversal_variant;  /* 0: 16 LDS stack entries, 5 waves/SIMD; 1: 12, 6; 2: 8, 6 */
    uint32_t refill;             /* idle lanes of a wave that trigger new rays (1..64)       */
    uint32_t tri_batch;          /* lanes with pending triangles that keep a step testing (1..65) */
    uint32_t passes_in_flight;   /* path capacity used when a render passes 0 (0 = automatic) */
    uint32_t fetch_chunk;        /* queue entries a traversal wave takes at a time (64..65536) */
    uint32_t leaf_size;          /* most triangles per leaf of the traversal trees of scenes
                                    created afterwards (1..8); 0: traverse the reference-shaped
                                    tree.  Tied hits are ranked by the reference tree either way */
    uint32_t tail_paths;         /* the wavefront's remaining bounces run thread per path in one
                                    launch once its queue holds at most this many paths
                                    (0xFFFFFFFF, the default: two per thread of the traversal
                                    grid; 0: never)                                           */
} s3dg_tuning;

int s3dg_get_tuning(s3dg_context *ctx, s3dg_tuning *out);
int s3dg_set_tuning(s3dg_context *ctx, const s3dg_tuning *tuning);
int s3dg_autotune(s3dg_scene *scene, const s3dg_render_config *cfg, s3dg_tuning *best, double *best_ms);

/* ---------------------------------------------------------------------------
 * Scene sets and batched rendering (no reference counterpart: the reference
 * renders one scene per process).  A set merges existing scenes of one
 * context into one set of device arrays (the scenes stay valid and
 * independent); s3dg_render_batch_device renders a list of frames of its
 * members, each with its own resolution, passes, seed and sum buffer, with the
 * paths of all frames sharing every launch.  A frame's sum equals that of
 * s3dg_render_device on the member scene with the same settings.  From
 * `common` only max_bounce_count, rr_start_depth, flags and passes_in_flight
 * (the path capacity in units of the largest frame; 0 = automatic) are used.
 * ------------------------------------------------------------------------- */
typedef struct s3dg_scene_set s3dg_scene_set;

typedef struct s3dg_frame_job {
    uint32_t scene;            /* member index in the set                      */
    int32_t  width;
    int32_t  height;
    int32_t  spp;
    uint64_t seed;
    int32_t  first_pass;
    int32_t  reserved;
    float   *d_accum_rgb;      /* device sum buffer, width*height*3 floats     */
} s3dg_frame_job;

int s3dg_scene_set_create(s3dg_context *ctx, uint32_t count, s3dg_scene *const *scenes, s3dg_scene_set **out);
int s3dg_scene_set_destroy(s3dg_scene_set *set);
int s3dg_render_batch_device(s3dg_scene_set *set, uint32_t frame_count, const s3dg_frame_job *frames,
                             const s3dg_render_config *common, s3dg_render_stats *stats);

/* ---------------------------------------------------------------------------
 * Post: replaces ToneMapper::Map (src/src/s3d_tonemapper.cpp:206-232) and the
 * sRGB/8-bit conversion of PathTracer::Capture (src/src/s3d_pt.cpp:137-171).
 * out_rgb (optional) receives the tonemapped float image, out_rgb8 (optional)
 * the clamped, sRGB-encoded 8-bit image.
 * ------------------------------------------------------------------------- */
int s3dg_tonemap(s3dg_context *ctx, int32_t type, int32_t width, int32_t height,
                 const float *in_rgb, float *out_rgb, uint8_t *out_rgb8);
int s3dg_tonemap_device(s3dg_context *ctx, int32_t type, int32_t width, int32_t height,
                        const float *d_in_rgb, float *d_out_rgb, uint8_t *d_out_rgb8);

/* s3dg_tonemap_device that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its reduction scratch (12,352 bytes: 1,024 double and 1,024 float partials
 * plus the two results) is a file-scope device array, one per GPU for the
 * life of the process, shared by every call on that GPU: calls that overlap
 * on one GPU (other streams, other contexts) race.  Images up to 2^28 pixels;
 * larger ones are rejected with S3DG_ERR_INVALID.  Results are identical to
 * s3dg_tonemap_device's. */
int s3dg_tonemap_device_noalloc(s3dg_context *ctx, int32_t type, int32_t width, int32_t height,
                                const float *d_in_rgb, float *d_out_rgb, uint8_t *d_out_rgb8);

/* Replaces FilterNLM(int w, int h, float coeff, const Color4*, Color4*)
 * (src/include/s3d_denoiser.h:16, src/src/s3d_denoiser.cpp:19-95).  The
 * reference fixes search radius 6 and patch radius 2; both are parameters
 * here.  src and dst are RGBA floats and must not alias. */
int s3dg_nlm(s3dg_context *ctx, int32_t width, int32_t height, float coeff,
             int32_t search_radius, int32_t patch_radius,
             const float *src_rgba, float *dst_rgba);
int s3dg_nlm_device(s3dg_context *ctx, int32_t width, int32_t height, float coeff,
                    int32_t search_radius, int32_t patch_radius,
                    const float *d_src_rgba, float *d_dst_rgba);

/* ---------------------------------------------------------------------------
 * Learned denoiser (no reference counterpart: the learned stand-in for
 * FilterNLM).  A 4-level U-Net (3x3 convolutions, 32/64/96/128 channels, ReLU,
 * average pooling, nearest upsampling, concatenated skips, 1x1 head to log
 * radiance) run with BF16 MFMA kernels and FP32 accumulation.
 * Input features: N x H x W x 10 floats (radiance 3, albedo 3, normal 3,
 * depth 1), H and W multiples of 8; output N x H x W x 3 linear radiance.
 * Training minimises the relative squared error
 * (pred - target)^2 / (pred^2 + 0.01), denominator held constant, with Adam;
 * pixels whose depth guide is 0 (camera ray missed) pass their input
 * radiance through and take no part in the loss.
 * ------------------------------------------------------------------------- */
typedef struct s3dg_cnn s3dg_cnn;

int s3dg_cnn_create(s3dg_context *ctx, uint64_t seed, s3dg_cnn **out);
int s3dg_cnn_destroy(s3dg_cnn *cnn);
int s3dg_cnn_param_count(const s3dg_cnn *cnn, uint64_t *count);
/* Flat FP32 parameters: per layer W[9][cout][cin] then b[cout], head last. */
int s3dg_cnn_get_params(s3dg_cnn *cnn, float *host, uint64_t count);
int s3dg_cnn_set_params(s3dg_cnn *cnn, const float *host, uint64_t count);
/* The second call with the same shape and buffers captures the launches
 * into a HIP graph, later ones replay it (same results). */
int s3dg_cnn_infer_device(s3dg_cnn *cnn, int32_t n, int32_t h, int32_t w, const float *d_features,
                          float *d_out_rgb);
/* No reference counterpart: the denoiser's input features [n][h][w][10] from
 * a render's device buffers (accumulated RGB sum of spp passes, guides from
 * s3dg_render_guides_device): sum / spp, albedo, normal, depth, as a caller
 * assembles them on the host, bit for bit. */
int s3dg_denoiser_features_device(s3dg_context *ctx, int32_t n, int32_t h, int32_t w, const float *d_accum,
                                  int32_t spp, const float *d_albedo, const float *d_normal, const float *d_depth,
                                  float *d_features);
/* Mean loss and the gradient of every parameter (host copy optional). */
int s3dg_cnn_gradient_device(s3dg_cnn *cnn, int32_t n, int32_t h, int32_t w, const float *d_features,
                             const float *d_target_rgb, double *loss, float *host_grad, uint64_t count);
int s3dg_cnn_train_step_device(s3dg_cnn *cnn, int32_t n, int32_t h, int32_t w, const float *d_features,
                               const float *d_target_rgb, float learning_rate, double *loss);

/* ---------------------------------------------------------------------------
 * Per-object calls of the reference's host interfaces, evaluated on the
 * device with the renderer's own device functions, for a C++ (or other) host
 * layer that keeps those interfaces.  Host arrays in and out; one kernel per
 * call; at most 2^24 queries per call.
 * ------------------------------------------------------------------------- */

/* One ShadingArg (src/include/s3d_material.h:21-29). */
typedef struct s3dg_shade_arg {
    float    input[3];      /* incoming direction                                 */
    float    normal[3];     /* shading normal                                     */
    float    texcoord[2];   /* where a TexturedMaterial samples its texture       */
    uint64_t random;        /* PCG32 state (PCG::m_State), advanced by the draws  */
    float    output[3];     /* out: sampled direction (ShadingArg::output)        */
    float    weight[3];     /* out: Shade's return value / GetBaseColor           */
    int32_t  dice;          /* out: ShadingArg::dice                              */
    int32_t  pad;
} s3dg_shade_arg;

/* IMaterial::Shade (src/include/s3d_material.h:38; Lambert, Mirror, Glass,
 * Phong, Plastic: src/src/s3d_{lambert,mirror,glass,phong,plastic}.cpp) for n
 * shading events of one material; with a texture it is a TexturedMaterial
 * (src/src/s3d_texturedmaterial.cpp:63-87).  material->texture is ignored. */
int s3dg_shade(s3dg_context *ctx, const s3dg_material *material, const s3dg_texture_desc *texture,
               uint32_t n, s3dg_shade_arg *args);
/* IMaterial::GetBaseColor(texcoord) (src/include/s3d_material.h:40) into
 * args[i].weight; only texcoord is read. */
int s3dg_base_color(s3dg_context *ctx, const s3dg_material *material, const s3dg_texture_desc *texture,
                    uint32_t n, s3dg_shade_arg *args);

/* ICamera::GetRay(x, y) (src/include/s3d_camera.h:19-25, 95-100, 207-225):
 * org_dir receives origin and direction (6 floats per ray) through the
 * image-plane points xy (2 floats each, [-0.5, 0.5]).  A thin-lens camera
 * draws its lens sample from random[i] (advanced); a pinhole camera takes
 * random = NULL. */
int s3dg_camera_rays(s3dg_context *ctx, const s3dg_camera_desc *camera, uint32_t n, const float *xy,
                     uint64_t *random, float *org_dir);

/* Sphere::Sample(PCG&, Vector3*, float*) (src/src/s3d_sphere.cpp:106-115). */
int s3dg_sphere_sample(s3dg_context *ctx, const float center[3], float radius, uint32_t n, uint64_t *random,
                       float *position, float *pdf);

/* IShape::CalcParam (src/include/s3d_shape.h:66; Triangle
 * src/src/s3d_triangle.cpp:107-121, Sphere src/src/s3d_sphere.cpp:94-104) of a
 * hit on a shape of a scene: triangle = index in its mesh's input order (-2:
 * the shape is a sphere, whose parameters follow from position).  texcoord is
 * the blend of the vertex coordinates when the scene keeps them (any material
 * textured), else 0, 0. */
typedef struct s3dg_param_query {
    int32_t shape;
    int32_t triangle;
    float   position[3];
    float   barycentric[2];
} s3dg_param_query;
int s3dg_scene_calc_param(s3dg_scene *scene, uint32_t n, const s3dg_param_query *queries, float *normal,
                          float *texcoord);

/* Scene::SampleIBL(dir) (src/include/s3d_scene.h:80-82): the IBL texture's
 * lat-long lookup times 10, or the constant environment. */
int s3dg_scene_environment(s3dg_scene *scene, uint32_t n, const float *dirs, float *rgb);

#ifdef __cplusplus
}
#endif

#endif /* S3DG_H */
