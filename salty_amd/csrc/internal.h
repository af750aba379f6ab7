// internal.h -- context and scene objects behind the C ABI (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/s3dg.h"
#include "bvh.h"
#include "gpu_types.h"

struct s3dg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    // scratch reused across calls
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // small reduction buffer (tonemap partial sums)
    void *small = nullptr;
    size_t small_bytes = 0;
    // this device's copy of the file-scope tonemap scratch array (post.hip)
    void *noalloc_scratch = nullptr;
    // traversal kernels hand new rays to a wave once this many of its lanes are idle
    uint32_t refill = 16;
    // a traversal step keeps testing triangles while this many lanes have some pending
    uint32_t tri_batch = 16;
    // queue entries a traversal wave takes from the shared cursor at a time
    uint32_t fetch_chunk = 128;
    // traversal kernel variant (index into kVariants)
    uint32_t variant = 1;
    // passes traced together when a render leaves it to the library (0: automatic)
    uint32_t passes_in_flight = 0;
    // most triangles per leaf of the traversal trees of scenes created from
    // now on (1..8); 0 traverses the reference-shaped tree instead
    uint32_t leaf_size = 2;
    // the wavefront's remaining bounces run thread per path once its
    // extension queue holds at most this many paths (capped at one per
    // thread of the traversal grid; 0: never)
    uint32_t tail_paths = 0xffffffffu;
};

namespace s3dg {

// Per-scene part of the device view: where traversal starts, the lights and
// camera.  A scene has one; a scene set (several scenes merged into one set of
// arrays) has one per member, indexed by the scene number each path carries.
struct SceneRec {
    uint32_t root;          // top-level root node, or the mesh root of a single untransformed mesh
    int32_t single_shape;   // shape index of that mesh (-1 with a top-level tree)
    int32_t light_base;     // first light of the scene in the light array
    int32_t light_count;
    float env[3];
    int32_t env_tex;        // IBL texture (index into SceneDev::textures), -1: constant env
    GCamera cam;
};

// Device-side view of a scene or a scene set, passed by value to kernels.
struct SceneDev {
    const GNode *nodes;
    const GTri *tris;
    const GTriShade *shade;
    const GInst *insts;
    const GMat *mats;
    const GLight *lights;
    const GShapeRef *shape_ref;     // top-level shape index -> record + material override
    const SceneRec *recs;
    const uint32_t *tri_rank;   // depth-first rank of each stored triangle (tie order)
    const GTexture *textures;   // texture records (null when the scene has none)
    const float *texels;        // RGB texels of all textures
    const GTriUV *tri_uv;       // vertex texture coordinates per stored triangle (null: no textures)
    int32_t stack_need;     // deepest stack the trees can require
    uint32_t tri_batch;     // lanes with pending triangles that keep a step testing
};

// Words of Workspace::graph_key.
constexpr int kGraphKeyWords = 13;

// Device buffers of the traversal and render kernels, owned by a scene or a
// scene set and sized for the largest launch so far.
struct Workspace {
    void *d_gstack = nullptr;       // traversal stack overflow area (global memory)
    size_t gstack_bytes = 0;
    int32_t *d_error = nullptr;     // word 0: error flag; word 16: batch traversal work cursor
    void *d_render = nullptr;       // path state
    size_t render_bytes = 0;
    void *d_batch = nullptr;        // batch descriptor and segment table of the render
    size_t batch_bytes = 0;
    hipGraphExec_t graph = nullptr; // captured bounce loop
    uint64_t graph_key[kGraphKeyWords] = {};   // launch parameters the graph was captured with
};

struct SceneMeshInfo {
    uint32_t tri_offset;    // first depth-first triangle of the mesh in the global array
    uint32_t tri_count;
    uint32_t root;          // compressed root node
    std::vector<int32_t> order;   // depth-first -> input index
    std::vector<int32_t> stored_of_input;   // input index -> position in the stored (leaf) order
};

}  // namespace s3dg

struct s3dg_scene {
    s3dg_context *ctx = nullptr;
    s3dg::SceneDev dev{};
    s3dg::SceneRec rec{};
    void *d_rec = nullptr;
    std::vector<s3dg::SceneMeshInfo> meshes;
    s3dg_scene_stats stats{};
    // device allocations
    void *d_nodes = nullptr, *d_tris = nullptr, *d_shade = nullptr, *d_insts = nullptr;
    void *d_mats = nullptr, *d_lights = nullptr, *d_shape_ref = nullptr, *d_rank = nullptr;
    void *d_textures = nullptr, *d_texels = nullptr, *d_uv = nullptr;
    struct {
        uint32_t nodes, tris, insts, mats, lights, shapes;   // element counts of the device arrays
        uint32_t textures;                                    // texture records
        uint64_t texel_floats;                                // floats of the texel array
    } counts{};
    s3dg::Workspace ws;
    // host copies an instance update rebuilds the top level from
    struct {
        std::vector<s3dg::GInst> inst;     // shape records in shape order
        std::vector<int32_t> shape_mesh;   // mesh of every shape (-1: sphere)
        std::vector<s3dg::Aabb> wbox;      // world box of every shape
        std::vector<s3dg::Aabb> mesh_box;  // box of every mesh
        std::vector<s3dg::GLight> lights;
        uint32_t tlas_base = 0, tlas_cap = 0;   // first top-level node, nodes reserved for the top level
        int32_t blas_depth = 0;
    } host;
};

namespace s3dg {

// One frame of a render: scene (index into SceneDev::recs), resolution,
// passes, RNG key, first pass index and its RGB sum buffer in device memory.
struct RenderJob {
    uint32_t scene;
    int32_t width, height, spp;
    uint64_t seed;
    int32_t first_pass;
    float *d_accum;
};

int render_jobs(s3dg_context *ctx, const SceneDev &S, Workspace &ws, const RenderJob *jobs, size_t njobs,
                int32_t max_bounce, int32_t rr_start, uint32_t flags, uint32_t passes_in_flight,
                s3dg_render_stats *stats);
int set_error(s3dg_context *ctx, int code, const std::string &msg);
int workspace_init(s3dg_context *ctx, Workspace &ws);
void workspace_free(Workspace &ws);
int ensure_gstack(s3dg_context *ctx, Workspace &ws, uint32_t threads, int32_t stack_need, int lds);
int check_error_flag(s3dg_context *ctx, Workspace &ws);
int resident_grid(s3dg_context *ctx, const void *kernel, uint32_t *blocks);
uint32_t batch_grid(uint32_t n);
int hip_check(s3dg_context *ctx, hipError_t e, const char *what);
int tonemap_scratch_init(s3dg_context *ctx);
// Device records built from the ABI descriptions (scene.cpp).
GCamera make_camera(const s3dg_camera_desc &c);
void make_material(const s3dg_material &m, GMat &g);
int ensure_buffer(s3dg_context *ctx, void **ptr, size_t *have, size_t need);

constexpr int kBlock = 256;

// Traversal kernel variants: stack entries held in LDS per lane before
// spilling to global memory, and the waves per SIMD the register allocation
// aims at.  Fewer LDS entries and registers let more waves hide memory
// latency; the autotuner picks one per context.
struct TraversalVariant {
    int lds, waves;
};
constexpr TraversalVariant kVariants[] = {{16, 5}, {12, 6}, {8, 6}};
constexpr int kVariantCount = 3;
constexpr int kDefaultVariant = 1;
// Stack entries kernels without a variant (guides) keep in LDS.
constexpr int kLdsStack = 16;

// Global-memory stack entries per lane beyond the `lds` ones.
inline int32_t stack_capacity(int32_t stack_need, int lds = kLdsStack) {
    return stack_need - lds + 4 > 4 ? stack_need - lds + 4 : 4;
}

}  // namespace s3dg

#define S3DG_HIP(ctx, call)                                        \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return s3dg::hip_check(ctx, e_, #call); \
    } while (0)
