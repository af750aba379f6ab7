// internal.h -- context and scene objects behind the C ABI (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/s3dg.h"
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
    // traversal kernels hand new rays to a wave once this many of its lanes are idle
    uint32_t refill = 32;
    // a traversal step keeps testing triangles while this many lanes have some pending
    uint32_t tri_batch = 16;
};

namespace s3dg {

// Device-side view of a scene, passed by value to kernels.
struct SceneDev {
    const GNode *nodes;
    const GTri *tris;
    const GTriShade *shade;
    const GInst *insts;
    const GMat *mats;
    const GLight *lights;
    const int32_t *inst_of_shape;   // top-level shape index -> instance record
    uint32_t root;          // top-level root node, or the mesh root when single_mesh
    int32_t single_mesh;    // 1: one untransformed mesh, traversal starts at its root
    int32_t single_shape;   // shape index of that mesh
    int32_t light_count;
    float env[3];
    int32_t stack_need;     // deepest stack the trees can require
    uint32_t tri_batch;     // lanes with pending triangles that keep a step testing
    GCamera cam;
};

struct SceneMeshInfo {
    uint32_t tri_offset;    // first depth-first triangle of the mesh in the global array
    uint32_t tri_count;
    uint32_t root;          // compressed root node
    std::vector<int32_t> order;   // depth-first -> input index
};

}  // namespace s3dg

struct s3dg_scene {
    s3dg_context *ctx = nullptr;
    s3dg::SceneDev dev{};
    std::vector<s3dg::SceneMeshInfo> meshes;
    s3dg_scene_stats stats{};
    // device allocations
    void *d_nodes = nullptr, *d_tris = nullptr, *d_shade = nullptr, *d_insts = nullptr;
    void *d_mats = nullptr, *d_lights = nullptr, *d_inst_of_shape = nullptr;
    // traversal stack overflow area (global memory), sized per launch
    void *d_gstack = nullptr;
    size_t gstack_bytes = 0;
    int32_t *d_error = nullptr;
    // render state, sized for the largest frame rendered so far
    void *d_render = nullptr;
    size_t render_bytes = 0;
};

namespace s3dg {

int set_error(s3dg_context *ctx, int code, const std::string &msg);
int hip_check(s3dg_context *ctx, hipError_t e, const char *what);
int ensure_buffer(s3dg_context *ctx, void **ptr, size_t *have, size_t need);

// Stack entries held in LDS per lane before spilling to global memory.
constexpr int kLdsStack = 16;
constexpr int kBlock = 256;

}  // namespace s3dg

#define S3DG_HIP(ctx, call)                                        \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return s3dg::hip_check(ctx, e_, #call); \
    } while (0)
