// sceneset.hip -- several scenes merged into one set of device arrays, and
// the batched render over a list of frames of those scenes.
//
// No counterpart in the reference (it renders one scene per process).  The
// member scenes' nodes, triangles, shading records, top-level records,
// materials and lights are copied one after another into merged arrays; the
// indices stored inside them (child and triangle ranges of the nodes, tree
// roots, shape numbers, materials, light triangles) are moved by the member's
// offsets on the device.  Each member keeps a SceneRec (root, lights, camera,
// environment); a path carries its member's number, so one traversal launch
// serves rays of every member.  Within a member the order of shapes and
// triangles is unchanged, so ties resolve as in the member's own render and a
// frame rendered in a batch has the same sum as rendered alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "internal.h"

using namespace s3dg;

struct s3dg_scene_set {
    s3dg_context *ctx = nullptr;
    SceneDev dev{};
    std::vector<SceneRec> recs;
    void *d_nodes = nullptr, *d_tris = nullptr, *d_shade = nullptr, *d_insts = nullptr;
    void *d_mats = nullptr, *d_lights = nullptr, *d_shape_ref = nullptr, *d_recs = nullptr;
    void *d_rank = nullptr, *d_textures = nullptr, *d_texels = nullptr, *d_uv = nullptr;
    Workspace ws;
};

namespace {

__global__ void k_rebase_nodes(GNode *__restrict__ nodes, uint32_t n, uint32_t node_off, uint32_t tri_off,
                               uint32_t inst_off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        GNode &g = nodes[i];
        const uint32_t top = g.child_base & kTopLevelBit;
        if (g.imask) g.child_base = ((g.child_base & ~kTopLevelBit) + node_off) | top;
        g.leaf_base += top ? inst_off : tri_off;
    }
}

__global__ void k_rebase_insts(GInst *__restrict__ insts, uint32_t n, uint32_t node_off, int32_t shape_off,
                               int32_t mat_off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        GInst &g = insts[i];
        if (g.type != INST_SPHERE) g.blas_root += node_off;
        g.shape += shape_off;
        if (g.material >= 0) g.material += mat_off;
    }
}

__global__ void k_rebase_shade(GTriShade *__restrict__ shade, uint32_t n, int32_t mat_off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        shade[i].material += mat_off;
}

__global__ void k_rebase_lights(GLight *__restrict__ lights, uint32_t n, int32_t tri_off, int32_t inst_off,
                                int32_t shape_off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        GLight &g = lights[i];
        if (g.type == LIGHT_SPHERE) {
            g.tri[0] += inst_off;
        } else {
            for (int k = 0; k < g.count; ++k) g.tri[k] += tri_off;
        }
        g.shape += shape_off;
    }
}

__global__ void k_rebase_mats(GMat *__restrict__ mats, uint32_t n, int32_t tex_off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (mats[i].texture >= 0) mats[i].texture += tex_off;
}

__global__ void k_rebase_textures(GTexture *__restrict__ t, uint32_t n, uint64_t texel_off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        t[i].offset += texel_off;
}

__global__ void k_add_u32(uint32_t *__restrict__ a, uint32_t n, uint32_t off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] += off;
}

__global__ void k_rebase_shape_refs(GShapeRef *__restrict__ refs, uint32_t n, int32_t inst_off, int32_t mat_off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        refs[i].inst += inst_off;
        if (refs[i].material >= 0) refs[i].material += mat_off;
    }
}

unsigned grid_for(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 4096u)); }

void free_set(s3dg_scene_set *set) {
    void *ptrs[] = {set->d_nodes, set->d_tris,   set->d_shade,          set->d_insts, set->d_mats,
                    set->d_lights, set->d_shape_ref, set->d_recs, set->d_rank,
                    set->d_textures, set->d_texels, set->d_uv};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    workspace_free(set->ws);
    delete set;
}

}  // namespace

extern "C" int s3dg_scene_set_create(s3dg_context *ctx, uint32_t n, s3dg_scene *const *scenes,
                                     s3dg_scene_set **out) {
    if (!ctx || !scenes || !out || n == 0) return S3DG_ERR_INVALID;
    *out = nullptr;
    for (uint32_t k = 0; k < n; ++k)
        if (!scenes[k] || scenes[k]->ctx != ctx) return set_error(ctx, S3DG_ERR_INVALID, "scene of another context");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    s3dg_scene_set *set = new (std::nothrow) s3dg_scene_set();
    if (!set) return set_error(ctx, S3DG_ERR_NOMEM, "host allocation");
    set->ctx = ctx;
    // element offsets of every member
    struct Off {
        uint32_t nodes, tris, insts, mats, lights, shapes, textures;
        uint64_t texels;
    };
    std::vector<Off> off(n);
    Off tot{0, 0, 0, 0, 0, 0, 0, 0};
    bool any_uv = false;
    int32_t stack_need = 4;
    for (uint32_t k = 0; k < n; ++k) {
        const auto &c = scenes[k]->counts;
        off[k] = tot;
        tot.nodes += c.nodes;
        tot.tris += c.tris;
        tot.insts += c.insts;
        tot.mats += c.mats;
        tot.lights += c.lights;
        tot.shapes += c.shapes;
        tot.textures += c.textures;
        tot.texels += c.texel_floats;
        any_uv = any_uv || scenes[k]->d_uv != nullptr;
        stack_need = std::max(stack_need, scenes[k]->dev.stack_need);
    }
    auto alloc = [&](void **p, size_t bytes) -> hipError_t { return hipMalloc(p, std::max<size_t>(bytes, 256)); };
    hipError_t e = hipSuccess;
    if ((e = alloc(&set->d_nodes, (size_t)tot.nodes * sizeof(GNode))) != hipSuccess ||
        (e = alloc(&set->d_tris, (size_t)tot.tris * sizeof(GTri))) != hipSuccess ||
        (e = alloc(&set->d_shade, (size_t)tot.tris * sizeof(GTriShade))) != hipSuccess ||
        (e = alloc(&set->d_insts, (size_t)tot.insts * sizeof(GInst))) != hipSuccess ||
        (e = alloc(&set->d_mats, (size_t)tot.mats * sizeof(GMat))) != hipSuccess ||
        (e = alloc(&set->d_lights, (size_t)tot.lights * sizeof(GLight))) != hipSuccess ||
        (e = alloc(&set->d_shape_ref, (size_t)tot.shapes * sizeof(GShapeRef))) != hipSuccess ||
        (e = alloc(&set->d_recs, (size_t)n * sizeof(SceneRec))) != hipSuccess ||
        (e = alloc(&set->d_rank, (size_t)tot.tris * sizeof(uint32_t))) != hipSuccess ||
        (tot.textures && (e = alloc(&set->d_textures, (size_t)tot.textures * sizeof(GTexture))) != hipSuccess) ||
        (tot.texels && (e = alloc(&set->d_texels, (size_t)tot.texels * sizeof(float))) != hipSuccess) ||
        (any_uv && (e = alloc(&set->d_uv, (size_t)tot.tris * sizeof(GTriUV))) != hipSuccess)) {
        free_set(set);
        return hip_check(ctx, e, "hipMalloc(scene set)");
    }
    int rc = workspace_init(ctx, set->ws);
    if (rc) {
        free_set(set);
        return rc;
    }
    hipStream_t s = ctx->stream;
    set->recs.resize(n);
    for (uint32_t k = 0; k < n; ++k) {
        const s3dg_scene *sc = scenes[k];
        const auto &c = sc->counts;
        const Off &o = off[k];
        GNode *nodes = (GNode *)set->d_nodes + o.nodes;
        GTriShade *shade = (GTriShade *)set->d_shade + o.tris;
        GInst *insts = (GInst *)set->d_insts + o.insts;
        GLight *lights = (GLight *)set->d_lights + o.lights;
        uint32_t *rank = (uint32_t *)set->d_rank + o.tris;
        GShapeRef *refs = (GShapeRef *)set->d_shape_ref + o.shapes;
        struct {
            void *dst;
            const void *src;
            size_t bytes;
        } copies[] = {{nodes, sc->d_nodes, (size_t)c.nodes * sizeof(GNode)},
                      {(GTri *)set->d_tris + o.tris, sc->d_tris, (size_t)c.tris * sizeof(GTri)},
                      {shade, sc->d_shade, (size_t)c.tris * sizeof(GTriShade)},
                      {insts, sc->d_insts, (size_t)c.insts * sizeof(GInst)},
                      {(GMat *)set->d_mats + o.mats, sc->d_mats, (size_t)c.mats * sizeof(GMat)},
                      {lights, sc->d_lights, (size_t)c.lights * sizeof(GLight)},
                      {refs, sc->d_shape_ref, (size_t)c.shapes * sizeof(GShapeRef)},
                      {rank, sc->d_rank, (size_t)c.tris * sizeof(uint32_t)},
                      {(GTexture *)set->d_textures + o.textures, sc->d_textures, (size_t)c.textures * sizeof(GTexture)},
                      {(float *)set->d_texels + o.texels, sc->d_texels, (size_t)c.texel_floats * sizeof(float)},
                      {(GTriUV *)set->d_uv + o.tris, sc->d_uv, sc->d_uv ? (size_t)c.tris * sizeof(GTriUV) : 0}};
        for (auto &cp : copies)
            if (cp.bytes) {
                e = hipMemcpyAsync(cp.dst, cp.src, cp.bytes, hipMemcpyDeviceToDevice, s);
                if (e != hipSuccess) {
                    free_set(set);
                    return hip_check(ctx, e, "hipMemcpyAsync(scene set)");
                }
            }
        if (c.nodes)
            hipLaunchKernelGGL(k_rebase_nodes, dim3(grid_for(c.nodes)), dim3(256), 0, s, nodes, c.nodes, o.nodes,
                               o.tris, o.insts);
        if (c.insts)
            hipLaunchKernelGGL(k_rebase_insts, dim3(grid_for(c.insts)), dim3(256), 0, s, insts, c.insts, o.nodes,
                               (int32_t)o.shapes, (int32_t)o.mats);
        if (c.tris) {
            hipLaunchKernelGGL(k_rebase_shade, dim3(grid_for(c.tris)), dim3(256), 0, s, shade, c.tris,
                               (int32_t)o.mats);
            hipLaunchKernelGGL(k_add_u32, dim3(grid_for(c.tris)), dim3(256), 0, s, rank, c.tris, o.tris);
        }
        if (c.lights)
            hipLaunchKernelGGL(k_rebase_lights, dim3(grid_for(c.lights)), dim3(256), 0, s, lights, c.lights,
                               (int32_t)o.tris, (int32_t)o.insts, (int32_t)o.shapes);
        if (c.mats)
            hipLaunchKernelGGL(k_rebase_mats, dim3(grid_for(c.mats)), dim3(256), 0, s, (GMat *)set->d_mats + o.mats,
                               c.mats, (int32_t)o.textures);
        if (c.textures)
            hipLaunchKernelGGL(k_rebase_textures, dim3(grid_for(c.textures)), dim3(256), 0, s,
                               (GTexture *)set->d_textures + o.textures, c.textures, o.texels);
        if (c.shapes)
            hipLaunchKernelGGL(k_rebase_shape_refs, dim3(grid_for(c.shapes)), dim3(256), 0, s, refs, c.shapes,
                               (int32_t)o.insts, (int32_t)o.mats);
        SceneRec r = sc->rec;
        r.root += o.nodes;
        if (r.single_shape >= 0) r.single_shape += (int32_t)o.shapes;
        r.light_base = (int32_t)o.lights;
        if (r.env_tex >= 0) r.env_tex += (int32_t)o.textures;
        set->recs[k] = r;
    }
    e = hipMemcpyAsync(set->d_recs, set->recs.data(), n * sizeof(SceneRec), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free_set(set);
        return hip_check(ctx, e, "scene set merge");
    }
    SceneDev &dv = set->dev;
    dv.nodes = (const GNode *)set->d_nodes;
    dv.tris = (const GTri *)set->d_tris;
    dv.shade = (const GTriShade *)set->d_shade;
    dv.insts = (const GInst *)set->d_insts;
    dv.mats = (const GMat *)set->d_mats;
    dv.lights = (const GLight *)set->d_lights;
    dv.shape_ref = (const GShapeRef *)set->d_shape_ref;
    dv.recs = (const SceneRec *)set->d_recs;
    dv.tri_rank = (const uint32_t *)set->d_rank;
    dv.textures = (const GTexture *)set->d_textures;
    dv.texels = (const float *)set->d_texels;
    dv.tri_uv = (const GTriUV *)set->d_uv;
    dv.stack_need = stack_need;
    dv.tri_batch = ctx->tri_batch;
    *out = set;
    return S3DG_OK;
}

extern "C" int s3dg_scene_set_destroy(s3dg_scene_set *set) {
    if (!set) return S3DG_ERR_INVALID;
    (void)hipSetDevice(set->ctx->device);
    free_set(set);
    return S3DG_OK;
}

extern "C" int s3dg_render_batch_device(s3dg_scene_set *set, uint32_t njobs, const s3dg_frame_job *jobs,
                                        const s3dg_render_config *common, s3dg_render_stats *stats) {
    if (!set || !common || (njobs && !jobs)) return S3DG_ERR_INVALID;
    s3dg_context *ctx = set->ctx;
    if (common->max_bounce_count < 0) return set_error(ctx, S3DG_ERR_INVALID, "bad bounce count");
    if (njobs == 0) {
        // an empty batch renders nothing
        if (stats) memset(stats, 0, sizeof(*stats));
        return S3DG_OK;
    }
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<RenderJob> rj(njobs);
    for (uint32_t j = 0; j < njobs; ++j) {
        const s3dg_frame_job &f = jobs[j];
        if (f.scene >= set->recs.size()) return set_error(ctx, S3DG_ERR_INVALID, "frame names no scene of the set");
        rj[j].scene = f.scene;
        rj[j].width = f.width;
        rj[j].height = f.height;
        rj[j].spp = f.spp;
        rj[j].seed = f.seed;
        rj[j].first_pass = f.first_pass;
        rj[j].d_accum = f.d_accum_rgb;
    }
    return render_jobs(ctx, set->dev, set->ws, rj.data(), rj.size(), common->max_bounce_count,
                       common->rr_start_depth, common->flags, common->passes_in_flight, stats);
}
