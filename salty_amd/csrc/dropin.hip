// dropin.hip -- the per-object calls of the reference's host interfaces,
// evaluated on the device for the C++ drop-in layer: IMaterial::Shade and
// GetBaseColor, ICamera::GetRay, Sphere::Sample, Triangle / Sphere::CalcParam
// and Scene::SampleIBL.  Each entry point takes host arrays of n queries,
// copies them into the context's scratch memory, runs one kernel (the device
// functions the renderer itself uses, so the results are the renderer's) and
// copies the answers back.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "internal.h"
#include "shading.h"

using namespace s3dg;

namespace {

constexpr uint32_t kMaxQueries = 1u << 24;

__global__ __launch_bounds__(kBlock) void k_shade_args(GMat m, const float *__restrict__ texels, GTexture tex,
                                                        uint32_t n, s3dg_shade_arg *args, int32_t base_only) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s3dg_shade_arg a = args[i];
    // TexturedMaterial: Shade and GetBaseColor times SampleColor(texcoord)
    // (src/src/s3d_texturedmaterial.cpp:63-87); 1 without a texture
    V3 tcol = v3(1.0f, 1.0f, 1.0f);
    if (texels) tcol = tex_sample(texels, tex, a.texcoord[0], a.texcoord[1]);
    V3 w;
    if (base_only) {
        w = base_color(m) * tcol;
    } else {
        Pcg rng;
        rng.state = a.random;
        V3 out = v3(0.0f, 0.0f, 0.0f);
        bool dice = false;
        w = shade(m, ld3(a.input), ld3(a.normal), rng, out, dice) * tcol;
        a.random = rng.state;
        a.output[0] = out.x; a.output[1] = out.y; a.output[2] = out.z;
        a.dice = dice ? 1 : 0;
    }
    a.weight[0] = w.x; a.weight[1] = w.y; a.weight[2] = w.z;
    args[i] = a;
}

__global__ __launch_bounds__(kBlock) void k_camera_rays(GCamera c, uint32_t n, const float *__restrict__ xy,
                                                         uint64_t *random, float *org_dir) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Pcg rng;
    rng.state = random ? random[i] : 0ull;
    V3 o, d;
    camera_ray(c, xy[2 * i], xy[2 * i + 1], rng, o, d);
    if (random) random[i] = rng.state;
    float *r = org_dir + 6 * (size_t)i;
    r[0] = o.x; r[1] = o.y; r[2] = o.z;
    r[3] = d.x; r[4] = d.y; r[5] = d.z;
}

__global__ __launch_bounds__(kBlock) void k_sphere_samples(float cx, float cy, float cz, float radius, uint32_t n,
                                                            uint64_t *random, float *pos, float *pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Pcg rng;
    rng.state = random[i];
    const V3 p = sphere_sample_point(v3(cx, cy, cz), radius, rng);
    random[i] = rng.state;
    pos[3 * i] = p.x; pos[3 * i + 1] = p.y; pos[3 * i + 2] = p.z;
    pdf[i] = sphere_sample_pdf(radius);
}

__global__ __launch_bounds__(kBlock) void k_calc_param(SceneDev S, uint32_t n, const int32_t *__restrict__ stored,
                                                        const s3dg_param_query *__restrict__ q, float *normal,
                                                        float *texcoord) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const s3dg_param_query a = q[i];
    V3 nn;
    float u = 0.0f, v = 0.0f;
    if (a.triangle == -2) {
        // Sphere::CalcParam (src/src/s3d_sphere.cpp:94-104)
        const GInst &in = S.insts[S.shape_ref[a.shape].inst];
        nn = unit3(ld3(a.position) - ld3(in.center));
        sphere_texcoord(nn, u, v);
    } else {
        // Triangle::CalcParam (src/src/s3d_triangle.cpp:107-121)
        const int32_t p = stored[i];
        nn = tri_normal(S.shade[p], a.barycentric[0], a.barycentric[1]);
        if (S.tri_uv) tri_texcoord(S.tri_uv[p], a.barycentric[0], a.barycentric[1], u, v);
    }
    normal[3 * i] = nn.x; normal[3 * i + 1] = nn.y; normal[3 * i + 2] = nn.z;
    texcoord[2 * i] = u; texcoord[2 * i + 1] = v;
}

__global__ __launch_bounds__(kBlock) void k_environment(SceneDev S, uint32_t n, const float *__restrict__ dirs,
                                                         float *rgb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SceneRec &rec = S.recs[0];
    V3 c;
    if (rec.env_tex < 0) c = ld3(rec.env);
    else c = tex_sample_dir(S.texels, S.textures[rec.env_tex], ld3(dirs + 3 * (size_t)i)) * 10.0f;
    rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z;
}

uint32_t blocks(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// Host-to-scratch staging of one call: regions are carved 256-byte aligned
// from the context's scratch buffer.
struct Staging {
    std::vector<std::pair<size_t, size_t>> parts;   // offset, bytes
    size_t total = 0;
    size_t add(size_t bytes) {
        const size_t off = total;
        parts.push_back({off, bytes});
        total += (bytes + 255) & ~(size_t)255;
        return off;
    }
};

int check_count(s3dg_context *ctx, uint32_t n) {
    if (n > kMaxQueries) return set_error(ctx, S3DG_ERR_INVALID, "more than 2^24 queries in one call");
    return S3DG_OK;
}

// Material (and its texture) for one call of the material entry points.
int material_call(s3dg_context *ctx, const s3dg_material *material, const s3dg_texture_desc *texture, uint32_t n,
                  s3dg_shade_arg *args, int32_t base_only) {
    if (!ctx || !material || (n && !args)) return S3DG_ERR_INVALID;
    int rc = check_count(ctx, n);
    if (rc || n == 0) return rc;
    if (material->type < S3DG_MAT_LAMBERT || material->type > S3DG_MAT_PLASTIC)
        return set_error(ctx, S3DG_ERR_INVALID, "unknown material type");
    if (texture && (texture->width <= 0 || texture->height <= 0 || !texture->rgb))
        return set_error(ctx, S3DG_ERR_INVALID, "bad texture");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    GMat m;
    make_material(*material, m);
    const size_t tex_floats = texture ? (size_t)texture->width * texture->height * 3 : 0;
    Staging st;
    const size_t o_args = st.add((size_t)n * sizeof(s3dg_shade_arg));
    const size_t o_tex = st.add(tex_floats * sizeof(float));
    rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, st.total);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    s3dg_shade_arg *d_args = (s3dg_shade_arg *)(base + o_args);
    float *d_tex = texture ? (float *)(base + o_tex) : nullptr;
    GTexture g{};
    if (texture) {
        g.offset = 0;
        g.width = texture->width;
        g.height = texture->height;
        S3DG_HIP(ctx, hipMemcpyAsync(d_tex, texture->rgb, tex_floats * sizeof(float), hipMemcpyHostToDevice,
                                     ctx->stream));
    }
    S3DG_HIP(ctx, hipMemcpyAsync(d_args, args, (size_t)n * sizeof(s3dg_shade_arg), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_shade_args, dim3(blocks(n)), dim3(kBlock), 0, ctx->stream, m, d_tex, g, n, d_args, base_only);
    S3DG_HIP(ctx, hipGetLastError());
    S3DG_HIP(ctx, hipMemcpyAsync(args, d_args, (size_t)n * sizeof(s3dg_shade_arg), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

}  // namespace

extern "C" int s3dg_shade(s3dg_context *ctx, const s3dg_material *material, const s3dg_texture_desc *texture,
                          uint32_t n, s3dg_shade_arg *args) {
    return material_call(ctx, material, texture, n, args, 0);
}

extern "C" int s3dg_base_color(s3dg_context *ctx, const s3dg_material *material, const s3dg_texture_desc *texture,
                               uint32_t n, s3dg_shade_arg *args) {
    return material_call(ctx, material, texture, n, args, 1);
}

extern "C" int s3dg_camera_rays(s3dg_context *ctx, const s3dg_camera_desc *camera, uint32_t n, const float *xy,
                                uint64_t *random, float *org_dir) {
    if (!ctx || !camera || (n && (!xy || !org_dir))) return S3DG_ERR_INVALID;
    int rc = check_count(ctx, n);
    if (rc || n == 0) return rc;
    if (camera->type == S3DG_CAMERA_THINLENS && camera->lens_radius > 0.0f && !random)
        return set_error(ctx, S3DG_ERR_INVALID, "a thin-lens camera needs random states");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    const GCamera g = make_camera(*camera);
    Staging st;
    const size_t o_xy = st.add((size_t)n * 2 * sizeof(float));
    const size_t o_rng = st.add(random ? (size_t)n * sizeof(uint64_t) : 0);
    const size_t o_out = st.add((size_t)n * 6 * sizeof(float));
    rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, st.total);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    float *d_xy = (float *)(base + o_xy);
    uint64_t *d_rng = random ? (uint64_t *)(base + o_rng) : nullptr;
    float *d_out = (float *)(base + o_out);
    S3DG_HIP(ctx, hipMemcpyAsync(d_xy, xy, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    if (random)
        S3DG_HIP(ctx, hipMemcpyAsync(d_rng, random, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_camera_rays, dim3(blocks(n)), dim3(kBlock), 0, ctx->stream, g, n, d_xy, d_rng, d_out);
    S3DG_HIP(ctx, hipGetLastError());
    S3DG_HIP(ctx, hipMemcpyAsync(org_dir, d_out, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (random)
        S3DG_HIP(ctx, hipMemcpyAsync(random, d_rng, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_sphere_sample(s3dg_context *ctx, const float center[3], float radius, uint32_t n,
                                  uint64_t *random, float *pos, float *pdf) {
    if (!ctx || !center || (n && (!random || !pos || !pdf)) || !(radius > 0.0f)) return S3DG_ERR_INVALID;
    int rc = check_count(ctx, n);
    if (rc || n == 0) return rc;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    Staging st;
    const size_t o_rng = st.add((size_t)n * sizeof(uint64_t));
    const size_t o_pos = st.add((size_t)n * 3 * sizeof(float));
    const size_t o_pdf = st.add((size_t)n * sizeof(float));
    rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, st.total);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    uint64_t *d_rng = (uint64_t *)(base + o_rng);
    float *d_pos = (float *)(base + o_pos), *d_pdf = (float *)(base + o_pdf);
    S3DG_HIP(ctx, hipMemcpyAsync(d_rng, random, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_sphere_samples, dim3(blocks(n)), dim3(kBlock), 0, ctx->stream, center[0], center[1],
                       center[2], radius, n, d_rng, d_pos, d_pdf);
    S3DG_HIP(ctx, hipGetLastError());
    S3DG_HIP(ctx, hipMemcpyAsync(random, d_rng, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipMemcpyAsync(pos, d_pos, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipMemcpyAsync(pdf, d_pdf, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_scene_calc_param(s3dg_scene *sc, uint32_t n, const s3dg_param_query *queries, float *normal,
                                     float *texcoord) {
    if (!sc || (n && (!queries || !normal || !texcoord))) return S3DG_ERR_INVALID;
    s3dg_context *ctx = sc->ctx;
    int rc = check_count(ctx, n);
    if (rc || n == 0) return rc;
    // stored position of each queried triangle (host tables of the scene)
    std::vector<int32_t> stored(n, -1);
    for (uint32_t i = 0; i < n; ++i) {
        const s3dg_param_query &q = queries[i];
        if (q.shape < 0 || (size_t)q.shape >= sc->host.shape_mesh.size())
            return set_error(ctx, S3DG_ERR_INVALID, "query names no shape of the scene");
        const int32_t mesh = sc->host.shape_mesh[q.shape];
        if (q.triangle == -2) {
            if (mesh >= 0) return set_error(ctx, S3DG_ERR_INVALID, "sphere query on a mesh shape");
            continue;
        }
        if (mesh < 0) return set_error(ctx, S3DG_ERR_INVALID, "triangle query on a sphere");
        const SceneMeshInfo &mi = sc->meshes[mesh];
        if (q.triangle < 0 || (uint32_t)q.triangle >= mi.tri_count)
            return set_error(ctx, S3DG_ERR_INVALID, "triangle index out of range");
        stored[i] = (int32_t)(mi.tri_offset + (uint32_t)mi.stored_of_input[q.triangle]);
    }
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    Staging st;
    const size_t o_q = st.add((size_t)n * sizeof(s3dg_param_query));
    const size_t o_s = st.add((size_t)n * sizeof(int32_t));
    const size_t o_n = st.add((size_t)n * 3 * sizeof(float));
    const size_t o_t = st.add((size_t)n * 2 * sizeof(float));
    rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, st.total);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    s3dg_param_query *d_q = (s3dg_param_query *)(base + o_q);
    int32_t *d_s = (int32_t *)(base + o_s);
    float *d_n = (float *)(base + o_n), *d_t = (float *)(base + o_t);
    S3DG_HIP(ctx, hipMemcpyAsync(d_q, queries, (size_t)n * sizeof(s3dg_param_query), hipMemcpyHostToDevice,
                                 ctx->stream));
    S3DG_HIP(ctx, hipMemcpyAsync(d_s, stored.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_calc_param, dim3(blocks(n)), dim3(kBlock), 0, ctx->stream, sc->dev, n, d_s, d_q, d_n, d_t);
    S3DG_HIP(ctx, hipGetLastError());
    S3DG_HIP(ctx, hipMemcpyAsync(normal, d_n, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipMemcpyAsync(texcoord, d_t, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_scene_environment(s3dg_scene *sc, uint32_t n, const float *dirs, float *rgb) {
    if (!sc || (n && (!dirs || !rgb))) return S3DG_ERR_INVALID;
    s3dg_context *ctx = sc->ctx;
    int rc = check_count(ctx, n);
    if (rc || n == 0) return rc;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    Staging st;
    const size_t o_d = st.add((size_t)n * 3 * sizeof(float));
    const size_t o_c = st.add((size_t)n * 3 * sizeof(float));
    rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, st.total);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    float *d_d = (float *)(base + o_d), *d_c = (float *)(base + o_c);
    S3DG_HIP(ctx, hipMemcpyAsync(d_d, dirs, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_environment, dim3(blocks(n)), dim3(kBlock), 0, ctx->stream, sc->dev, n, d_d, d_c);
    S3DG_HIP(ctx, hipGetLastError());
    S3DG_HIP(ctx, hipMemcpyAsync(rgb, d_c, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}
