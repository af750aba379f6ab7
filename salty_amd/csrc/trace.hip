// trace.hip -- context entry points and the batch traversal kernels
// (closest hit and any hit) behind s3dg_trace / s3dg_trace_device.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <new>

#include "internal.h"
#include "traverse.h"

using namespace s3dg;

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
extern "C" int s3dg_create(s3dg_context **out, int device) {
    if (!out) return S3DG_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return S3DG_ERR_NODEVICE;
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) return S3DG_ERR_NODEVICE;
    }
    if (device >= count) return S3DG_ERR_NODEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return S3DG_ERR_NODEVICE;
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) return S3DG_ERR_NODEVICE;
    s3dg_context *ctx = new (std::nothrow) s3dg_context();
    if (!ctx) return S3DG_ERR_NOMEM;
    ctx->device = device;
    // S3DG_REFILL (1..64) overrides the idle-lane count at which traversal waves take new rays
    if (const char *env = std::getenv("S3DG_REFILL")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v >= 1 && v <= 64) ctx->refill = (uint32_t)v;
    }
    // S3DG_FETCH_CHUNK (64..65536) overrides the queue entries a traversal wave takes at a time
    if (const char *env = std::getenv("S3DG_FETCH_CHUNK")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v >= 64 && v <= 65536) ctx->fetch_chunk = (uint32_t)v;
    }
    // S3DG_VARIANT (0..2) overrides the traversal kernel variant
    if (const char *env = std::getenv("S3DG_VARIANT")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v >= 0 && v < kVariantCount) ctx->variant = (uint32_t)v;
    }
    // S3DG_TAIL_PATHS overrides the queue length below which the remaining
    // bounces run thread per path (0: never)
    if (const char *env = std::getenv("S3DG_TAIL_PATHS")) {
        const long long v = std::strtoll(env, nullptr, 10);
        if (v >= 0) ctx->tail_paths = v > 0xffffffffll ? 0xffffffffu : (uint32_t)v;
    }
    // S3DG_LEAF_SIZE (0..8) overrides the leaf size of the traversal trees
    if (const char *env = std::getenv("S3DG_LEAF_SIZE")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v >= 0 && v <= 8) ctx->leaf_size = (uint32_t)v;
    }
    // S3DG_TRI_BATCH (1..65) overrides the pending-triangle lane count that keeps a step testing
    if (const char *env = std::getenv("S3DG_TRI_BATCH")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v >= 1 && v <= 65) ctx->tri_batch = (uint32_t)v;
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return S3DG_ERR_DEVICE;
    }
    if (tonemap_scratch_init(ctx) != S3DG_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return S3DG_ERR_DEVICE;
    }
    *out = ctx;
    return S3DG_OK;
}

extern "C" int s3dg_destroy(s3dg_context *ctx) {
    if (!ctx) return S3DG_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->small) (void)hipFree(ctx->small);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return S3DG_OK;
}

extern "C" const char *s3dg_last_error(const s3dg_context *ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" int s3dg_synchronize(s3dg_context *ctx) {
    if (!ctx) return S3DG_ERR_INVALID;
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_malloc(s3dg_context *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return S3DG_ERR_INVALID;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    S3DG_HIP(ctx, hipMalloc(out, bytes ? bytes : 1));
    return S3DG_OK;
}

extern "C" int s3dg_free(s3dg_context *ctx, void *p) {
    if (!ctx) return S3DG_ERR_INVALID;
    if (p) S3DG_HIP(ctx, hipFree(p));
    return S3DG_OK;
}

extern "C" int s3dg_memcpy_htod(s3dg_context *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return S3DG_ERR_INVALID;
    S3DG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_memcpy_dtoh(s3dg_context *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return S3DG_ERR_INVALID;
    S3DG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" void *s3dg_stream(s3dg_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// ---------------------------------------------------------------------------
// Batch traversal
// ---------------------------------------------------------------------------
namespace {

template <bool ANY, bool COUNT, int LDS, int WAVES>
__global__ __attribute__((amdgpu_flat_work_group_size(kBlock, kBlock), amdgpu_waves_per_eu(WAVES, WAVES))) void
k_trace(SceneDev S, const s3dg_ray *__restrict__ rays, uint32_t n, s3dg_hit *__restrict__ hits,
        int32_t *__restrict__ shapes, uint2 *gstack, int32_t gcap, int32_t *err, unsigned long long *counters,
        uint32_t *cursor, FetchParams fetch) {
    __shared__ uint2 lds[LDS * kBlock];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack<LDS> st;
    st.init(lds, threadIdx.x, gstack + gtid, stride, gcap, err);
    Tracer<ANY, COUNT> tr;
    tr.reset_counts();
    uint32_t i = 0;
    refill_loop(
        cursor, n, fetch,
        [&](uint32_t k) {
            i = k;
            const float4 a = reinterpret_cast<const float4 *>(rays + i)[0];
            const float4 b = reinterpret_cast<const float4 *>(rays + i)[1];
            DRay r;
            r.ox = a.x; r.oy = a.y; r.oz = a.z;
            r.dx = b.x; r.dy = b.y; r.dz = b.z;
            st.sp = 0;
            tr.start(S.recs[0], r, a.w, b.w);
            return true;
        },
        [&]() { return tr.step(S, st); },
        [&]() {
            int32_t prim = tr.best.prim;
            if (prim >= 0) prim = S.shade[prim].input_index;
            float4 o;
            o.x = __int_as_float(prim);
            o.y = tr.best.t;
            o.z = tr.best.u;
            o.w = tr.best.v;
            reinterpret_cast<float4 *>(hits)[i] = o;
            if (shapes) shapes[i] = tr.best.shape;
        });
    tr.flush(counters);
}

// Batch kernels of one traversal variant, [any][count].
struct BatchKernels {
    const void *k[2][2];
    int lds;
};

template <int V>
BatchKernels batch_kernels_of() {
    constexpr int L = kVariants[V].lds, W = kVariants[V].waves;
    return BatchKernels{{{reinterpret_cast<const void *>(&k_trace<false, false, L, W>),
                          reinterpret_cast<const void *>(&k_trace<false, true, L, W>)},
                         {reinterpret_cast<const void *>(&k_trace<true, false, L, W>),
                          reinterpret_cast<const void *>(&k_trace<true, true, L, W>)}},
                        L};
}

BatchKernels batch_kernels(uint32_t v) {
    static_assert(kVariantCount == 3, "one case per variant");
    switch (v) {
        case 0: return batch_kernels_of<0>();
        case 2: return batch_kernels_of<2>();
        default: return batch_kernels_of<1>();
    }
}

__global__ void k_clear_u32(uint32_t *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// Work cursor of the batch kernel: a word of the workspace's small status
// block (the error flag lives at its start).
uint32_t *trace_cursor(Workspace &ws) { return reinterpret_cast<uint32_t *>(ws.d_error) + 16; }

}  // namespace

namespace s3dg {

// Grid for persistent batch kernels: enough blocks to fill every CU several times.
uint32_t batch_grid(uint32_t n) {
    const uint32_t want = (n + kBlock - 1) / kBlock;
    return std::max(1u, std::min(want, 256u * 8u));
}

// Blocks of `kernel` (kBlock threads) that fit on the device at once: the
// grid of a persistent kernel.
int resident_grid(s3dg_context *ctx, const void *kernel, uint32_t *blocks) {
    int per_cu = 0, cus = 0;
    S3DG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0));
    S3DG_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    *blocks = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
    return S3DG_OK;
}

int workspace_init(s3dg_context *ctx, Workspace &ws) {
    S3DG_HIP(ctx, hipMalloc((void **)&ws.d_error, 256));
    S3DG_HIP(ctx, hipMemset(ws.d_error, 0, 256));
    return S3DG_OK;
}

void workspace_free(Workspace &ws) {
    if (ws.graph) (void)hipGraphExecDestroy(ws.graph);
    void *ptrs[] = {ws.d_gstack, ws.d_error, ws.d_render, ws.d_batch};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    ws = Workspace{};
}

int ensure_gstack(s3dg_context *ctx, Workspace &ws, uint32_t threads, int32_t stack_need, int lds) {
    const size_t need = (size_t)threads * stack_capacity(stack_need, lds) * sizeof(uint2);
    return ensure_buffer(ctx, &ws.d_gstack, &ws.gstack_bytes, need);
}

int check_error_flag(s3dg_context *ctx, Workspace &ws) {
    int32_t flag = 0;
    S3DG_HIP(ctx, hipMemcpyAsync(&flag, ws.d_error, sizeof(flag), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (flag) {
        (void)hipMemsetAsync(ws.d_error, 0, sizeof(int32_t), ctx->stream);
        return set_error(ctx, S3DG_ERR_UNSUPPORTED, "traversal stack overflow");
    }
    return S3DG_OK;
}

int launch_trace(s3dg_scene *sc, const s3dg_ray *d_rays, uint32_t n, int32_t mode, s3dg_hit *d_hits,
                 int32_t *d_shapes, unsigned long long *d_counters) {
    s3dg_context *ctx = sc->ctx;
    if (n == 0) return S3DG_OK;
    const BatchKernels K = batch_kernels(ctx->variant);
    const void *kern = K.k[mode == S3DG_TRACE_ANY ? 1 : 0][d_counters ? 1 : 0];
    uint32_t grid = 0;
    int rc = resident_grid(ctx, kern, &grid);
    if (rc) return rc;
    grid = std::min(grid, (n + kBlock - 1) / kBlock);
    rc = ensure_gstack(ctx, sc->ws, grid * kBlock, sc->dev.stack_need, K.lds);
    if (rc) return rc;
    uint2 *gs = (uint2 *)sc->ws.d_gstack;
    int32_t cap = stack_capacity(sc->dev.stack_need, K.lds);
    uint32_t *cursor = trace_cursor(sc->ws);
    FetchParams fetch{ctx->refill, ctx->fetch_chunk};
    SceneDev S = sc->dev;
    S.tri_batch = ctx->tri_batch;
    int32_t *d_err = sc->ws.d_error;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, ctx->stream, cursor);
    void *args[] = {&S, &d_rays, &n, &d_hits, &d_shapes, &gs, &cap, &d_err, &d_counters, &cursor, &fetch};
    S3DG_HIP(ctx, hipLaunchKernel(kern, dim3(grid), dim3(kBlock), args, 0, ctx->stream));
    S3DG_HIP(ctx, hipGetLastError());
    return S3DG_OK;
}

}  // namespace s3dg

extern "C" int s3dg_trace_device(s3dg_scene *sc, const s3dg_ray *d_rays, uint32_t n, int32_t mode,
                                 s3dg_hit *d_hits, int32_t *d_shapes) {
    if (!sc || (n && (!d_rays || !d_hits))) return S3DG_ERR_INVALID;
    if (mode != S3DG_TRACE_CLOSEST && mode != S3DG_TRACE_ANY) return set_error(sc->ctx, S3DG_ERR_INVALID, "bad mode");
    S3DG_HIP(sc->ctx, hipSetDevice(sc->ctx->device));
    int rc = launch_trace(sc, d_rays, n, mode, d_hits, d_shapes, nullptr);
    if (rc) return rc;
    return check_error_flag(sc->ctx, sc->ws);
}

// Variant that also counts node visits and triangle tests (counters[0], [1]),
// for the algorithmic-bytes figure of the roofline.
extern "C" int s3dg_trace_counted_device(s3dg_scene *sc, const s3dg_ray *d_rays, uint32_t n, int32_t mode,
                                         s3dg_hit *d_hits, uint64_t *d_counters) {
    if (!sc || !d_counters || (n && (!d_rays || !d_hits))) return S3DG_ERR_INVALID;
    S3DG_HIP(sc->ctx, hipSetDevice(sc->ctx->device));
    int rc = launch_trace(sc, d_rays, n, mode, d_hits, nullptr, (unsigned long long *)d_counters);
    if (rc) return rc;
    return check_error_flag(sc->ctx, sc->ws);
}

extern "C" int s3dg_trace(s3dg_scene *sc, const s3dg_ray *rays, uint32_t n, int32_t mode, s3dg_hit *hits,
                          int32_t *shapes) {
    if (!sc || (n && (!rays || !hits))) return S3DG_ERR_INVALID;
    if (n == 0) return S3DG_OK;
    s3dg_context *ctx = sc->ctx;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    const size_t rb = (size_t)n * sizeof(s3dg_ray), hb = (size_t)n * sizeof(s3dg_hit), sb = (size_t)n * 4;
    int rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, rb + hb + sb);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    s3dg_ray *d_rays = (s3dg_ray *)base;
    s3dg_hit *d_hits = (s3dg_hit *)(base + rb);
    int32_t *d_shapes = (int32_t *)(base + rb + hb);
    S3DG_HIP(ctx, hipMemcpyAsync(d_rays, rays, rb, hipMemcpyHostToDevice, ctx->stream));
    rc = s3dg_trace_device(sc, d_rays, n, mode, d_hits, shapes ? d_shapes : nullptr);
    if (rc) return rc;
    S3DG_HIP(ctx, hipMemcpyAsync(hits, d_hits, hb, hipMemcpyDeviceToHost, ctx->stream));
    if (shapes) S3DG_HIP(ctx, hipMemcpyAsync(shapes, d_shapes, sb, hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}
