// post.hip -- tonemapping + sRGB capture (ToneMapper::Map, PathTracer::Capture)
// and the non-local-means denoiser (FilterNLM) on gfx950.
#include <hip/hip_runtime.h>
#include <float.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "internal.h"

using namespace s3dg;

namespace {

constexpr int kLumBlocks = 1024;
// Largest image the tonemap entry points take (pixel indices times 3 stay
// below 2^32).
constexpr uint32_t kMaxTonemapPixels = 1u << 28;
// Tonemap scratch: k_lum_partial's per-block log sums (double) and maxima
// (float), then the log-average and maximum luminance k_lum_final writes.
constexpr size_t kTonemapScratchBytes = kLumBlocks * (sizeof(double) + sizeof(float)) + 64;
constexpr float kLumEps = 0.00001f;   // epsilon passed by every operator (s3d_tonemapper.cpp:248,291,331)

// RGBToY with BT.601 weights (s3d_tonemapper.cpp:20,31-33)
__device__ __forceinline__ float rgb_to_y(float r, float g, float b) {
    return (0.29900f * r) + (0.58700f * g) + (0.11400f * b);
}

// Per-block partial sums of log(eps + Lw) in double and the maximum Lw
// (ComputeLogarithmicAverage, s3d_tonemapper.cpp:38-95).  The sum is kept in
// double and reduced in a fixed order, so it is deterministic.
__global__ __launch_bounds__(256) void k_lum_partial(const float *__restrict__ in, uint32_t n, double *psum,
                                                      float *pmax) {
    __shared__ double ssum[256];
    __shared__ float smax[256];
    double s = 0.0;
    float m = 0.0f;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float lw = rgb_to_y(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
        if (lw != lw) lw = kLumEps;
        if (lw < FLT_EPSILON) lw = kLumEps;
        m = fmaxf(m, lw);
        s += (double)logf(kLumEps + lw);
    }
    ssum[threadIdx.x] = s;
    smax[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        psum[blockIdx.x] = ssum[0];
        pmax[blockIdx.x] = smax[0];
    }
}

__global__ __launch_bounds__(256) void k_lum_final(const double *psum, const float *pmax, int nb, uint32_t n,
                                                    float *out /* [0] aveLw, [1] maxLw */) {
    __shared__ double ssum[256];
    __shared__ float smax[256];
    double s = 0.0;
    float m = 0.0f;
    for (int i = threadIdx.x; i < nb; i += 256) {
        s += psum[i];
        m = fmaxf(m, pmax[i]);
    }
    ssum[threadIdx.x] = s;
    smax[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float ave = (float)ssum[0];
        if (ave != ave) ave = kLumEps;
        ave /= (float)n;
        out[0] = expf(ave);
        out[1] = smax[0];
    }
}

__device__ __forceinline__ float saturate(float v) {
    const float lo = (1.0f < v) ? 1.0f : v;   // s3d::Min(maxi, value)
    return (0.0f > lo) ? 0.0f : lo;           // s3d::Max(mini, ...)
}
// ACESFilm (s3d_tonemapper.cpp:117-130)
__device__ __forceinline__ float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f, f = 0.665406f, g = 12.0f;
    return saturate((x * (a * x * f / g + b)) / (x * f / g * (c * x * f + d) + e));
}
// Uncharted2Tonemap (s3d_tonemapper.cpp:100-112)
__device__ __forceinline__ float uncharted2(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.01f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
// sRGB encode and 8-bit quantisation of PathTracer::Capture (s3d_pt.cpp:150-166)
__device__ __forceinline__ uint8_t capture8(float r) {
    if (r > 1.0f) r = 1.0f;
    if (r < 0.0f) r = 0.0f;
    r = (r <= 0.0031308f) ? 12.92f * r : powf(1.055f * r, 1.0f / 2.4f) - 0.055f;
    return (uint8_t)(r * 255.0f + 0.5f);
}

__global__ __launch_bounds__(256) void k_tonemap(int32_t type, const float *__restrict__ in, uint32_t n,
                                                  const float *lum, float *out, uint8_t *out8) {
    const float ave = lum[0], maxl = lum[1];
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float c[3] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
        if (type == S3DG_TONEMAP_ACES) {
            const float coeff = 0.27f / ave;
            for (int k = 0; k < 3; ++k) c[k] = aces(c[k] * coeff);
        } else if (type == S3DG_TONEMAP_UNCHARTED2) {
            const float coeff = 0.18f / ave;
            const float white = uncharted2(11.2f);
            for (int k = 0; k < 3; ++k) c[k] = uncharted2(2.0f * (c[k] * coeff)) / white;
        } else {
            const float coeff = 0.18f / ave;
            float m2 = maxl * coeff;
            m2 *= m2;
            for (int k = 0; k < 3; ++k) {
                const float L = c[k] * coeff;
                c[k] = L * (1.0f + (L / m2)) / (1.0f + L);
            }
        }
        if (out) {
            out[3 * i] = c[0];
            out[3 * i + 1] = c[1];
            out[3 * i + 2] = c[2];
        }
        if (out8) {
            out8[3 * i] = capture8(c[0]);
            out8[3 * i + 1] = capture8(c[1]);
            out8[3 * i + 2] = capture8(c[2]);
        }
    }
}

// ---------------------------------------------------------------------------
// Non-local means: FilterNLM (src/src/s3d_denoiser.cpp:19-95), search radius
// rs, patch radius rp, the source tile clamped to the image and staged in LDS.
// Patch radius 2 (the reference's) runs as box sums (k_nlm_box below); other
// patch radii walk the search window and the patch per pixel in the
// reference's order (x offset outer, y offset inner; RGBA distance summed as
// (xx+yy)+(zz+ww)), one thread per pixel (k_nlm).
// ---------------------------------------------------------------------------
constexpr int kTile = 16;
constexpr int kMaxHalo = 12;   // search radius + patch radius

__global__ __launch_bounds__(kTile *kTile) void k_nlm(int32_t W, int32_t H, float frac_h2, float sigma2, int32_t rs,
                                                      int32_t rp, const float4 *__restrict__ src,
                                                      float4 *__restrict__ dst) {
    __shared__ float4 tile[(kTile + 2 * kMaxHalo) * (kTile + 2 * kMaxHalo)];
    const int halo = rs + rp;
    const int tw = kTile + 2 * halo;
    const int x0 = blockIdx.x * kTile - halo, y0 = blockIdx.y * kTile - halo;
    for (int i = threadIdx.x; i < tw * tw; i += kTile * kTile) {
        const int tx = i % tw, ty = i / tw;
        const int gx = min(max(x0 + tx, 0), W - 1);
        const int gy = min(max(y0 + ty, 0), H - 1);
        tile[i] = src[(size_t)gy * W + gx];
    }
    __syncthreads();
    const int lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
    const int x = blockIdx.x * kTile + lx, y = blockIdx.y * kTile + ly;
    if (x >= W || y >= H) return;
    const int cx = lx + halo, cy = ly + halo;   // pixel position inside the tile
    float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float sw = 0.0f;
    for (int ox = -rs; ox <= rs; ++ox) {
        for (int oy = -rs; oy <= rs; ++oy) {
            float acc = 0.0f;
            for (int px = -rp; px <= rp; ++px) {
                for (int py = -rp; py <= rp; ++py) {
                    const float4 a = tile[(cy + py) * tw + (cx + px)];
                    const float4 b = tile[(cy + oy + py) * tw + (cx + ox + px)];
                    const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
                    acc += ((d0 * d0) + (d1 * d1)) + ((d2 * d2) + (d3 * d3));
                }
            }
            const float arg = -fmaxf(acc - 2.0f * sigma2, 0.0f) * frac_h2;
            const float w = expf(arg);
            sw += w;
            const float4 s = tile[(cy + oy) * tw + (cx + ox)];
            sum.x += s.x * w;
            sum.y += s.y * w;
            sum.z += s.z * w;
            sum.w += s.w * w;
        }
    }
    dst[(size_t)y * W + x] = make_float4(sum.x / sw, sum.y / sw, sum.z / sw, sum.w / sw);
}

// Non-local means for patch radius 2 (the reference's) by box sums.  For one
// search offset o the patch distance of pixel p is the sum, over the 5 x 5
// texels u around p, of D_o(u) = |I(u) - I(u + o)|^2 (RGBA, (xx+yy)+(zz+ww),
// coordinates clamped to the image as above), so each D_o(u) is formed once
// and shared by the 25 patches containing u instead of once per patch.  A
// wave's 64 lanes are 64 consecutive columns u (60 output pixels and two
// columns either side); the horizontal 5-sums come from the neighbouring
// lanes (whole-wave DPP shifts by one lane), the vertical ones from a lane's
// own rows: a band of 8 output rows takes 12 rows of D_o.  The four waves of
// a workgroup share one band (a 60 x 8 pixel tile) and take a quarter of the
// search offsets each, in the reference's order (x offset outer, y offset
// inner), and their weighted sums are added in wave order.  (A band per
// wave and all offsets per wave, 60 x 32 tiles, left a 1080p frame's 1,088
// equal workgroups 2.1 rounds of the 512 that fit at once; quarter-offset
// workgroups make 8.4 rounds: 0.64 -> 0.55 ms at r_s = 7.)  The patch
// distance is summed horizontally, then vertically, with fused multiply-adds,
// and the weights come from the hardware exponential, so results differ from
// the reference's by rounding only.
constexpr int kBoxLanes = 64, kBoxOutCols = kBoxLanes - 4, kBoxBand = 8, kBoxWaves = 4;

// v from the lane one below (CTRL 0x138, wave_shr:1) or one above (0x130,
// wave_shl:1); lanes with no source receive 0
template <int CTRL>
__device__ __forceinline__ float wave_shift1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// sum of v over lanes l - 2 .. l + 2 (exact only for lanes 2 .. 61)
__device__ __forceinline__ float lane_sum5(float v) {
    const float a1 = wave_shift1<0x138>(v), b1 = wave_shift1<0x130>(v);
    const float a2 = wave_shift1<0x138>(a1), b2 = wave_shift1<0x130>(b1);
    return ((a2 + a1) + v) + (b1 + b2);
}

template <int RS>
__global__ __launch_bounds__(kBoxLanes *kBoxWaves, 2) void k_nlm_box(int32_t W, int32_t H, float frac_h2,
                                                                     float sigma2, int32_t rs_arg,
                                                                     const float4 *__restrict__ src,
                                                                     float4 *__restrict__ dst) {
    constexpr int SPLIT = kBoxWaves, kOutRows = kBoxBand;
    constexpr int kRsMax = RS >= 0 ? RS : kMaxHalo - 2;
    __shared__ float4 tile[(kOutRows + 4 + 2 * kRsMax) * (kBoxLanes + 2 * kRsMax)];
    // the weighted sums of waves 1 .. SPLIT-1
    __shared__ float4 part_sum[(SPLIT - 1) * kBoxBand * kBoxLanes];
    __shared__ float part_sw[(SPLIT - 1) * kBoxBand * kBoxLanes];
    const int rs = RS >= 0 ? RS : rs_arg;
    const int tw = kBoxLanes + 2 * rs, th = kOutRows + 4 + 2 * rs;
    const int x0 = blockIdx.x * kBoxOutCols, y0 = blockIdx.y * kOutRows;
    // tile origin: two patch texels and the search radius before the first output pixel
    const int tx0 = x0 - 2 - rs, ty0 = y0 - 2 - rs;
    for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
        const int tx = i % tw, ty = i / tw;
        const int gx = min(max(tx0 + tx, 0), W - 1), gy = min(max(ty0 + ty, 0), H - 1);
        tile[i] = src[(size_t)gy * W + gx];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int part = wave;
    // this lane's column u = x0 - 2 + lane, the band's D rows y0 - 2 + r
    // (r = 0 .. 11): tile column lane + rs, tile rows rs + r
    const float4 *const base = tile + rs * tw + lane + rs;
    float4 c[kBoxBand + 4];
#pragma unroll
    for (int r = 0; r < kBoxBand + 4; ++r) c[r] = base[r * tw];
    float4 sum[kBoxBand];
    float sw[kBoxBand];
#pragma unroll
    for (int j = 0; j < kBoxBand; ++j) {
        sum[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        sw[j] = 0.0f;
    }
    // this wave's offsets: k = (ox + rs) (2 rs + 1) + (oy + rs) in [k0, k1)
    const int side = 2 * rs + 1, noff = side * side;
    const int k0 = part * noff / SPLIT, k1 = (part + 1) * noff / SPLIT;
    for (int k = k0; k < k1; ++k) {
        const int ox = k / side - rs, oy = k % side - rs;
        const float4 *const nb = base + oy * tw + ox;
        float hs[kBoxBand + 4];
        float4 s[kBoxBand];   // I(p + o) of the band's output pixels p
#pragma unroll
        for (int r = 0; r < kBoxBand + 4; ++r) {
            const float4 n = nb[r * tw];
            if (r >= 2 && r < kBoxBand + 2) s[r - 2] = n;
            const float d0 = c[r].x - n.x, d1 = c[r].y - n.y, d2 = c[r].z - n.z, d3 = c[r].w - n.w;
            hs[r] = lane_sum5(fmaf(d0, d0, d1 * d1) + fmaf(d2, d2, d3 * d3));
        }
#pragma unroll
        for (int j = 0; j < kBoxBand; ++j) {
            const float acc = (((hs[j] + hs[j + 1]) + hs[j + 2]) + hs[j + 3]) + hs[j + 4];
            // the hardware exponential: the weights lie in (0, 1], and the
            // centre offset's is exactly 1, so its relative error (a few
            // ulp) is all that reaches the result
            const float w = __expf(-fmaxf(acc - 2.0f * sigma2, 0.0f) * frac_h2);
            sw[j] += w;
            sum[j].x = fmaf(s[j].x, w, sum[j].x);
            sum[j].y = fmaf(s[j].y, w, sum[j].y);
            sum[j].z = fmaf(s[j].z, w, sum[j].z);
            sum[j].w = fmaf(s[j].w, w, sum[j].w);
        }
    }
    {
        if (part > 0) {
#pragma unroll
            for (int j = 0; j < kBoxBand; ++j) {
                const int e = ((part - 1) * kBoxBand + j) * kBoxLanes + lane;
                part_sum[e] = sum[j];
                part_sw[e] = sw[j];
            }
        }
        __syncthreads();
        if (part > 0) return;
#pragma unroll
        for (int q = 0; q < SPLIT - 1; ++q) {
#pragma unroll
            for (int j = 0; j < kBoxBand; ++j) {
                const int e = (q * kBoxBand + j) * kBoxLanes + lane;
                const float4 ps = part_sum[e];
                sum[j].x += ps.x;
                sum[j].y += ps.y;
                sum[j].z += ps.z;
                sum[j].w += ps.w;
                sw[j] += part_sw[e];
            }
        }
    }
    const int x = x0 + lane - 2;
    if (lane >= 2 && lane < kBoxLanes - 2 && x < W) {
#pragma unroll
        for (int j = 0; j < kBoxBand; ++j) {
            const int y = y0 + j;
            if (y < H)
                dst[(size_t)y * W + x] =
                    make_float4(sum[j].x / sw[j], sum[j].y / sw[j], sum[j].z / sw[j], sum[j].w / sw[j]);
        }
    }
}

template <int RS>
void launch_nlm_box(hipStream_t stream, int32_t w, int32_t h, float frac_h2, float sigma2, int32_t rs,
                    const float4 *s4, float4 *d4) {
    const dim3 grid((w + kBoxOutCols - 1) / kBoxOutCols, (h + kBoxBand - 1) / kBoxBand);
    hipLaunchKernelGGL((k_nlm_box<RS>), grid, dim3(kBoxLanes * kBoxWaves), 0, stream, w, h, frac_h2, sigma2, rs, s4,
                       d4);
}

// Checks a tonemap call's arguments; returns the pixel count through n.
int tonemap_args(s3dg_context *ctx, int32_t type, int32_t w, int32_t h, const float *d_in, uint32_t &n) {
    if (!ctx || !d_in || w <= 0 || h <= 0) return S3DG_ERR_INVALID;
    if (type < S3DG_TONEMAP_REINHARD || type > S3DG_TONEMAP_ACES)
        return set_error(ctx, S3DG_ERR_INVALID, "bad tonemap type");
    if ((uint64_t)w * (uint64_t)h > kMaxTonemapPixels)
        return set_error(ctx, S3DG_ERR_INVALID, "tonemap image larger than 2^28 pixels");
    n = (uint32_t)w * (uint32_t)h;
    return S3DG_OK;
}

// The three tonemap launches over a scratch area of kTonemapScratchBytes.
int launch_tonemap(s3dg_context *ctx, int32_t type, uint32_t n, const float *d_in, float *d_out, uint8_t *d_out8,
                   void *scratch_area) {
    const int nb = (int)std::min<uint32_t>(kLumBlocks, (n + 255) / 256);
    double *psum = (double *)scratch_area;
    float *pmax = (float *)(psum + nb);
    float *lum = pmax + nb;
    hipLaunchKernelGGL(k_lum_partial, dim3(nb), dim3(256), 0, ctx->stream, d_in, n, psum, pmax);
    hipLaunchKernelGGL(k_lum_final, dim3(1), dim3(256), 0, ctx->stream, psum, pmax, nb, n, lum);
    const uint32_t grid = std::min<uint32_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_tonemap, dim3(grid), dim3(256), 0, ctx->stream, type, d_in, n, lum, d_out, d_out8);
    S3DG_HIP(ctx, hipGetLastError());
    return S3DG_OK;
}

}  // namespace

// Device scratch of s3dg_tonemap_device_noalloc, one copy per GPU for the
// life of the process.  Host code never uses this name as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kTonemapScratchBytes];

namespace {
std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_addr;   // device address of `scratch`, by device
}  // namespace

namespace s3dg {

// Looks up (once per device, with that device current) the address of this
// device's copy of `scratch` and gives it to the context.
int tonemap_scratch_init(s3dg_context *ctx) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_addr.size() <= (size_t)ctx->device) g_scratch_addr.resize((size_t)ctx->device + 1, nullptr);
    if (!g_scratch_addr[ctx->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_check(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                    "hipGetSymbolAddress(scratch)");
        g_scratch_addr[ctx->device] = p;
    }
    ctx->noalloc_scratch = g_scratch_addr[ctx->device];
    return S3DG_OK;
}

}  // namespace s3dg

extern "C" int s3dg_tonemap_device(s3dg_context *ctx, int32_t type, int32_t w, int32_t h, const float *d_in,
                                   float *d_out, uint8_t *d_out8) {
    uint32_t n = 0;
    int rc = tonemap_args(ctx, type, w, h, d_in, n);
    if (rc) return rc;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    rc = ensure_buffer(ctx, &ctx->small, &ctx->small_bytes, kTonemapScratchBytes);
    if (rc) return rc;
    return launch_tonemap(ctx, type, n, d_in, d_out, d_out8, ctx->small);
}

extern "C" int s3dg_tonemap_device_noalloc(s3dg_context *ctx, int32_t type, int32_t w, int32_t h, const float *d_in,
                                           float *d_out, uint8_t *d_out8) {
    uint32_t n = 0;
    int rc = tonemap_args(ctx, type, w, h, d_in, n);
    if (rc) return rc;
    if (!ctx->noalloc_scratch) return set_error(ctx, S3DG_ERR_DEVICE, "device scratch address unknown");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    // every scratch word the kernels read is written first in the same call
    // (partials of blocks 0..nb-1, then the two results), so nothing is cleared
    return launch_tonemap(ctx, type, n, d_in, d_out, d_out8, ctx->noalloc_scratch);
}

extern "C" int s3dg_tonemap(s3dg_context *ctx, int32_t type, int32_t w, int32_t h, const float *in, float *out,
                            uint8_t *out8) {
    if (!ctx || !in || w <= 0 || h <= 0) return S3DG_ERR_INVALID;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    const size_t fb = n * 3 * sizeof(float), b8 = n * 3;
    int rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, 2 * fb + b8 + 256);
    if (rc) return rc;
    float *d_in = (float *)ctx->scratch;
    float *d_out = (float *)((char *)ctx->scratch + fb);
    uint8_t *d_8 = (uint8_t *)((char *)ctx->scratch + 2 * fb);
    S3DG_HIP(ctx, hipMemcpyAsync(d_in, in, fb, hipMemcpyHostToDevice, ctx->stream));
    rc = s3dg_tonemap_device(ctx, type, w, h, d_in, out ? d_out : nullptr, out8 ? d_8 : nullptr);
    if (rc) return rc;
    if (out) S3DG_HIP(ctx, hipMemcpyAsync(out, d_out, fb, hipMemcpyDeviceToHost, ctx->stream));
    if (out8) S3DG_HIP(ctx, hipMemcpyAsync(out8, d_8, b8, hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_nlm_device(s3dg_context *ctx, int32_t w, int32_t h, float coeff, int32_t rs, int32_t rp,
                               const float *d_src, float *d_dst) {
    if (!ctx || !d_src || !d_dst || w <= 0 || h <= 0) return S3DG_ERR_INVALID;
    if (d_src == d_dst) return set_error(ctx, S3DG_ERR_INVALID, "NLM source and destination alias");
    if (rs < 0 || rp < 0 || rs + rp > kMaxHalo)
        return set_error(ctx, S3DG_ERR_UNSUPPORTED, "NLM search + patch radius must be <= 12");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    // parameters exactly as FilterNLM derives them (s3d_denoiser.cpp:23-26)
    const float param_h = std::max(0.0001f, coeff);
    const float sigma = std::max(0.0001f, coeff);
    const float frac_h2 = 1.0f / (param_h * param_h);
    const float sigma2 = sigma * sigma;
    const float4 *s4 = (const float4 *)d_src;
    float4 *d4 = (float4 *)d_dst;
    if (rp == 2) {
        // box sums (rs <= 10 here); the reference's search radius (6) and the
        // denoiser comparison's (7) have their own instantiations
        if (rs == 6)
            launch_nlm_box<6>(ctx->stream, w, h, frac_h2, sigma2, rs, s4, d4);
        else if (rs == 7)
            launch_nlm_box<7>(ctx->stream, w, h, frac_h2, sigma2, rs, s4, d4);
        else
            launch_nlm_box<-1>(ctx->stream, w, h, frac_h2, sigma2, rs, s4, d4);
    } else {
        const dim3 grid((w + kTile - 1) / kTile, (h + kTile - 1) / kTile);
        hipLaunchKernelGGL(k_nlm, grid, dim3(kTile * kTile), 0, ctx->stream, w, h, frac_h2, sigma2, rs, rp, s4, d4);
    }
    S3DG_HIP(ctx, hipGetLastError());
    return S3DG_OK;
}

extern "C" int s3dg_nlm(s3dg_context *ctx, int32_t w, int32_t h, float coeff, int32_t rs, int32_t rp, const float *src,
                        float *dst) {
    if (!ctx || !src || !dst || w <= 0 || h <= 0) return S3DG_ERR_INVALID;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    const size_t b = (size_t)w * h * 4 * sizeof(float);
    int rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, 2 * b);
    if (rc) return rc;
    float *d_src = (float *)ctx->scratch;
    float *d_dst = (float *)((char *)ctx->scratch + b);
    S3DG_HIP(ctx, hipMemcpyAsync(d_src, src, b, hipMemcpyHostToDevice, ctx->stream));
    rc = s3dg_nlm_device(ctx, w, h, coeff, rs, rp, d_src, d_dst);
    if (rc) return rc;
    S3DG_HIP(ctx, hipMemcpyAsync(dst, d_dst, b, hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}
