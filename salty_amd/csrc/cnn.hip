// cnn.hip -- albedo/normal/depth-guided U-Net denoiser with hand-written BF16
// MFMA kernels (gfx950 v_mfma_f32_16x16x32_bf16, FP32 accumulation) for the
// forward pass, the backward pass (data and weight gradients) and Adam.
//
// Not in the reference (it has only FilterNLM); this is the learned stand-in
// for NLM.  Network: 4 levels, 3x3 convolutions with 32/64/96/128 channels,
// ReLU, 2x2 average pooling, nearest upsampling, concatenated skips, and a
// 1x1 head to log radiance.  Activations are NHWC bf16; weights keep an FP32
// master copy and two bf16 copies: [tap][cout][cin] for the forward pass and
// the flipped transpose [tap][cin][cout] for the data gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <type_traits>
#include <vector>

#include "internal.h"

using namespace s3dg;

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

namespace {

constexpr int kInChannels = 10;     // radiance 3, albedo 3, normal 3, depth 1
constexpr int kInPadded = 32;       // first layer input padded to one MFMA K step
constexpr int kOut = 3;

__device__ __forceinline__ float b2f(bf16 x) { return (float)x; }

// log(1 + x) for the input features (x >= 0), by Goldberg's method: with
// y = 1 + x rounded, log(y) * x / (y - 1) cancels the rounding of y (a few
// ulps, against the library log1pf's ~25 instructions; the packing kernel
// and the first convolution's own packing both use it, so they agree bit for
// bit, and bf16 rounding of the result leaves a few ulps of FP32 unseen).
__device__ __forceinline__ float log1p_feature(float x) {
    const float y = 1.0f + x;
    const float d = y - 1.0f;
    return d == 0.0f ? x : __logf(y) * __fdividef(x, d);
}
__device__ __forceinline__ bf16 f2b(float x) { return (bf16)x; }

// ---------------------------------------------------------------------------
// Virtual convolution input: concat(upsample_s1(in1), in2) along channels.
// in1: [N, H/s1, W/s1, C1], in2: [N, H, W, C2] (may be null with C2 = 0);
// C1 and C2 are multiples of 32, so that a 32-channel slice lies in one source.
// ---------------------------------------------------------------------------
struct ConvIn {
    const bf16 *in1;
    const bf16 *in2;
    int32_t C1, C2, s1;   // s1 is 1 or 2
    // the first layer at inference: the 10 FP32 input features per pixel
    // (8-byte aligned), packed while staged as k_pack_input packs them
    const float *feat;
};

// Workgroup barrier for LDS only: this wave's LDS operations complete
// (lgkmcnt(0); gfx9 s_waitcnt fields: vmcnt 63, expcnt 7, lgkmcnt 0), then
// s_barrier.  __syncthreads() also fences global memory, i.e. waits for every
// outstanding global load and store of the wave (vmcnt(0)), which would make
// each barrier wait for the next step's prefetched loads and the previous
// step's output stores.  No wave here reads global memory another wave of the
// workgroup writes, so the LDS ordering is all a barrier needs.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
}

// A wave-uniform pointer as a value held in scalar registers.
__device__ __forceinline__ const bf16 *uniform_ptr(const bf16 *p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<const bf16 *>(((uint64_t)hi << 32) | lo);
}

// 8 consecutive channels c..c+7 (all in the same source) at pixel (n, y, x),
// without the zeroing outside the image: `inside` says whether the value is
// the pixel's (outside, an in-image address is read).  One load from whichever
// source holds the channels: with a load on each side of a branch the
// compiler has to wait for the first before the second may write the same
// registers, which serialised the staging loads.
__device__ __forceinline__ bf16x8 load8_raw(const ConvIn &I, int n, int y, int x, int H, int W, int c,
                                            bool &inside) {
    inside = y >= 0 && y < H && x >= 0 && x < W;
    const int yy = inside ? y : 0, xx = inside ? x : 0;
    const bool up = c < I.C1;
    // upsampled input: coordinates halved when s1 = 2
    const int sh = I.s1 >> 1;
    const int h1 = (H + sh) >> sh, w1 = (W + sh) >> sh;
    const size_t off = up ? (((size_t)n * h1 + (yy >> sh)) * w1 + (xx >> sh)) * I.C1 + c
                          : (((size_t)n * H + yy) * W + xx) * I.C2 + (c - I.C1);
    // the two source pointers as values in scalar registers: selecting
    // between the struct's fields became a per-lane load of the pointer,
    // with a wait for every load before it
    const bf16 *base = up ? uniform_ptr(I.in1) : uniform_ptr(I.in2);
    // a global-memory load (through a generic pointer it became a flat load,
    // whose wait also waits for LDS and scalar traffic)
    typedef __attribute__((address_space(1))) const bf16x8 gbf16x8;
    return *(gbf16x8 *)(base + off);
}

// The same, zero outside the image.
__device__ __forceinline__ bf16x8 load8(const ConvIn &I, int n, int y, int x, int H, int W, int c) {
    bool inside;
    bf16x8 v = load8_raw(I, n, y, x, H, W, c, inside);
    if (!inside) {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = f2b(0.0f);
    }
    return v;
}

__device__ __forceinline__ float load1(const ConvIn &I, int n, int y, int x, int H, int W, int c) {
    if (y < 0 || y >= H || x < 0 || x >= W) return 0.0f;
    if (c < I.C1) {
        const int h1 = (H + I.s1 - 1) / I.s1, w1 = (W + I.s1 - 1) / I.s1;
        return b2f(I.in1[(((size_t)n * h1 + y / I.s1) * w1 + x / I.s1) * I.C1 + c]);
    }
    return b2f(I.in2[(((size_t)n * H + y) * W + x) * I.C2 + (c - I.C1)]);
}

// ---------------------------------------------------------------------------
// 3x3 convolution (stride 1, zero padding) as an implicit GEMM: M = pixels,
// N = output channels, K = 9 taps x Cin, on v_mfma_f32_16x16x32_bf16 with one
// 16-pixel row of the tile as M and a whole 32-channel input slice as K.  A
// workgroup (four waves) computes a ROWS x 16 pixel tile for COT output
// channels (ROWS x COT = 1024); each wave owns 8 rows x 32 channels.  Per
// input slice it stages the (ROWS + 2) x 18 halo and the 9 x COT x 32
// weights in LDS (pixel and channel rows padded from 64 to 80 bytes, so that
// the 16-byte reads of consecutive lanes hit distinct banks).  For each column
// shift dx a wave reads the 3 x 2 weight operands once and, per half of its
// rows, the 6 input rows it needs, and runs the 3 x 4 x 2 MFMAs they make
// (each input row serves up to three taps): 18 LDS reads per 48 MFMAs.  The
// grid is persistent and each step's global loads (one tile and 32-channel
// slice) are in flight while the previous step computes; with one slice and
// one output-channel block (SINGLE) the weights are staged once per
// workgroup and the halos of the next two tiles are in flight, in two
// register stages used in turn.  Used for the forward pass
// (weights [tap][cout][cin]) and for the data gradient (flipped transposed
// weights, dY as input; with `mask`, the result passes the ReLU whose output
// mask is).  The weights are the MFMA's first operand, so each lane ends with
// 8 consecutive output channels of one pixel (with the rows of its two
// channel tiles mapped by chan_of) and stores them directly, 16 bytes.
// ---------------------------------------------------------------------------
constexpr int kTile = 16;                 // tile width (pixels)
constexpr int kHalo = kTile + 2;
constexpr int kRow = 40;                  // LDS row: 32 channels + 8 bf16 padding

// Output channel (within a wave's 32) of row m of channel tile t: rows
// 4g .. 4g+3 of tile 0 and of tile 1 are channels 8g .. 8g+7, so that the lane
// holding rows 4g .. 4g+3 of both tiles holds 8 consecutive channels.
__device__ __forceinline__ int chan_of(int t, int m) { return 8 * (m >> 2) + 4 * t + (m & 3); }

// The 1x1 inference head fused into the last convolution (Cout = 32): with
// `out` set, that convolution's output is not stored; each pixel's 32
// channels (bf16-rounded, as stored) go through the head instead, and the
// prediction expm1(head), clamped at 0, or the input radiance where the camera
// ray missed (depth guide 0), is written as 3 floats.  The sums of the
// training head (k_head), in another order.
struct HeadArgs {
    const float *w;      // [3][32]
    const float *b;      // [3]
    const float *feat;   // [pixels][10]
    float *out;          // [pixels][3], or null: no head
};

// Epilogue of a convolution: bias + ReLU (+ 2x2 pooling), the fused head, the
// data gradient's ReLU mask, or the plain data gradient; one kernel
// instantiation each, so that each carries only its own registers.
enum { EPI_FORWARD = 0, EPI_HEAD = 1, EPI_MASK = 2, EPI_LINEAR = 3 };

template <int COT, int EPI, bool FEAT, bool SINGLE>
__global__ __launch_bounds__(256, 2) void k_conv3x3(ConvIn I, int32_t N, int32_t H, int32_t W,
                                                    const bf16 *__restrict__ wt, const float *__restrict__ bias,
                                                    int32_t Cout, const bf16 *__restrict__ mask,
                                                    bf16 *__restrict__ out, HeadArgs head, bf16 *__restrict__ pool) {
    // bias and ReLU in the forward pass (and the head), neither in the data gradient
    constexpr bool kAct = EPI == EPI_FORWARD || EPI == EPI_HEAD;
    constexpr int ROWS = 1024 / COT;                 // tile rows: 32 (COT 32) or 16 (COT 64)
    constexpr int WR = ROWS / 8;                     // waves along the rows
    constexpr int kInElems = (ROWS + 2) * kHalo * kRow, kWElems = 9 * COT * kRow;
    __shared__ __attribute__((aligned(16))) bf16 smem[kInElems + kWElems];
    bf16 *const s_in = smem;
    bf16 *const s_w = smem + kInElems;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (W + kTile - 1) / kTile, tiles_y = (H + ROWS - 1) / ROWS;
    const int co_blocks = (Cout + COT - 1) / COT;
    const int items = N * tiles_x * tiles_y * co_blocks;   // (tile, output-channel block) pairs
    const int Cin = I.C1 + I.C2;
    const int rw0 = (wave % WR) * 8;                 // this wave's first tile row
    const int cw0 = (wave / WR) * 32;                // ... and first channel of the block
    const int l16 = lane & 15, kq = 8 * (lane >> 4);
    bf16x8 zero;
#pragma unroll
    for (int j = 0; j < 8; ++j) zero[j] = f2b(0.0f);
    // work item -> image, tile origin, first output channel (output-channel
    // blocks innermost, so that consecutive items share the input tile)
    struct Item {
        int n, y0, x0, co0;
    };
    const auto item_of = [&](int it) {
        Item r;
        r.co0 = (it % co_blocks) * COT;
        const int t = it / co_blocks;
        r.x0 = (t % tiles_x) * kTile;
        r.y0 = ((t / tiles_x) % tiles_y) * ROWS;
        r.n = t / (tiles_x * tiles_y);
        return r;
    };
    // staging: all global loads of a step (one item, one 32-channel slice)
    // are in flight while the previous step computes; SINGLE (one slice, one
    // output-channel block: the weights never change) keeps the weights in
    // LDS for the whole launch and two items' halo loads in flight
    constexpr int kHalfCh = (ROWS + 2) * kHalo * 4, kHalfPer = (kHalfCh + 255) / 256;
    constexpr int kWCh = 9 * COT * 4, kWPer = (kWCh + 255) / 256;
    // FEAT: one halo pixel (10 floats) per thread and round instead
    constexpr int kFeatPix = (ROWS + 2) * kHalo, kFeatPer = (kFeatPix + 255) / 256;
    struct Stage {
        bf16x8 hv[FEAT ? 1 : kHalfPer];
        float fv[FEAT ? kFeatPer : 1][kInChannels];
        bf16x8 wv[kWPer];
        uint32_t hin, win;   // chunks whose value is real (inside the image / an existing channel)
    };
    // every chunk load is issued unconditionally and raw (the last round's
    // spare threads load the last chunk again and do not store it;
    // out-of-range chunks are zeroed when stored), so that nothing between
    // the loads and the next step's LDS stores waits for them
    const auto issue_weights = [&](Stage &sg, const Item &it, int c0) {
        typedef __attribute__((address_space(1))) const bf16x8 gbf16x8;
        sg.win = 0;
#pragma unroll
        for (int k = 0; k < kWPer; ++k) {
            const int e = min((int)threadIdx.x + 256 * k, kWCh - 1);
            const int row = e >> 2, q = e & 3;
            const int tap = row / COT, co = it.co0 + row % COT;
            const uint32_t off = (((uint32_t)tap * (uint32_t)Cout + (uint32_t)min(co, Cout - 1)) * (uint32_t)Cin +
                                  (uint32_t)q * 8u) * 2u;
            sg.wv[k] = *(const gbf16x8 *)((const char *)uniform_ptr(wt + c0) + off);
            sg.win |= (uint32_t)(co < Cout) << k;
        }
    };
    const auto issue_halo = [&](Stage &sg, const Item &it, int c0) {
        sg.hin = 0;
        if constexpr (FEAT) {
            typedef float f32x2 __attribute__((ext_vector_type(2)));
            typedef __attribute__((address_space(1))) const f32x2 gfloat2;
            const float *feat = reinterpret_cast<const float *>(uniform_ptr(reinterpret_cast<const bf16 *>(I.feat)));
#pragma unroll
            for (int k = 0; k < kFeatPer; ++k) {
                const int pix = min((int)threadIdx.x + 256 * k, kFeatPix - 1);
                const int y = it.y0 + pix / kHalo - 1, x = it.x0 + pix % kHalo - 1;
                const bool inside = y >= 0 && y < H && x >= 0 && x < W;
                const uint32_t sp = ((uint32_t)it.n * (uint32_t)H + (uint32_t)(inside ? y : 0)) * (uint32_t)W +
                                    (uint32_t)(inside ? x : 0);
                const gfloat2 *src = (const gfloat2 *)(feat + (size_t)sp * kInChannels);
#pragma unroll
                for (int j = 0; j < kInChannels / 2; ++j) {
                    const f32x2 v = src[j];
                    sg.fv[k][2 * j] = v.x;
                    sg.fv[k][2 * j + 1] = v.y;
                }
                sg.hin |= (uint32_t)inside << k;
            }
        } else {
            // the slice lies in one source (C1 is a multiple of 32), chosen
            // once per step; pixel indices in 32 bits (N*H*W <= 2^27)
            const bool up = c0 < I.C1;
            const bf16 *src = up ? uniform_ptr(I.in1) : uniform_ptr(I.in2);
            const uint32_t C = (uint32_t)(up ? I.C1 : I.C2);
            const int cs = up ? c0 : c0 - I.C1;
            const int sh = up ? (I.s1 >> 1) : 0;   // upsampled input: coordinates halved
            const uint32_t hs = (uint32_t)(H + sh) >> sh, ws = (uint32_t)(W + sh) >> sh;
            typedef __attribute__((address_space(1))) const bf16x8 gbf16x8;
            // addresses: one uniform base, the source pixel left of and above
            // the tile origin's (it may lie before the tensor; nothing reads
            // it), plus a 32-bit byte offset per chunk.  Tile origins are
            // even, so a chunk's offset from the base is the same for every
            // tile; a chunk outside the image reads the tile origin's pixel
            // instead (in the image) and is zeroed when stored.
            const int sy0 = it.y0 >> sh, sx0 = it.x0 >> sh;
            // signed 64-bit throughout: the base pixel's row and column can be -1
            const int64_t bpix = ((int64_t)it.n * (int64_t)hs + sy0 - 1) * (int64_t)ws + sx0 - 1;
            const uint64_t hbase = (uint64_t)(uintptr_t)src + (uint64_t)((bpix * (int64_t)C + cs) * 2);
            const char *hb = (const char *)(uintptr_t)hbase;
            const uint32_t origin_off = ((ws + 1u) * C) * 2u;
#pragma unroll
            for (int k = 0; k < kHalfPer; ++k) {
                const int e = min((int)threadIdx.x + 256 * k, kHalfCh - 1);
                const int pix = e >> 2, q = e & 3;
                const int dy = pix / kHalo - 1, dx = pix % kHalo - 1;
                const bool inside = (uint32_t)(it.y0 + dy) < (uint32_t)H && (uint32_t)(it.x0 + dx) < (uint32_t)W;
                const uint32_t ry = (uint32_t)((dy >> sh) + 1), rx = (uint32_t)((dx >> sh) + 1);
                const uint32_t off = (inside ? ((ry * ws + rx) * C) * 2u : origin_off) + (uint32_t)q * 16u;
                sg.hv[k] = *(const gbf16x8 *)(hb + off);
                sg.hin |= (uint32_t)inside << k;
            }
        }
    };
    const auto store_weights = [&](const Stage &sg) {
#pragma unroll
        for (int k = 0; k < kWPer; ++k) {
            const int e = threadIdx.x + 256 * k;
            if (e < kWCh)
                *reinterpret_cast<bf16x8 *>(&s_w[(e >> 2) * kRow + (e & 3) * 8]) = ((sg.win >> k) & 1u) ? sg.wv[k] : zero;
        }
    };
    const auto store_halo = [&](const Stage &sg) {
        if constexpr (FEAT) {
            // k_pack_input's packing: log1p of radiance and depth, bf16,
            // zero padding to 32 channels (and outside the image)
#pragma unroll
            for (int k = 0; k < kFeatPer; ++k) {
                const int pix = threadIdx.x + 256 * k;
                if (pix < kFeatPix) {
                    bf16x8 o[kInPadded / 8];
#pragma unroll
                    for (int c = 0; c < kInPadded; ++c) {
                        float v = 0.0f;
                        if (c < kInChannels) {
                            v = sg.fv[k][c];
                            if (c < 3 || c == 9) v = log1p_feature(fmaxf(v, 0.0f));
                        }
                        o[c / 8][c % 8] = ((sg.hin >> k) & 1u) ? f2b(v) : f2b(0.0f);
                    }
#pragma unroll
                    for (int q = 0; q < kInPadded / 8; ++q)
                        *reinterpret_cast<bf16x8 *>(&s_in[pix * kRow + q * 8]) = o[q];
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kHalfPer; ++k) {
                const int e = threadIdx.x + 256 * k;
                if (e < kHalfCh)
                    *reinterpret_cast<bf16x8 *>(&s_in[(e >> 2) * kRow + (e & 3) * 8]) =
                        ((sg.hin >> k) & 1u) ? sg.hv[k] : zero;
            }
        }
    };
    // written first by the MFMAs of the first slice's taps (0, 0) to (2, 0),
    // which add to zero
    f32x4 acc[8][2];
    // the taps of one staged slice (the LDS barrier before them makes the
    // stores visible, the one after frees the stage for the next stores)
    const auto compute = [&](int c0) {
        __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of the taps
        lds_barrier();
        // the taps of column shift dx; FIRST: the item's first (dy = 0
        // starts the sums from zero)
        const auto taps = [&](int dx, auto first) {
            constexpr bool FIRST = decltype(first)::value;
            const f32x4 zacc = {0.0f, 0.0f, 0.0f, 0.0f};
            // operands: halo rows rw0 .. rw0+9 shifted by dx (A: lane =
            // pixel column, 8 channels), weights of taps (dy, dx) (B: lane
            // = output channel, 8 channels)
            bf16x8 b[3][2];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    b[dy][t] = *reinterpret_cast<const bf16x8 *>(
                        &s_w[((dy * 3 + dx) * COT + cw0 + chan_of(t, l16)) * kRow + kq]);
            // two halves of 4 output rows (6 input rows each), so that
            // consecutive MFMAs go to 8 different accumulators
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                bf16x8 a[6];
#pragma unroll
                for (int i = 0; i < 6; ++i)
                    a[i] = *reinterpret_cast<const bf16x8 *>(&s_in[((rw0 + 4 * h + i) * kHalo + l16 + dx) * kRow + kq]);
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int t = 0; t < 2; ++t)
                            acc[4 * h + i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                b[dy][t], a[i + dy], (FIRST && dy == 0) ? zacc : acc[4 * h + i][t], 0, 0, 0);
            }
        };
        int dx0 = 0;
        if (c0 == 0) {
            taps(0, std::true_type{});
            dx0 = 1;
        }
#pragma unroll 1
        for (int dx = dx0; dx < 3; ++dx) taps(dx, std::false_type{});
        lds_barrier();
    };
    // Epilogue straight from the accumulators: with the weights as the
    // MFMA's first operand the result is D[channel][pixel]: lane l holds
    // rows 4*(l/16) .. +3 of both channel tiles for pixel column l % 16,
    // and with the rows mapped by chan_of those are channels
    // 8*(l/16) .. +7, 16 contiguous bytes of the NHWC output; bias and
    // ReLU applied, or the mask.
    const auto epilogue = [&](const Item &it) {
        const int co = it.co0 + cw0 + 8 * (lane >> 4);
        const int x = it.x0 + l16;
        float bv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) bv[e] = (kAct && co + e < Cout) ? bias[co + e] : 0.0f;
        if (EPI == EPI_HEAD) {
            // fused head as one GEMM per row on the matrix cores: the
            // weights are the first operand with output k in row 4k (so
            // that lane group k ends with output k of its pixel column),
            // split into bf16 high and low parts (two MFMAs, about 16
            // bits of the FP32 weights); the second operand is this
            // lane's 8 channels of the pixel, bf16 as stored
            const int g = lane >> 4;
            bf16x8 whi, wlo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float wv = (l16 & 3) == 0 && l16 < 4 * kOut ? head.w[(l16 >> 2) * 32 + 8 * g + e] : 0.0f;
                whi[e] = f2b(wv);
                wlo[e] = f2b(wv - b2f(whi[e]));
            }
            const float hb = g < kOut ? head.b[g] : 0.0f;
            // the 8 rows' depth guides and input radiance loaded before
            // any is used (rows past the image read row H - 1, columns
            // past it column W - 1)
            const int xc = min(x, W - 1);
            float depth[8], rad[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const size_t p = ((size_t)it.n * H + min(it.y0 + rw0 + i, H - 1)) * W + xc;
                depth[i] = head.feat[p * kInChannels + 9];
                rad[i] = head.feat[p * kInChannels + min(g, kOut - 1)];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int y = it.y0 + rw0 + i;
                if (y >= H) break;
                bf16x8 v;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    v[e] = f2b(fmaxf(acc[i][e >> 2][e & 3] + bv[e], 0.0f));
                }
                f32x4 hs = {0.0f, 0.0f, 0.0f, 0.0f};
                hs = __builtin_amdgcn_mfma_f32_16x16x32_bf16(whi, v, hs, 0, 0, 0);
                hs = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wlo, v, hs, 0, 0, 0);
                if (g < kOut && x < W) {
                    const size_t p = ((size_t)it.n * H + y) * W + x;
                    head.out[p * kOut + g] = depth[i] == 0.0f ? rad[i] : fmaxf(expm1f(hb + hs[0]), 0.0f);
                }
            }
        } else if (x < W && co < Cout) {
            // addresses: the tile origin's element (uniform) plus this
            // lane's 32-bit offset (row rw0, column l16, its 8 channels)
            // and a row stride (W <= 65536 keeps them in 32 bits)
            typedef __attribute__((address_space(1))) bf16x8 gbf16x8w;
            typedef __attribute__((address_space(1))) const bf16x8 gbf16x8r;
            const size_t t0 = (((size_t)it.n * H + it.y0) * W + it.x0) * Cout;
            const uint32_t lo = ((uint32_t)rw0 * (uint32_t)W + (uint32_t)l16) * (uint32_t)Cout + (uint32_t)co;
            const uint32_t rs = (uint32_t)W * (uint32_t)Cout;
            char *const ob = (char *)(out + t0);
            // the ReLU mask of all 8 rows loaded before any is used (rows
            // past the image read row H - 1 and are not stored)
            bf16x8 m[8];
            if (EPI == EPI_MASK) {
                const char *mb = (const char *)(mask + t0);
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const uint32_t r = (uint32_t)min(rw0 + i, H - 1 - it.y0) - (uint32_t)rw0;
                    m[i] = *(const gbf16x8r *)(mb + (lo + r * rs) * 2u);
                }
            }
            bf16x8 prev;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int y = it.y0 + rw0 + i;
                if (y >= H) break;
                bf16x8 v;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    float f = acc[i][e >> 2][e & 3] + bv[e];
                    if (kAct) f = fmaxf(f, 0.0f);
                    v[e] = f2b(f);
                }
                if (EPI == EPI_MASK) {
                    // gradient through the ReLU whose output is `mask` (same shape)
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        if (!(b2f(m[i][e]) > 0.0f)) v[e] = f2b(0.0f);
                }
                *(gbf16x8w *)(ob + (lo + (uint32_t)i * rs) * 2u) = v;
                if (EPI == EPI_FORWARD && pool) {
                    // 2x2 average pooling of the stored values, summed top
                    // left, top right, bottom left, bottom right, times
                    // 0.25: the tile's rows pair up within a lane (y0 and
                    // rw0 are even), its columns with lane ^ 1 (x0 is even,
                    // DPP quad permutation 1, 0, 3, 2); H and W are even,
                    // so both lanes of a pair are here
                    if (i & 1) {
                        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                        const u32x4 pv = __builtin_bit_cast(u32x4, prev), cv = __builtin_bit_cast(u32x4, v);
                        u32x4 ptr, pbr;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            ptr[j] = (uint32_t)__builtin_amdgcn_mov_dpp((int)pv[j], 0xB1, 0xF, 0xF, false);
                            pbr[j] = (uint32_t)__builtin_amdgcn_mov_dpp((int)cv[j], 0xB1, 0xF, 0xF, false);
                        }
                        const bf16x8 tr = __builtin_bit_cast(bf16x8, ptr), br = __builtin_bit_cast(bf16x8, pbr);
                        if ((l16 & 1) == 0) {
                            bf16x8 o;
#pragma unroll
                            for (int e = 0; e < 8; ++e)
                                o[e] = f2b(0.25f * (b2f(prev[e]) + b2f(tr[e]) + b2f(v[e]) + b2f(br[e])));
                            const size_t p0 = (((size_t)it.n * (H / 2) + (it.y0 >> 1)) * (W / 2) + (it.x0 >> 1)) * Cout;
                            const uint32_t plo = ((uint32_t)((rw0 + i) >> 1) * (uint32_t)(W / 2) + (uint32_t)(l16 >> 1)) *
                                                     (uint32_t)Cout + (uint32_t)co;
                            *(gbf16x8w *)((char *)(pool + p0) + plo * 2u) = o;
                        }
                    }
                    prev = v;
                }
            }
        }
    };
    if constexpr (SINGLE) {
        // one slice and one output-channel block (Cin = 32, Cout <= COT): the
        // weights are staged once; items alternate between two halo stages,
        // each loaded two items ahead
        const int grid = (int)gridDim.x;
        int itn = blockIdx.x;
        if (itn >= items) return;
        Stage sa, sb;
        issue_weights(sa, item_of(itn), 0);
        issue_halo(sa, item_of(itn), 0);
        if (itn + grid < items) issue_halo(sb, item_of(itn + grid), 0);
        store_weights(sa);
        for (;;) {
            store_halo(sa);
            if (itn + 2 * grid < items) issue_halo(sa, item_of(itn + 2 * grid), 0);
            compute(0);
            epilogue(item_of(itn));
            itn += grid;
            if (itn >= items) break;
            store_halo(sb);
            if (itn + 2 * grid < items) issue_halo(sb, item_of(itn + 2 * grid), 0);
            compute(0);
            epilogue(item_of(itn));
            itn += grid;
            if (itn >= items) break;
        }
    } else {
        Stage sg;
        if ((int)blockIdx.x < items) {
            issue_halo(sg, item_of(blockIdx.x), 0);
            issue_weights(sg, item_of(blockIdx.x), 0);
        }
        for (int itn = blockIdx.x; itn < items; itn += gridDim.x) {
            const Item it = item_of(itn);
            for (int c0 = 0; c0 < Cin; c0 += 32) {
                store_halo(sg);
                store_weights(sg);
                // the next step: the next slice of this item, or slice 0 of
                // the next item
                {
                    const bool more = c0 + 32 < Cin;
                    const int nitem = more ? itn : itn + (int)gridDim.x;
                    if (nitem < items) {
                        issue_halo(sg, item_of(nitem), more ? c0 + 32 : 0);
                        issue_weights(sg, item_of(nitem), more ? c0 + 32 : 0);
                    }
                }
                compute(c0);
            }
            epilogue(it);
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[tap][co][ci] += sum_p dPre[p][co] * X[p + off(tap)][ci]
// and db[co] += sum_p dPre[p][co]: a GEMM per tap with M = co, N = ci,
// K = pixels, on v_mfma_f32_32x32x16_bf16.  Both operands need K (pixels)
// consecutive in a lane while NHWC keeps channels consecutive, so the tiles
// are staged in LDS as they lie in memory (16-byte copies) and read with the
// gfx950 transposing LDS load ds_read_b64_tr_b16: in each group of 16 lanes,
// lane l receives element l % 4 of the 8-byte rows addressed by lanes
// 4k + l / 4, k = 0..3.  A lane that addresses pixel kr + 4h, channels
// cc .. cc+3 therefore delivers, after the exchange, 4 consecutive pixels of
// one channel, and two such loads make one MFMA operand.
//
// A workgroup owns 32 output x 32 input channels x all 9 taps and walks a
// strided set of 16 x 16 pixel tiles; wave w takes tile rows 4w .. 4w+3 as
// the K steps (one 16-pixel row each).  Input rows r + dy serve taps dy = 0..2
// of three consecutive output rows, so each K step loads one new halo row
// (three column shifts) and the dPre row: 8 transposing loads per 9 MFMAs.
// The bias gradient is one more MFMA per K step, against a ones operand, in
// the workgroups of the first input-channel block.  The four waves' partial
// sums meet in LDS; one atomic add per weight and workgroup.
// ---------------------------------------------------------------------------
constexpr int kDwPitch = 48;   // LDS pixel row: 32 channels + 16 bf16 padding (96 B, 4 pixels on distinct banks)

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s lds_v4s;

// 8 bf16 of one MFMA operand: two transposing loads, the second `4 * step`
// elements (4 pixels) after the first.  Assembled from whole dwords: built
// element by element from the two short4 results, the operand came out with
// its dwords duplicated.
__device__ __forceinline__ bf16x8 tr_operand(const bf16 *p, int step) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const u32x2 lo = __builtin_bit_cast(u32x2, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)(p)));
    const u32x2 hi = __builtin_bit_cast(u32x2, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)(p + 4 * step)));
    const u32x4 w = {lo.x, lo.y, hi.x, hi.y};
    return __builtin_bit_cast(bf16x8, w);
}

__global__ __launch_bounds__(256, 2) void k_conv_dw(ConvIn I, int32_t N, int32_t H, int32_t W,
                                                     const bf16 *__restrict__ dpre, int32_t Cout, int32_t bias_grad,
                                                     float *__restrict__ part, float *__restrict__ bpart) {
    // one array: the cross-wave reduction at the end reuses both tiles
    __shared__ __attribute__((aligned(16))) bf16 smem[(kTile * kTile + kHalo * kHalo) * kDwPitch];
    bf16 *const s_d = smem;                              // dPre tile [pixel][32 co]
    bf16 *const s_x = smem + kTile * kTile * kDwPitch;   // input halo [pixel][32 ci]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Cin = I.C1 + I.C2;
    const int nci = Cin / 32;
    const int ci0 = (blockIdx.x % nci) * 32;
    const int co0 = (blockIdx.x / nci) * 32;
    const bool bias = bias_grad && ci0 == 0;
    const int tiles_x = (W + kTile - 1) / kTile, tiles_y = (H + kTile - 1) / kTile;
    const long long ntiles = (long long)N * tiles_x * tiles_y;
    // transposing-load addressing of this lane: pixel kr (+4 for the second
    // load) of a 16-pixel K step, channels cc .. cc+3 of the 32-channel block
    const int j = lane & 15, g = lane >> 4;
    const int kr = 8 * (g >> 1) + (j >> 2);
    const int cc = 16 * (g & 1) + 4 * (j & 3);
    f32x16 acc[9], accb;
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) accb[i] = 0.0f;
    bf16x8 ones, zero;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        ones[e] = f2b(1.0f);
        zero[e] = f2b(0.0f);
    }
    constexpr int kDCh = kTile * kTile * 4, kDPer = kDCh / 256;              // 16-byte chunks of the dPre tile
    constexpr int kXCh = kHalo * kHalo * 4, kXPer = (kXCh + 255) / 256;      // ... of the halo
    for (long long tile = blockIdx.y; tile < ntiles; tile += gridDim.y) {
        const int n = (int)(tile / ((long long)tiles_x * tiles_y));
        const int rem = (int)(tile % ((long long)tiles_x * tiles_y));
        const int y0 = (rem / tiles_x) * kTile, x0 = (rem % tiles_x) * kTile;
        // every global load of the tile is issued before the first LDS store
        bf16x8 dv[kDPer], xv[kXPer];
#pragma unroll
        for (int k = 0; k < kDPer; ++k) {
            const int e = threadIdx.x + 256 * k;
            const int px = e >> 2, q = e & 3;
            const int y = y0 + px / kTile, x = x0 + px % kTile;
            dv[k] = zero;
            if (y < H && x < W) dv[k] = *reinterpret_cast<const bf16x8 *>(dpre + (((size_t)n * H + y) * W + x) * Cout + co0 + 8 * q);
        }
        // unconditional halo loads (spare threads of the last round reload
        // the last chunk and do not store it): no load waits in a branch
#pragma unroll
        for (int k = 0; k < kXPer; ++k) {
            const int e = min((int)threadIdx.x + 256 * k, kXCh - 1);
            const int hp = e >> 2, q = e & 3;
            xv[k] = load8(I, n, y0 + hp / kHalo - 1, x0 + hp % kHalo - 1, H, W, ci0 + 8 * q);
        }
#pragma unroll
        for (int k = 0; k < kDPer; ++k) {
            const int e = threadIdx.x + 256 * k;
            *reinterpret_cast<bf16x8 *>(&s_d[(e >> 2) * kDwPitch + (e & 3) * 8]) = dv[k];
        }
#pragma unroll
        for (int k = 0; k < kXPer; ++k) {
            const int e = threadIdx.x + 256 * k;
            if (e < kXCh) *reinterpret_cast<bf16x8 *>(&s_x[(e >> 2) * kDwPitch + (e & 3) * 8]) = xv[k];
        }
        __syncthreads();
        // K steps: output rows r0 .. r0+3; xw[i][dx] holds halo row r0 + i
        // (output row r0 + i - 1) shifted by dx
        const int r0 = 4 * wave;
        bf16x8 xw[6][3];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
                xw[i][dx] = tr_operand(&s_x[((r0 + i) * kHalo + kr + dx) * kDwPitch + cc], kDwPitch);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
                xw[i + 2][dx] = tr_operand(&s_x[((r0 + i + 2) * kHalo + kr + dx) * kDwPitch + cc], kDwPitch);
            const bf16x8 a = tr_operand(&s_d[((r0 + i) * kTile + kr) * kDwPitch + cc], kDwPitch);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx)
                    acc[dy * 3 + dx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xw[i + dy][dx], acc[dy * 3 + dx], 0, 0, 0);
            if (bias) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, ones, accb, 0, 0, 0);
        }
        __syncthreads();
    }
    // the four waves' sums: waves 1..3 in turn through LDS (the tile buffers
    // are free now) into wave 0
    float *red = reinterpret_cast<float *>(smem);
    static_assert(10 * 16 * 64 * sizeof(float) <= sizeof(smem), "reduction buffer");
    for (int w = 1; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int t = 0; t < 10; ++t)
#pragma unroll
                for (int i = 0; i < 16; i += 4) {
                    const f32x16 &v = t < 9 ? acc[t] : accb;
                    *reinterpret_cast<f32x4 *>(&red[((t * 4 + i / 4) * 64 + lane) * 4]) =
                        f32x4{v[i], v[i + 1], v[i + 2], v[i + 3]};
                }
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int t = 0; t < 10; ++t)
#pragma unroll
                for (int i = 0; i < 16; i += 4) {
                    const f32x4 v = *reinterpret_cast<const f32x4 *>(&red[((t * 4 + i / 4) * 64 + lane) * 4]);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if (t < 9) acc[t][i + k] += v[k];
                        else accb[i + k] += v[k];
                    }
                }
        }
        __syncthreads();
    }
    if (wave != 0) return;
    // this slice's partial sums (plain stores; k_conv_dw_reduce adds the
    // slices up: same-address atomics from hundreds of workgroups serialise).
    // D layout: row (output channel) = 8*(i/4) + 4*(lane/32) + i%4, column
    // (input channel) = lane % 32
    float *pw = part + (size_t)blockIdx.y * 9 * Cout * Cin;
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int co = co0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            pw[((size_t)t * Cout + co) * Cin + ci0 + (lane & 31)] = acc[t][i];
        }
    // bias: every column of accb is the row sum
    if (bias && (lane & 31) == 0) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
            bpart[(size_t)blockIdx.y * Cout + co0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)] = accb[i];
    }
}

// dst[i] += sum over slices s of part[s * n + i], i < n (n a multiple of 4),
// for two arrays at once (workgroups [0, blocks_a) take the first): 16 float4
// columns per workgroup and 16 groups of slices (a slice every 16th), four
// independent loads in flight per thread, the groups summed in a fixed order.
struct SliceSum {
    const float *part;
    int32_t slices, n;
    float *dst;
};

__global__ __launch_bounds__(256) void k_slice_sum(SliceSum a, SliceSum b, int32_t blocks_a) {
    __shared__ f32x4 red[16][16];
    const bool first = (int)blockIdx.x < blocks_a;
    const SliceSum &S = first ? a : b;
    const int col = (first ? (int)blockIdx.x : (int)blockIdx.x - blocks_a) * 16 + (threadIdx.x & 15);
    const int grp = threadIdx.x >> 4;
    f32x4 acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if (4 * col < S.n) {
        const float *src = S.part + 4 * col;
        int k = grp;
        for (; k + 48 < S.slices; k += 64) {
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] += *reinterpret_cast<const f32x4 *>(src + (size_t)(k + 16 * u) * S.n);
        }
        for (; k < S.slices; k += 16) acc[0] += *reinterpret_cast<const f32x4 *>(src + (size_t)k * S.n);
    }
    red[grp][threadIdx.x & 15] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (grp == 0 && 4 * col < S.n) {
        f32x4 t = red[0][threadIdx.x];
#pragma unroll
        for (int g = 1; g < 16; ++g) t += red[g][threadIdx.x];
        f32x4 *d = reinterpret_cast<f32x4 *>(S.dst + 4 * col);
        *d += t;
    }
}

// dX += upsample(dY) / 4, the last contribution to dX, then (with mask) the
// gradient through the ReLU whose output is `mask`.  One thread per pixel and
// group of 8 channels (16-byte accesses), 32-bit indexing (N*H*W <= 2^27,
// C <= 224).
__global__ void k_pool2_back(const bf16 *__restrict__ dy, int32_t N, int32_t H, int32_t W, int32_t C,
                             const bf16 *__restrict__ mask, bf16 *__restrict__ dx) {
    const uint32_t G = (uint32_t)C / 8, Ho = (uint32_t)H / 2, Wo = (uint32_t)W / 2;
    const uint32_t n = (uint32_t)N * H * W * G;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t g = i % G, p = i / G;
        const uint32_t x = p % (uint32_t)W, y = (p / (uint32_t)W) % (uint32_t)H, b = p / ((uint32_t)W * H);
        const size_t off = (size_t)p * C + 8 * g;
        bf16x8 v = *reinterpret_cast<const bf16x8 *>(dx + off);
        bf16x8 d;
        if (y / 2 < Ho && x / 2 < Wo) {
            d = *reinterpret_cast<const bf16x8 *>(dy + (((size_t)b * Ho + y / 2) * Wo + x / 2) * C + 8 * g);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) d[j] = f2b(0.0f);
        }
        bf16x8 m;
        if (mask) m = *reinterpret_cast<const bf16x8 *>(mask + off);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            v[j] = (mask && !(b2f(m[j]) > 0.0f)) ? f2b(0.0f) : f2b(b2f(v[j]) + 0.25f * b2f(d[j]));
        *reinterpret_cast<bf16x8 *>(dx + off) = v;
    }
}
// Split the data gradient of a concat(upsample2(low), skip) input:
// dLow = 2x2 sum of the first C1 channels (with low_mask: zero where that
// ReLU output is not positive, dLow being its whole gradient), dSkip = the
// rest.  One thread per (low pixel or skip pixel, group of 8 channels),
// 16-byte accesses, 32-bit indexing (N*H*W <= 2^27, C1 <= 128, C2 <= 96).
__global__ void k_split_grad(const bf16 *__restrict__ dcat, int32_t N, int32_t H, int32_t W, int32_t C1, int32_t C2,
                             const bf16 *__restrict__ low_mask, bf16 *__restrict__ dlow, bf16 *__restrict__ dskip) {
    const uint32_t C = (uint32_t)(C1 + C2), G1 = (uint32_t)C1 / 8, G2 = (uint32_t)C2 / 8;
    const uint32_t Hl = (uint32_t)(H + 1) / 2, Wl = (uint32_t)(W + 1) / 2;
    const uint32_t n_low = (uint32_t)N * Hl * Wl * G1;
    const uint32_t total = n_low + (uint32_t)N * H * W * G2;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < total; i += gridDim.x * 256u) {
        if (i < n_low) {
            const uint32_t g = i % G1, p = i / G1;
            const uint32_t xl = p % Wl, yl = (p / Wl) % Hl, b = p / (Wl * Hl);
            float s[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t y = 2 * yl + (k >> 1), x = 2 * xl + (k & 1);
                if (y < (uint32_t)H && x < (uint32_t)W) {
                    const bf16x8 v =
                        *reinterpret_cast<const bf16x8 *>(dcat + (((size_t)b * H + y) * W + x) * C + 8 * g);
#pragma unroll
                    for (int j = 0; j < 8; ++j) s[j] += b2f(v[j]);
                }
            }
            const size_t off = (size_t)p * C1 + 8 * g;
            bf16x8 m, o;
            if (low_mask) m = *reinterpret_cast<const bf16x8 *>(low_mask + off);
#pragma unroll
            for (int j = 0; j < 8; ++j) o[j] = (low_mask && !(b2f(m[j]) > 0.0f)) ? f2b(0.0f) : f2b(s[j]);
            *reinterpret_cast<bf16x8 *>(dlow + off) = o;
        } else {
            const uint32_t j = i - n_low, g = j % G2, p = j / G2;
            *reinterpret_cast<bf16x8 *>(dskip + (size_t)p * C2 + 8 * g) =
                *reinterpret_cast<const bf16x8 *>(dcat + (size_t)p * C + C1 + 8 * g);
        }
    }
}

// The denoiser's 10 input features per pixel from a render's device buffers:
// mean radiance (the accumulated sum divided by the sample count, an IEEE
// division as the host's mean), albedo, normal, depth.  One thread per pixel.
__global__ void k_features(long long npix, const float *__restrict__ accum, float spp,
                           const float *__restrict__ albedo, const float *__restrict__ normal,
                           const float *__restrict__ depth, float *__restrict__ out) {
    for (long long p = blockIdx.x * 256LL + threadIdx.x; p < npix; p += (long long)gridDim.x * 256) {
        float *o = out + p * kInChannels;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            o[c] = accum[3 * p + c] / spp;
            o[3 + c] = albedo[3 * p + c];
            o[6 + c] = normal[3 * p + c];
        }
        o[9] = depth[p];
    }
}

// Input features -> padded bf16: log1p radiance, albedo, normal, log1p depth.
template <bool PAIRS>
__global__ void k_pack_input(const float *__restrict__ f, long long npix, bf16 *__restrict__ x) {
    // one thread per pixel: its 10 features in (with PAIRS, for 8-byte
    // aligned features, as five 8-byte loads), 32 bf16 out as four 16-byte
    // stores
    static_assert(kInChannels == 10, "five float2 per pixel");
    for (long long p = blockIdx.x * 256LL + threadIdx.x; p < npix; p += (long long)gridDim.x * 256) {
        float in[kInChannels];
        if (PAIRS) {
            const float2 *src = reinterpret_cast<const float2 *>(f + p * kInChannels);
#pragma unroll
            for (int k = 0; k < kInChannels / 2; ++k) {
                const float2 v = src[k];
                in[2 * k] = v.x;
                in[2 * k + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int c = 0; c < kInChannels; ++c) in[c] = f[p * kInChannels + c];
        }
        bf16x8 o[kInPadded / 8];
#pragma unroll
        for (int c = 0; c < kInPadded; ++c) {
            float v = 0.0f;
            if (c < kInChannels) {
                v = in[c];
                if (c < 3 || c == 9) v = log1p_feature(fmaxf(v, 0.0f));
            }
            o[c / 8][c % 8] = f2b(v);
        }
        bf16x8 *dst = reinterpret_cast<bf16x8 *>(x + p * kInPadded);
#pragma unroll
        for (int q = 0; q < kInPadded / 8; ++q) dst[q] = o[q];
    }
}

// 1x1 head (32 -> 3, FP32), its input the last ReLU's output.  The
// prediction is expm1 of the head output;
// inference writes it clamped at 0.  Pixels whose camera ray missed the scene
// (depth guide exactly 0) carry the exact environment radiance in every
// sample, so their input radiance is passed through and they take no part in
// the loss.  Training uses the relative squared error
// (pred - target)^2 / (pred^2 + kRelEps) with the denominator held constant
// (unbiased under noisy targets), writes dL/dX masked by X > 0 (bf16: the
// gradient of the last convolution's pre-activation) and accumulates the head
// gradients and the loss.
constexpr float kRelEps = 1e-2f;

// Eight lanes per pixel, each holding 4 of its 32 input channels (8-byte
// loads and stores, 32 pixels per workgroup pass); the head sums are
// completed across the eight lanes with shuffles, and the head gradients and
// loss are reduced per wave, then per workgroup, into one partial per
// workgroup (no atomics: thousands of workgroups adding into the same hundred
// addresses serialise).
__global__ __launch_bounds__(256) void k_head(const bf16 *__restrict__ x, long long npix, const float *__restrict__ wh,
                                               const float *__restrict__ bh, const float *__restrict__ feat,
                                               const float *__restrict__ target,
                                               float *__restrict__ out, bf16 *__restrict__ dx,
                                               float *__restrict__ part, double *__restrict__ lpart, float inv_count) {
    __shared__ float sg[4][kOut * 33];
    __shared__ double sl[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = threadIdx.x & 7;   // channels 4q .. 4q+3
    float w[kOut][4], bias[kOut], gw[kOut][4], gb[kOut];
#pragma unroll
    for (int k = 0; k < kOut; ++k) {
        bias[k] = bh[k];
        gb[k] = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            w[k][j] = wh[k * 32 + 4 * q + j];
            gw[k][j] = 0.0f;
        }
    }
    double l = 0.0;
    // four pixels per lane group and pass, their loads issued together
    constexpr int U = 4;
    const long long stride = gridDim.x * 32LL;
    for (long long p0 = blockIdx.x * 32LL + (threadIdx.x >> 3); p0 < npix; p0 += U * stride) {
        bf16x4 xv[U];
        float depth[U], tg[U][kOut];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long p = p0 + u * stride;
            if (p < npix) {
                xv[u] = *reinterpret_cast<const bf16x4 *>(x + p * 32 + 4 * q);
                depth[u] = feat[p * kInChannels + 9];
#pragma unroll
                for (int k = 0; k < kOut; ++k) tg[u][k] = target[p * kOut + k];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long p = p0 + u * stride;
            if (p >= npix) break;   // the same for the eight lanes of a pixel
            float xi[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) xi[j] = b2f(xv[u][j]);
            float o[kOut];
#pragma unroll
            for (int k = 0; k < kOut; ++k) {
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; ++j) s += w[k][j] * xi[j];
                // the eight lanes of this pixel
                s += __shfl_xor(s, 1);
                s += __shfl_xor(s, 2);
                s += __shfl_xor(s, 4);
                o[k] = s + bias[k];
            }
            bf16x4 dv;
            if (depth[u] == 0.0f) {
                if (out && q < kOut) out[p * kOut + q] = feat[p * kInChannels + q];
#pragma unroll
                for (int j = 0; j < 4; ++j) dv[j] = f2b(0.0f);
            } else {
                float g[kOut];
#pragma unroll
                for (int k = 0; k < kOut; ++k) {
                    const float t = fmaxf(tg[u][k], 0.0f);
                    // outputs beyond e^30 are treated as e^30 (keeps pred^2 finite)
                    const float oc = fminf(o[k], 30.0f);
                    const float pred = expm1f(oc);
                    const float d = pred - t;
                    const float den = pred * pred + kRelEps;
                    if (q == 0) l += (double)(d * d / den);
                    g[k] = 2.0f * d / den * expf(oc) * inv_count;
                    if (out && q == k) out[p * kOut + k] = fmaxf(pred, 0.0f);
                    if (q == 0) gb[k] += g[k];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float s = 0.0f;
#pragma unroll
                    for (int k = 0; k < kOut; ++k) {
                        s += w[k][j] * g[k];
                        gw[k][j] += g[k] * xi[j];
                    }
                    // through the ReLU whose output x is: the gradient of the
                    // last convolution's pre-activation
                    dv[j] = (xi[j] > 0.0f) ? f2b(s) : f2b(0.0f);
                }
            }
            *reinterpret_cast<bf16x4 *>(dx + p * 32 + 4 * q) = dv;
        }
    }
    // lanes q, q + 8, ..., q + 56 hold the same channels
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) {
#pragma unroll
        for (int k = 0; k < kOut; ++k) {
            gb[k] += __shfl_xor(gb[k], off);
#pragma unroll
            for (int j = 0; j < 4; ++j) gw[k][j] += __shfl_xor(gw[k][j], off);
        }
        l += __shfl_xor(l, off);
    }
    if (lane < 8) {
#pragma unroll
        for (int k = 0; k < kOut; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) sg[wave][k * 33 + 4 * lane + j] = gw[k][j];
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kOut; ++k) sg[wave][k * 33 + 32] = gb[k];
        sl[wave] = l;
    }
    __syncthreads();
    // this workgroup's sums, added up by k_head_reduce
    if (threadIdx.x < kOut * 33)
        part[(size_t)blockIdx.x * (kOut * 33) + threadIdx.x] =
            sg[0][threadIdx.x] + sg[1][threadIdx.x] + sg[2][threadIdx.x] + sg[3][threadIdx.x];
    if (threadIdx.x == 0) lpart[blockIdx.x] = sl[0] + sl[1] + sl[2] + sl[3];
}

// Sums k_head's per-workgroup partials: workgroup v < kOut * 33 adds head
// gradient v (weight k * 32 + c, or bias k for c = 32) into dwh / dbh,
// workgroup kOut * 33 adds the loss.  Fixed order, so the sums are
// reproducible; one writer per value.
__global__ __launch_bounds__(256) void k_head_reduce(const float *__restrict__ part, const double *__restrict__ lpart,
                                                     int32_t nwg, float *__restrict__ dwh, float *__restrict__ dbh,
                                                     double *__restrict__ loss) {
    __shared__ double red[256];
    const int v = blockIdx.x;
    double s = 0.0;
    for (int w = threadIdx.x; w < nwg; w += 256)
        s += v < kOut * 33 ? (double)part[(size_t)w * (kOut * 33) + v] : lpart[w];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (v == kOut * 33) {
            *loss += red[0];
        } else {
            const int k = v / 33, c = v % 33;
            if (c < 32) dwh[k * 32 + c] += (float)red[0];
            else dbh[k] += (float)red[0];
        }
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// bf16 copies of one layer: wf[tap][co][ci] and wb[8 - tap][ci][co].
__global__ void k_pack_weights(const float *__restrict__ w, int32_t Cout, int32_t Cin, bf16 *__restrict__ wf,
                               bf16 *__restrict__ wb) {
    const long long n = 9LL * Cout * Cin;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ci = (int)(i % Cin);
        const int co = (int)((i / Cin) % Cout);
        const int tap = (int)(i / ((long long)Cin * Cout));
        const bf16 b = f2b(w[i]);
        wf[i] = b;
        wb[((size_t)(8 - tap) * Cin + ci) * Cout + co] = b;
    }
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
namespace {

struct ConvLayer {
    int cin, cout;
    size_t w_off, b_off;      // offsets of W and b in the flat parameter vector
    bf16 *wf = nullptr, *wb = nullptr;
};

enum {
    L_E0A, L_E0B, L_E1A, L_E1B, L_E2A, L_E2B, L_B0, L_B1,
    L_D2A, L_D2B, L_D1A, L_D1B, L_D0A, L_D0B, L_COUNT
};

}  // namespace

struct s3dg_cnn {
    s3dg_context *ctx = nullptr;
    ConvLayer layer[L_COUNT];
    size_t head_w = 0, head_b = 0;
    size_t nparam = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // FP32 master, grad, Adam moments
    bf16 *wbuf = nullptr;
    int step = 0;
    // activation workspace
    void *ws = nullptr;
    size_t ws_bytes = 0;
    double *d_loss = nullptr;
    // inference launches: run directly the first time a (shape, buffers) key
    // is seen, captured into a HIP graph the second time, replayed after that
    struct InferGraph {
        uint64_t key[6];
        hipGraphExec_t exec;
        uint64_t used;   // tick of the last launch (least recently used is replaced)
    };
    static constexpr int kGraphs = 4, kSeen = 8;
    InferGraph graphs[kGraphs] = {};
    uint64_t seen[kSeen][6] = {};   // keys run directly (ring)
    int seen_next = 0;
    uint64_t tick = 0;
};

namespace {

const int kLayerIO[L_COUNT][2] = {
    {kInPadded, 32}, {32, 32}, {32, 64}, {64, 64}, {64, 96}, {96, 96}, {96, 128}, {128, 128},
    {128 + 96, 96}, {96, 96}, {96 + 64, 64}, {64, 64}, {64 + 32, 32}, {32, 32}};

size_t grid1(long long n) { return (size_t)std::max<long long>(1, std::min<long long>((n + 255) / 256, 65536)); }

struct Acts {
    // forward activations (post-ReLU), bf16 NHWC
    bf16 *x, *e0a, *e0b, *p0, *e1a, *e1b, *p1, *e2a, *e2b, *p2, *b0, *b1, *d2a, *d2b, *d1a, *d1b, *d0a, *d0b;
    // gradients
    bf16 *g_cat, *g_a, *g_b, *g_e0b, *g_e1b, *g_e2b, *g_low;
    // head partial sums, one row per k_head workgroup
    float *head_part;
    double *loss_part;
    // weight-gradient partial sums, one block per k_conv_dw slice
    float *dw_part, *db_part;
};

constexpr int kHeadWGs = 2048;   // k_head workgroups at most
constexpr int kDwWGs = 512;      // k_conv_dw workgroups (slices x channel blocks) at most
constexpr long long kDwPartFloats = (long long)kDwWGs * 9 * 32 * 32;   // slices x 9 x Cout x Cin <= this
constexpr long long kDbPartFloats = (long long)kDwWGs * 32;            // slices x Cout <= this

struct DwParts {
    float *w, *b;
};

size_t act_bytes(long long n0) {
    // element counts at full, 1/4, 1/16, 1/64 resolution
    const long long n1 = n0 / 4 + 64, n2 = n0 / 16 + 64, n3 = n0 / 64 + 64;
    long long e = 0;
    e += n0 * kInPadded + n0 * 32 * 2;           // x, e0a, e0b
    e += n1 * 32 + n1 * 64 * 2;                  // p0, e1a, e1b
    e += n2 * 64 + n2 * 96 * 2;                  // p1, e2a, e2b
    e += n3 * 96 + n3 * 128 * 2;                 // p2, b0, b1
    e += n2 * 96 * 2 + n1 * 64 * 2 + n0 * 32 * 2;  // decoder
    // gradients: concat buffer (largest: full res x 96), two layer buffers, three skips, low
    e += n0 * 224 + 2 * n0 * 96 + n0 * 32 + n1 * 64 + n2 * 96 + n0 * 128;
    return (size_t)e * 2 + (size_t)kHeadWGs * (kOut * 33 * sizeof(float) + sizeof(double)) +
           (size_t)(kDwPartFloats + kDbPartFloats) * sizeof(float) + 64 * 256;
}

Acts carve_acts(void *base, long long n0) {
    char *p = (char *)base;
    auto take = [&](long long elems) {
        bf16 *r = (bf16 *)p;
        p += ((size_t)elems * 2 + 255) & ~(size_t)255;
        return r;
    };
    const long long n1 = n0 / 4 + 64, n2 = n0 / 16 + 64, n3 = n0 / 64 + 64;
    Acts a;
    a.x = take(n0 * kInPadded);
    a.e0a = take(n0 * 32); a.e0b = take(n0 * 32); a.p0 = take(n1 * 32);
    a.e1a = take(n1 * 64); a.e1b = take(n1 * 64); a.p1 = take(n2 * 64);
    a.e2a = take(n2 * 96); a.e2b = take(n2 * 96); a.p2 = take(n3 * 96);
    a.b0 = take(n3 * 128); a.b1 = take(n3 * 128);
    a.d2a = take(n2 * 96); a.d2b = take(n2 * 96);
    a.d1a = take(n1 * 64); a.d1b = take(n1 * 64);
    a.d0a = take(n0 * 32); a.d0b = take(n0 * 32);
    a.g_cat = take(n0 * 224); a.g_a = take(n0 * 96); a.g_b = take(n0 * 96);
    a.g_e0b = take(n0 * 32); a.g_e1b = take(n1 * 64); a.g_e2b = take(n2 * 96);
    a.g_low = take(n0 * 128);
    a.head_part = reinterpret_cast<float *>(take((long long)kHeadWGs * kOut * 33 * 2));
    a.loss_part = reinterpret_cast<double *>(take((long long)kHeadWGs * 4));
    a.dw_part = reinterpret_cast<float *>(take(kDwPartFloats * 2));
    a.db_part = reinterpret_cast<float *>(take(kDbPartFloats * 2));
    return a;
}

// The device's CU count (cached per thread and device).
int conv_cus(s3dg_cnn *c) {
    static thread_local int cached_device = -1, cached_cus = 0;
    if (cached_device != c->ctx->device) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->ctx->device) != hipSuccess) cus = 0;
        cached_device = c->ctx->device;
        cached_cus = cus;
    }
    return cached_cus;
}

// mask (optional): out is the gradient of a ReLU's output, passed on only
// where that output (mask, same shape as out) is positive.  Persistent grid:
// two workgroups per CU at most, each walking its share of the (tile,
// channel block) items.
void conv(s3dg_cnn *c, const ConvIn &I, int N, int H, int W, const bf16 *wt, const float *bias, int Cout, bool relu,
          bf16 *out, const bf16 *mask = nullptr, HeadArgs head = HeadArgs{nullptr, nullptr, nullptr, nullptr},
          bf16 *pool = nullptr) {
    hipStream_t s = c->ctx->stream;
    const long long wgs = 2LL * conv_cus(c);
    const long long tx = (W + kTile - 1) / kTile;
    // the epilogue: forward (bias + ReLU, optional pooling), fused head,
    // masked or plain data gradient (no bias, no ReLU)
    const int epi = head.out ? EPI_HEAD : (mask ? EPI_MASK : (relu ? EPI_FORWARD : EPI_LINEAR));
#define S3DG_CONV(COT_, EPI_, FEAT_, SINGLE_)                                                                          \
    hipLaunchKernelGGL((k_conv3x3<COT_, EPI_, FEAT_, SINGLE_>), grid, dim3(256), 0, s, I, N, H, W, wt, bias, Cout,      \
                       mask, out, head, pool)
    // the fused head needs the 32-channel tiles (all of a pixel's channels in one wave)
    if (Cout % 64 == 0 && epi != EPI_HEAD) {
        const long long items = (long long)N * tx * ((H + 15) / 16) * (Cout / 64);
        const dim3 grid((unsigned)std::min(items, wgs));
        if (epi == EPI_MASK) S3DG_CONV(64, EPI_MASK, false, false);
        else if (epi == EPI_LINEAR) S3DG_CONV(64, EPI_LINEAR, false, false);
        else S3DG_CONV(64, EPI_FORWARD, false, false);
    } else {
        const long long items = (long long)N * tx * ((H + 31) / 32) * ((Cout + 31) / 32);
        const dim3 grid((unsigned)std::min(items, wgs));
        // one input slice and one output-channel block: the weights stay in
        // LDS and two tiles' loads are in flight (the forward layers at
        // 32 -> 32 channels and the first layer, whose full-resolution
        // input comes from HBM; not the head, whose epilogue needs the
        // registers the second stage would take, nor the half-resolution
        // 32 -> 64 layer or the data gradients, measured no faster)
        const bool single = I.C1 + I.C2 == 32 && Cout <= 32;
        if (epi == EPI_HEAD) S3DG_CONV(32, EPI_HEAD, false, false);
        else if (epi == EPI_MASK) S3DG_CONV(32, EPI_MASK, false, false);
        else if (epi == EPI_LINEAR) S3DG_CONV(32, EPI_LINEAR, false, false);
        else if (I.feat) S3DG_CONV(32, EPI_FORWARD, true, true);
        else if (single) S3DG_CONV(32, EPI_FORWARD, false, true);
        else S3DG_CONV(32, EPI_FORWARD, false, false);
    }
#undef S3DG_CONV
}

// Weight and bias gradients of one layer (Cin, Cout multiples of 32): one
// workgroup per (32 output, 32 input channels) block and slice of the pixel
// tiles, kDwWGs workgroups in all (two per CU on a 256-CU device), each
// slice's sums into the partial buffers, then added up per weight.
void conv_dw(s3dg_cnn *c, const DwParts &parts, const ConvIn &I, int N, int H, int W, const bf16 *dpre, int Cout,
             float *dw, float *db) {
    hipStream_t s = c->ctx->stream;
    const int Cin = I.C1 + I.C2;
    const long long ntiles = (long long)N * ((W + kTile - 1) / kTile) * ((H + kTile - 1) / kTile);
    const int blocks_xy = (Cout / 32) * (Cin / 32);
    const int slices = (int)std::max<long long>(1, std::min<long long>(ntiles, kDwWGs / blocks_xy));
    hipLaunchKernelGGL(k_conv_dw, dim3((unsigned)blocks_xy, (unsigned)slices), dim3(256), 0, s, I, N, H, W, dpre, Cout,
                       1, parts.w, parts.b);
    const SliceSum sw{parts.w, slices, 9 * Cout * Cin, dw}, sb{parts.b, slices, Cout, db};
    const int blocks_w = (sw.n / 4 + 15) / 16, blocks_b = (sb.n / 4 + 15) / 16;
    hipLaunchKernelGGL(k_slice_sum, dim3((unsigned)(blocks_w + blocks_b)), dim3(256), 0, s, sw, sb, blocks_w);
}

ConvIn cin1(const bf16 *a, int C) { return ConvIn{a, nullptr, C, 0, 1, nullptr}; }
ConvIn cin_cat(const bf16 *low, int C1, const bf16 *skip, int C2) { return ConvIn{low, skip, C1, C2, 2, nullptr}; }
// the input features, packed by the first convolution itself
ConvIn cin_feat(const float *feat) { return ConvIn{nullptr, nullptr, kInPadded, 0, 1, feat}; }

void pack_weights(s3dg_cnn *c) {
    for (int l = 0; l < L_COUNT; ++l) {
        ConvLayer &L = c->layer[l];
        const long long n = 9LL * L.cout * L.cin;
        hipLaunchKernelGGL(k_pack_weights, dim3(grid1(n)), dim3(256), 0, c->ctx->stream, c->p + L.w_off, L.cout, L.cin,
                           L.wf, L.wb);
    }
}

// Forward pass; returns the decoder output (d0b) for the head.
// Forward pass; with d_out (inference) the last convolution runs the fused
// head into d_out instead of storing d0b.
void forward(s3dg_cnn *c, Acts &a, const float *feat, int N, int H, int W, float *d_out = nullptr) {
    hipStream_t s = c->ctx->stream;
    const long long n0 = (long long)N * H * W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    ConvLayer *L = c->layer;
    auto B = [&](int l) { return c->p + L[l].b_off; };
    // inference with 8-byte aligned features: the first convolution packs
    // them while staging (training keeps the packed input for its gradient)
    const bool fuse_pack = d_out && ((uintptr_t)feat & 7u) == 0;
    if (fuse_pack) {
        conv(c, cin_feat(feat), N, H, W, L[L_E0A].wf, B(L_E0A), 32, true, a.e0a);
    } else {
        if (((uintptr_t)feat & 7u) == 0)
            hipLaunchKernelGGL(k_pack_input<true>, dim3(grid1(n0)), dim3(256), 0, s, feat, n0, a.x);
        else
            hipLaunchKernelGGL(k_pack_input<false>, dim3(grid1(n0)), dim3(256), 0, s, feat, n0, a.x);
        conv(c, cin1(a.x, kInPadded), N, H, W, L[L_E0A].wf, B(L_E0A), 32, true, a.e0a);
    }
    // the encoder's second convolutions also write their 2x2-pooled output
    const HeadArgs nohead{nullptr, nullptr, nullptr, nullptr};
    conv(c, cin1(a.e0a, 32), N, H, W, L[L_E0B].wf, B(L_E0B), 32, true, a.e0b, nullptr, nohead, a.p0);
    conv(c, cin1(a.p0, 32), N, H1, W1, L[L_E1A].wf, B(L_E1A), 64, true, a.e1a);
    conv(c, cin1(a.e1a, 64), N, H1, W1, L[L_E1B].wf, B(L_E1B), 64, true, a.e1b, nullptr, nohead, a.p1);
    conv(c, cin1(a.p1, 64), N, H2, W2, L[L_E2A].wf, B(L_E2A), 96, true, a.e2a);
    conv(c, cin1(a.e2a, 96), N, H2, W2, L[L_E2B].wf, B(L_E2B), 96, true, a.e2b, nullptr, nohead, a.p2);
    conv(c, cin1(a.p2, 96), N, H3, W3, L[L_B0].wf, B(L_B0), 128, true, a.b0);
    conv(c, cin1(a.b0, 128), N, H3, W3, L[L_B1].wf, B(L_B1), 128, true, a.b1);
    conv(c, cin_cat(a.b1, 128, a.e2b, 96), N, H2, W2, L[L_D2A].wf, B(L_D2A), 96, true, a.d2a);
    conv(c, cin1(a.d2a, 96), N, H2, W2, L[L_D2B].wf, B(L_D2B), 96, true, a.d2b);
    conv(c, cin_cat(a.d2b, 96, a.e1b, 64), N, H1, W1, L[L_D1A].wf, B(L_D1A), 64, true, a.d1a);
    conv(c, cin1(a.d1a, 64), N, H1, W1, L[L_D1B].wf, B(L_D1B), 64, true, a.d1b);
    conv(c, cin_cat(a.d1b, 64, a.e0b, 32), N, H, W, L[L_D0A].wf, B(L_D0A), 32, true, a.d0a);
    conv(c, cin1(a.d0a, 32), N, H, W, L[L_D0B].wf, B(L_D0B), 32, true, a.d0b, nullptr,
         HeadArgs{c->p + c->head_w, c->p + c->head_b, feat, d_out});
}

// Backward of one conv layer whose pre-activation gradient is in dpre (the
// producers of dpre apply the ReLU mask): accumulates dW and db, and (if dx)
// writes the input gradient, masked by dx_mask when the input is the output
// of a ReLU whose whole gradient this is.
void layer_back(s3dg_cnn *c, const DwParts &P, int l, const ConvIn &I, int N, int H, int W, const bf16 *dpre,
                bf16 *dx, const bf16 *dx_mask) {
    ConvLayer &L = c->layer[l];
    conv_dw(c, P, I, N, H, W, dpre, L.cout, c->g + L.w_off, c->g + L.b_off);
    if (dx) conv(c, cin1(dpre, L.cout), N, H, W, L.wb, nullptr, L.cin, false, dx, dx_mask);
}

// Splits the gradient of a concat(upsample2(low), skip) input; dLow is masked
// by low (a ReLU output whose whole gradient it is).
void split(s3dg_cnn *c, const bf16 *dcat, int N, int H, int W, int C1, int C2, const bf16 *low, bf16 *dlow,
           bf16 *dskip) {
    const long long total = (long long)N * ((H + 1) / 2) * ((W + 1) / 2) * (C1 / 8) + (long long)N * H * W * (C2 / 8);
    hipLaunchKernelGGL(k_split_grad, dim3(grid1(total)), dim3(256), 0, c->ctx->stream, dcat, N, H, W, C1, C2, low,
                       dlow, dskip);
}

// dX += upsample(dY) / 4, the last contribution to dX, then masked by the
// ReLU output `mask`.
void pool_back(s3dg_cnn *c, const bf16 *dy, int N, int H, int W, int C, const bf16 *mask, bf16 *dx) {
    hipLaunchKernelGGL(k_pool2_back, dim3(grid1((long long)N * H * W * (C / 8))), dim3(256), 0, c->ctx->stream, dy, N, H, W,
                       C, mask, dx);
}

// Backward pass.  a.g_a holds the gradient of d0b's pre-activation from the
// head on entry.  Skip gradients get the decoder's contribution first (the
// skip buffers are zeroed before), then the pooling path's, which applies
// their ReLU mask.
void backward(s3dg_cnn *c, Acts &a, int N, int H, int W) {
    const DwParts P{a.dw_part, a.db_part};
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    layer_back(c, P, L_D0B, cin1(a.d0a, 32), N, H, W, a.g_a, a.g_b, a.d0a);
    layer_back(c, P, L_D0A, cin_cat(a.d1b, 64, a.e0b, 32), N, H, W, a.g_b, a.g_cat, nullptr);
    split(c, a.g_cat, N, H, W, 64, 32, a.d1b, a.g_low, a.g_e0b);
    layer_back(c, P, L_D1B, cin1(a.d1a, 64), N, H1, W1, a.g_low, a.g_a, a.d1a);
    layer_back(c, P, L_D1A, cin_cat(a.d2b, 96, a.e1b, 64), N, H1, W1, a.g_a, a.g_cat, nullptr);
    split(c, a.g_cat, N, H1, W1, 96, 64, a.d2b, a.g_low, a.g_e1b);
    layer_back(c, P, L_D2B, cin1(a.d2a, 96), N, H2, W2, a.g_low, a.g_a, a.d2a);
    layer_back(c, P, L_D2A, cin_cat(a.b1, 128, a.e2b, 96), N, H2, W2, a.g_a, a.g_cat, nullptr);
    split(c, a.g_cat, N, H2, W2, 128, 96, a.b1, a.g_low, a.g_e2b);
    layer_back(c, P, L_B1, cin1(a.b0, 128), N, H3, W3, a.g_low, a.g_a, a.b0);
    layer_back(c, P, L_B0, cin1(a.p2, 96), N, H3, W3, a.g_a, a.g_b, nullptr);
    pool_back(c, a.g_b, N, H2, W2, 96, a.e2b, a.g_e2b);
    layer_back(c, P, L_E2B, cin1(a.e2a, 96), N, H2, W2, a.g_e2b, a.g_a, a.e2a);
    layer_back(c, P, L_E2A, cin1(a.p1, 64), N, H2, W2, a.g_a, a.g_b, nullptr);
    pool_back(c, a.g_b, N, H1, W1, 64, a.e1b, a.g_e1b);
    layer_back(c, P, L_E1B, cin1(a.e1a, 64), N, H1, W1, a.g_e1b, a.g_a, a.e1a);
    layer_back(c, P, L_E1A, cin1(a.p0, 32), N, H1, W1, a.g_a, a.g_b, nullptr);
    pool_back(c, a.g_b, N, H, W, 32, a.e0b, a.g_e0b);
    layer_back(c, P, L_E0B, cin1(a.e0a, 32), N, H, W, a.g_e0b, a.g_a, a.e0a);
    layer_back(c, P, L_E0A, cin1(a.x, kInPadded), N, H, W, a.g_a, nullptr, nullptr);
}

int check_dims(s3dg_context *ctx, int n, int h, int w) {
    if (n <= 0 || h <= 0 || w <= 0 || h % 8 || w % 8)
        return set_error(ctx, S3DG_ERR_INVALID, "denoiser input: N, H, W > 0 and H, W multiples of 8");
    if ((long long)n * h * w > (1LL << 27)) return set_error(ctx, S3DG_ERR_UNSUPPORTED, "denoiser batch too large");
    // the convolutions address a tile's rows with 32-bit offsets
    if (w > 65536) return set_error(ctx, S3DG_ERR_UNSUPPORTED, "denoiser input wider than 65536 pixels");
    return S3DG_OK;
}

int prepare(s3dg_cnn *c, long long n0, Acts &a) {
    int rc = ensure_buffer(c->ctx, &c->ws, &c->ws_bytes, act_bytes(n0));
    if (rc) return rc;
    a = carve_acts(c->ws, n0);
    return S3DG_OK;
}

// Loss and gradient of one batch into c->g (zeroed first); returns the mean loss.
int loss_and_grad(s3dg_cnn *c, int n, int h, int w, const float *d_feat, const float *d_target, double *loss) {
    s3dg_context *ctx = c->ctx;
    const long long n0 = (long long)n * h * w;
    Acts a;
    int rc = prepare(c, n0, a);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    S3DG_HIP(ctx, hipMemsetAsync(c->g, 0, c->nparam * sizeof(float), s));
    S3DG_HIP(ctx, hipMemsetAsync(c->d_loss, 0, sizeof(double), s));
    forward(c, a, d_feat, n, h, w);
    const float inv = 1.0f / (float)(n0 * kOut);
    const int head_wgs = (int)std::min<long long>((n0 + 31) / 32, kHeadWGs);
    hipLaunchKernelGGL(k_head, dim3((unsigned)head_wgs), dim3(256), 0, s, a.d0b, n0, c->p + c->head_w,
                       c->p + c->head_b, d_feat, d_target, nullptr, a.g_a, a.head_part, a.loss_part, inv);
    hipLaunchKernelGGL(k_head_reduce, dim3(kOut * 33 + 1), dim3(256), 0, s, a.head_part, a.loss_part, head_wgs,
                       c->g + c->head_w, c->g + c->head_b, c->d_loss);
    backward(c, a, n, h, w);
    S3DG_HIP(ctx, hipGetLastError());
    double l = 0.0;
    S3DG_HIP(ctx, hipMemcpyAsync(&l, c->d_loss, sizeof(double), hipMemcpyDeviceToHost, s));
    S3DG_HIP(ctx, hipStreamSynchronize(s));
    if (loss) *loss = l / (double)(n0 * kOut);
    return S3DG_OK;
}

}  // namespace

extern "C" int s3dg_cnn_create(s3dg_context *ctx, uint64_t seed, s3dg_cnn **out) {
    if (!ctx || !out) return S3DG_ERR_INVALID;
    *out = nullptr;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    s3dg_cnn *c = new (std::nothrow) s3dg_cnn();
    if (!c) return set_error(ctx, S3DG_ERR_NOMEM, "host allocation");
    c->ctx = ctx;
    size_t off = 0, wbf = 0;
    for (int l = 0; l < L_COUNT; ++l) {
        ConvLayer &L = c->layer[l];
        L.cin = kLayerIO[l][0];
        L.cout = kLayerIO[l][1];
        L.w_off = off;
        off += 9ull * L.cin * L.cout;
        L.b_off = off;
        off += L.cout;
        wbf += 2 * 9ull * L.cin * L.cout;
    }
    c->head_w = off;
    off += kOut * 32;
    c->head_b = off;
    off += kOut;
    c->nparam = off;
    // He-normal initialisation (fan-in 9 x real input channels), zero biases.
    std::vector<float> init(off, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    for (int l = 0; l < L_COUNT; ++l) {
        ConvLayer &L = c->layer[l];
        const int real_in = (l == L_E0A) ? kInChannels : L.cin;
        const float sd = sqrtf(2.0f / (9.0f * real_in));
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci)
                    init[L.w_off + ((size_t)tap * L.cout + co) * L.cin + ci] = ci < real_in ? nd(rng) * sd : 0.0f;
    }
    for (int i = 0; i < kOut * 32; ++i) init[c->head_w + i] = nd(rng) * sqrtf(1.0f / 32.0f);
    hipError_t e;
    if ((e = hipMalloc(&c->p, off * 4)) != hipSuccess || (e = hipMalloc(&c->g, off * 4)) != hipSuccess ||
        (e = hipMalloc(&c->m, off * 4)) != hipSuccess || (e = hipMalloc(&c->v, off * 4)) != hipSuccess ||
        (e = hipMalloc(&c->wbuf, wbf * 2)) != hipSuccess || (e = hipMalloc(&c->d_loss, 256)) != hipSuccess) {
        s3dg_cnn_destroy(c);
        return hip_check(ctx, e, "hipMalloc(cnn)");
    }
    bf16 *wp = c->wbuf;
    for (int l = 0; l < L_COUNT; ++l) {
        ConvLayer &L = c->layer[l];
        L.wf = wp;
        wp += 9ull * L.cin * L.cout;
        L.wb = wp;
        wp += 9ull * L.cin * L.cout;
    }
    S3DG_HIP(ctx, hipMemcpy(c->p, init.data(), off * 4, hipMemcpyHostToDevice));
    S3DG_HIP(ctx, hipMemset(c->m, 0, off * 4));
    S3DG_HIP(ctx, hipMemset(c->v, 0, off * 4));
    pack_weights(c);
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = c;
    return S3DG_OK;
}

extern "C" int s3dg_denoiser_features_device(s3dg_context *ctx, int32_t n, int32_t h, int32_t w,
                                             const float *d_accum, int32_t spp, const float *d_albedo,
                                             const float *d_normal, const float *d_depth, float *d_features) {
    if (!ctx || !d_accum || !d_albedo || !d_normal || !d_depth || !d_features || spp <= 0)
        return S3DG_ERR_INVALID;
    if (n <= 0 || h <= 0 || w <= 0 || (long long)n * h * w > (1LL << 27))
        return set_error(ctx, S3DG_ERR_INVALID, "denoiser features: bad size");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    const long long npix = (long long)n * h * w;
    hipLaunchKernelGGL(k_features, dim3(grid1(npix)), dim3(256), 0, ctx->stream, npix, d_accum, (float)spp,
                       d_albedo, d_normal, d_depth, d_features);
    S3DG_HIP(ctx, hipGetLastError());
    return S3DG_OK;
}

extern "C" int s3dg_cnn_destroy(s3dg_cnn *c) {
    if (!c) return S3DG_ERR_INVALID;
    for (auto &g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    void *ptrs[] = {c->p, c->g, c->m, c->v, c->wbuf, c->ws, c->d_loss};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete c;
    return S3DG_OK;
}

extern "C" int s3dg_cnn_param_count(const s3dg_cnn *c, uint64_t *n) {
    if (!c || !n) return S3DG_ERR_INVALID;
    *n = c->nparam;
    return S3DG_OK;
}

extern "C" int s3dg_cnn_get_params(s3dg_cnn *c, float *host, uint64_t n) {
    if (!c || !host || n != c->nparam) return S3DG_ERR_INVALID;
    S3DG_HIP(c->ctx, hipMemcpyAsync(host, c->p, n * 4, hipMemcpyDeviceToHost, c->ctx->stream));
    S3DG_HIP(c->ctx, hipStreamSynchronize(c->ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_cnn_set_params(s3dg_cnn *c, const float *host, uint64_t n) {
    if (!c || !host || n != c->nparam) return S3DG_ERR_INVALID;
    S3DG_HIP(c->ctx, hipMemcpyAsync(c->p, host, n * 4, hipMemcpyHostToDevice, c->ctx->stream));
    pack_weights(c);
    S3DG_HIP(c->ctx, hipStreamSynchronize(c->ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_cnn_infer_device(s3dg_cnn *c, int32_t n, int32_t h, int32_t w, const float *d_feat,
                                     float *d_out) {
    if (!c || !d_feat || !d_out) return S3DG_ERR_INVALID;
    int rc = check_dims(c->ctx, n, h, w);
    if (rc) return rc;
    S3DG_HIP(c->ctx, hipSetDevice(c->ctx->device));
    const long long n0 = (long long)n * h * w;
    Acts a;
    rc = prepare(c, n0, a);
    if (rc) return rc;
    hipStream_t s = c->ctx->stream;
    auto launch = [&]() { forward(c, a, d_feat, n, h, w, d_out); };
    // The ~20 launches of one inference as one graph launch once the same
    // shape and buffers come back (a denoiser runs frame after frame on the
    // same buffers).  The kernels read the weights through fixed pointers,
    // so a graph stays valid when training or set_params changes them.
    const uint64_t key[6] = {(uint64_t)n, (uint64_t)h, (uint64_t)w, (uint64_t)(uintptr_t)d_feat,
                             (uint64_t)(uintptr_t)d_out, (uint64_t)(uintptr_t)c->ws};
    ++c->tick;
    s3dg_cnn::InferGraph *hit = nullptr, *lru = &c->graphs[0];
    for (auto &g : c->graphs) {
        if (g.exec && memcmp(g.key, key, sizeof(key)) == 0) hit = &g;
        if (!g.exec || (lru->exec && g.used < lru->used)) lru = &g;
    }
    bool seen = false;
    for (const auto &k : c->seen) seen = seen || memcmp(k, key, sizeof(key)) == 0;
    if (hit) {
        hit->used = c->tick;
        S3DG_HIP(c->ctx, hipGraphLaunch(hit->exec, s));
    } else if (seen) {
        if (lru->exec) {
            (void)hipGraphExecDestroy(lru->exec);
            lru->exec = nullptr;
        }
        hipGraph_t g = nullptr;
        S3DG_HIP(c->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        launch();
        S3DG_HIP(c->ctx, hipStreamEndCapture(s, &g));
        const hipError_t ie = hipGraphInstantiate(&lru->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        S3DG_HIP(c->ctx, ie);
        memcpy(lru->key, key, sizeof(key));
        lru->used = c->tick;
        S3DG_HIP(c->ctx, hipGraphLaunch(lru->exec, s));
    } else {
        launch();
        memcpy(c->seen[c->seen_next], key, sizeof(key));
        c->seen_next = (c->seen_next + 1) % s3dg_cnn::kSeen;
    }
    S3DG_HIP(c->ctx, hipGetLastError());
    return S3DG_OK;
}

extern "C" int s3dg_cnn_gradient_device(s3dg_cnn *c, int32_t n, int32_t h, int32_t w, const float *d_feat,
                                        const float *d_target, double *loss, float *host_grad, uint64_t ngrad) {
    if (!c || !d_feat || !d_target) return S3DG_ERR_INVALID;
    if (host_grad && ngrad != c->nparam) return S3DG_ERR_INVALID;
    int rc = check_dims(c->ctx, n, h, w);
    if (rc) return rc;
    S3DG_HIP(c->ctx, hipSetDevice(c->ctx->device));
    rc = loss_and_grad(c, n, h, w, d_feat, d_target, loss);
    if (rc) return rc;
    if (host_grad) {
        S3DG_HIP(c->ctx, hipMemcpyAsync(host_grad, c->g, ngrad * 4, hipMemcpyDeviceToHost, c->ctx->stream));
        S3DG_HIP(c->ctx, hipStreamSynchronize(c->ctx->stream));
    }
    return S3DG_OK;
}

extern "C" int s3dg_cnn_train_step_device(s3dg_cnn *c, int32_t n, int32_t h, int32_t w, const float *d_feat,
                                          const float *d_target, float lr, double *loss) {
    if (!c || !d_feat || !d_target) return S3DG_ERR_INVALID;
    int rc = check_dims(c->ctx, n, h, w);
    if (rc) return rc;
    S3DG_HIP(c->ctx, hipSetDevice(c->ctx->device));
    rc = loss_and_grad(c, n, h, w, d_feat, d_target, loss);
    if (rc) return rc;
    c->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)c->step), c2 = 1.0f - powf(b2, (float)c->step);
    hipLaunchKernelGGL(k_adam, dim3(grid1((long long)c->nparam)), dim3(256), 0, c->ctx->stream, c->p, c->g, c->m, c->v,
                       (long long)c->nparam, lr, b1, b2, 1e-8f, c1, c2);
    pack_weights(c);
    S3DG_HIP(c->ctx, hipGetLastError());
    return S3DG_OK;
}
