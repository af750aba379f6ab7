// traverse.h -- gfx950 device code: exact ray/triangle and ray/sphere tests and
// the two-level traversal of compressed 8-wide BVHs with node-group stacks.
//
// Results are those of the reference's Scene::Intersect (src/include/s3d_scene.h:63-75):
// the reference walks its trees without clipping against the current hit, so its
// closest hit is the smallest t over all shapes under the Triangle::IsHit /
// Sphere::IsHit arithmetic, ties going to the first triangle in shape order and
// depth-first leaf order (triangles reject t >= best) while a sphere wins any
// tie it takes part in (spheres reject only t > best).  This traversal clips and
// orders children front to back for speed, and reproduces that tie order
// explicitly, so the hit it returns is the same triangle with the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "internal.h"

namespace s3dg {

constexpr float kFMaxF = 3.402823466e+38F;
constexpr float kRefHitMin = 1e-3f;   // F_HIT_MIN, src/include/s3d_math.h:45
constexpr float kRefHitMax = 1e12f;   // F_HIT_MAX, src/include/s3d_math.h:44

struct DRay {
    float ox, oy, oz;
    float dx, dy, dz;
};

struct DHit {
    float t, u, v;
    int32_t prim;    // global depth-first triangle index, -2 = sphere, -1 = none
    int32_t shape;   // top-level shape index
};

__device__ __forceinline__ DHit no_hit() {
    DHit h;
    h.t = kFMaxF;
    h.u = 0.0f;
    h.v = 0.0f;
    h.prim = -1;
    h.shape = -1;
    return h;
}

// Triangle::IsHit (src/src/s3d_triangle.cpp:73-105) with [tmin, tmax] for
// [F_HIT_MIN, F_HIT_MAX].  Every operation is the reference's, in its order,
// with IEEE division (no FMA: the library is built with -ffp-contract=off).
// Cheap sign/magnitude tests reject only candidates the exact test rejects.
__device__ __forceinline__ bool tri_test(const GTri *__restrict__ tris, uint32_t i, const DRay &r,
                                         float tmin, float tmax, float &t, float &u, float &v) {
    const float *p = reinterpret_cast<const float *>(tris + i);
    const float v0x = p[0], v0y = p[1], v0z = p[2];
    const float e0x = p[3], e0y = p[4], e0z = p[5];
    const float e1x = p[6], e1y = p[7], e1z = p[8];
    const float s1x = (r.dy * e1z) - (r.dz * e1y);
    const float s1y = (r.dz * e1x) - (r.dx * e1z);
    const float s1z = (r.dx * e1y) - (r.dy * e1x);
    const float div = (s1x * e0x) + (s1y * e0y) + (s1z * e0z);
    if (fabsf(div) <= FLT_EPSILON) return false;
    const float qx = r.ox - v0x, qy = r.oy - v0y, qz = r.oz - v0z;
    const float nb = (qx * s1x) + (qy * s1y) + (qz * s1z);
    const uint32_t sdiv = __float_as_uint(div) & 0x80000000u;
    // beta = nb/div <= 0 when the signs differ or nb is zero; >= 1 when |nb| >= |div|
    if (nb == 0.0f || (__float_as_uint(nb) & 0x80000000u) != sdiv) return false;
    const float adiv = fabsf(div);
    if (fabsf(nb) >= adiv) return false;
    const float s2x = (qy * e0z) - (qz * e0y);
    const float s2y = (qz * e0x) - (qx * e0z);
    const float s2z = (qx * e0y) - (qy * e0x);
    const float ng = (r.dx * s2x) + (r.dy * s2y) + (r.dz * s2z);
    if (ng == 0.0f || (__float_as_uint(ng) & 0x80000000u) != sdiv) return false;
    // beta + gamma >= 1 for certain (margin far above the rounding of the sum)
    if (fabsf(nb) + fabsf(ng) >= adiv * 1.0000010f) return false;
    const float beta = nb / div;
    if (beta <= 0.0f || beta >= 1.0f) return false;
    const float gamma = ng / div;
    if (gamma <= 0.0f || (beta + gamma) >= 1.0f) return false;
    const float dist = ((e1x * s2x) + (e1y * s2y) + (e1z * s2z)) / div;
    if (dist < tmin || dist > tmax) return false;
    t = dist;
    u = beta;
    v = gamma;
    return true;
}

// Sphere::IsHit (src/src/s3d_sphere.cpp:66-92): returns the candidate distance
// or -1 when the sphere is missed.
__device__ __forceinline__ float sphere_test(float cx, float cy, float cz, float radius, const DRay &r,
                                             float tmin) {
    const float px = cx - r.ox, py = cy - r.oy, pz = cz - r.oz;
    const float b = (px * r.dx) + (py * r.dy) + (pz * r.dz);
    const float D4 = b * b - ((px * px) + (py * py) + (pz * pz)) + radius * radius;
    if (D4 < 0.0f) return -1.0f;
    const float sq = sqrtf(D4);
    const float t1 = b - sq;
    const float t2 = b + sq;
    if (t1 < tmin && t2 < tmin) return -1.0f;
    return (t1 > tmin) ? t1 : t2;
}

// Instance::IsHit ray (src/src/s3d_instance.cpp:65-72): TransformCoord and
// TransformNormal by the inverse world matrix (src/include/s3d_math.h:2098-2125),
// then UnitVector.
__device__ __forceinline__ DRay to_local(const GInst &in, const DRay &r) {
    const float(*m)[4] = in.inv;
    const float X = (r.ox * m[0][0]) + (r.oy * m[1][0]) + (r.oz * m[2][0]) + m[3][0];
    const float Y = (r.ox * m[0][1]) + (r.oy * m[1][1]) + (r.oz * m[2][1]) + m[3][1];
    const float Z = (r.ox * m[0][2]) + (r.oy * m[1][2]) + (r.oz * m[2][2]) + m[3][2];
    const float W = (r.ox * m[0][3]) + (r.oy * m[1][3]) + (r.oz * m[2][3]) + m[3][3];
    const float nx = (r.dx * m[0][0]) + (r.dy * m[1][0]) + (r.dz * m[2][0]);
    const float ny = (r.dx * m[0][1]) + (r.dy * m[1][1]) + (r.dz * m[2][1]);
    const float nz = (r.dx * m[0][2]) + (r.dy * m[1][2]) + (r.dz * m[2][2]);
    const float mag = sqrtf(nx * nx + ny * ny + nz * nz);
    DRay l;
    l.ox = X / W;
    l.oy = Y / W;
    l.oz = Z / W;
    l.dx = nx / mag;
    l.dy = ny / mag;
    l.dz = nz / mag;
    return l;
}

// Culling bounds with slack: a box is skipped only when no triangle inside it
// can produce an accepted distance (boxes are also grown on the host).
__device__ __forceinline__ float cull_hi(float t) { return t + fabsf(t) * 1e-4f + 1e-6f; }
__device__ __forceinline__ float cull_lo(float t) { return t - fabsf(t) * 1e-4f - 1e-6f; }

// Bit permutation taking slot order to visiting order k = slot ^ oct.
__device__ __forceinline__ uint32_t to_korder(uint32_t m, uint32_t oct) {
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
    if (oct & 4u) m = ((m & 0x0fu) << 4) | ((m >> 4) & 0x0fu);
    return m;
}

struct RayPre {
    float ix, iy, iz;
    uint32_t oct;
};

__device__ __forceinline__ RayPre precompute(const DRay &r) {
    RayPre p;
    const float tiny = 1e-30f;
    const float dx = fabsf(r.dx) < tiny ? copysignf(tiny, r.dx) : r.dx;
    const float dy = fabsf(r.dy) < tiny ? copysignf(tiny, r.dy) : r.dy;
    const float dz = fabsf(r.dz) < tiny ? copysignf(tiny, r.dz) : r.dz;
    p.ix = __builtin_amdgcn_rcpf(dx);
    p.iy = __builtin_amdgcn_rcpf(dy);
    p.iz = __builtin_amdgcn_rcpf(dz);
    p.oct = (p.ix < 0.0f ? 1u : 0u) | (p.iy < 0.0f ? 2u : 0u) | (p.iz < 0.0f ? 4u : 0u);
    return p;
}

enum { ENT_GROUP = 0u, ENT_INST = 1u, ENT_RESTORE = 2u };

// Traversal stack: the first kLdsStack entries of a lane live in LDS (stride
// kBlock), the rest in a global overflow area (stride gstride).
struct Stack {
    uint2 *lds;
    uint2 *gst;
    uint32_t gstride;
    int32_t gcap;
    int32_t *err;
    int32_t sp;

    __device__ __forceinline__ void push(uint2 e) {
        if (sp < kLdsStack) {
            lds[sp * kBlock] = e;
        } else if (sp - kLdsStack < gcap) {
            gst[(size_t)(sp - kLdsStack) * gstride] = e;
        } else {
            atomicOr(err, 1);
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ uint2 top() const {
        const int32_t i = sp - 1;
        return i < kLdsStack ? lds[i * kBlock] : gst[(size_t)(i - kLdsStack) * gstride];
    }
    __device__ __forceinline__ void set_top(uint2 e) {
        const int32_t i = sp - 1;
        if (i < kLdsStack) lds[i * kBlock] = e;
        else gst[(size_t)(i - kLdsStack) * gstride] = e;
    }
};

// Closest hit (ANY = false) or first hit found (ANY = true) in [tmin, tmax].
template <bool ANY, bool COUNT>
__device__ __forceinline__ void traverse(const SceneDev &S, const DRay &wray, float tmin, float tmax,
                                         DHit &best, Stack &st, unsigned long long *counters) {
    DRay ray = wray;
    RayPre pre = precompute(ray);
    int32_t cur_shape = S.single_shape;
    uint32_t node = S.root;
    bool have_node = true;
    const float lo_cull = cull_lo(tmin);
    float bound = fminf(cull_hi(tmax), cull_hi(best.t));
    uint32_t visits = 0, tests = 0;

    for (;;) {
        if (!have_node) {
            if (st.sp == 0) break;
            uint2 e = st.top();
            const uint32_t kind = e.y >> 30;
            if (kind == ENT_GROUP) {
                const uint32_t imask = e.y & 0xffu;
                uint32_t rem = (e.y >> 8) & 0xffu;
                const uint32_t k0 = __builtin_ctz(rem);
                const uint32_t s0 = k0 ^ pre.oct;
                node = e.x + __builtin_popcount(imask & ((1u << s0) - 1u));
                rem &= rem - 1u;
                if (rem) st.set_top(make_uint2(e.x, (e.y & ~0xff00u) | (rem << 8)));
                else st.sp--;
                have_node = true;
            } else if (kind == ENT_INST) {
                st.sp--;
                const GInst &in = S.insts[e.x];
                if (in.type == INST_SPHERE) {
                    const float d = sphere_test(in.center[0], in.center[1], in.center[2], in.radius, ray, tmin);
                    if (d >= 0.0f && d <= tmax) {
                        const bool take = ANY ? true
                                              : (d < best.t || (d == best.t && (best.prim != -2 || in.shape > best.shape)));
                        if (take) {
                            best.t = d;
                            best.u = 0.0f;
                            best.v = 0.0f;
                            best.prim = -2;
                            best.shape = in.shape;
                            if (ANY) return;
                            bound = fminf(cull_hi(tmax), cull_hi(best.t));
                        }
                    }
                } else {
                    cur_shape = in.shape;
                    if (in.type == INST_TRANSFORMED) {
                        st.push(make_uint2(0u, ENT_RESTORE << 30));
                        ray = to_local(in, wray);
                        pre = precompute(ray);
                    }
                    node = in.blas_root;
                    have_node = true;
                }
            } else {   // ENT_RESTORE
                st.sp--;
                ray = wray;
                pre = precompute(ray);
            }
            continue;
        }

        // ---- visit one compressed node --------------------------------
        const uint4 *np = reinterpret_cast<const uint4 *>(S.nodes + node);
        const uint4 h0 = np[0];
        const uint4 h1 = np[1];
        const uint4 q0 = np[4];
        const uint4 q1 = np[5];
        const uint4 q2 = np[6];
        if (COUNT) ++visits;
        const float ox = __uint_as_float(h0.x), oy = __uint_as_float(h0.y), oz = __uint_as_float(h0.z);
        const uint32_t imask = h0.w >> 24;
        const float sx = __uint_as_float((h0.w & 0xffu) << 23);
        const float sy = __uint_as_float(((h0.w >> 8) & 0xffu) << 23);
        const float sz = __uint_as_float(((h0.w >> 16) & 0xffu) << 23);
        const float ax = sx * pre.ix, ay = sy * pre.iy, az = sz * pre.iz;
        const float bx = (ox - ray.ox) * pre.ix, by = (oy - ray.oy) * pre.iy, bz = (oz - ray.oz) * pre.iz;
        const uint32_t child_base = h1.x;
        const uint32_t cnt_lo = h1.y, cnt_hi = h1.z;
        const uint32_t top = h1.w & 1u;
        uint32_t hits = 0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int sh = (s & 3) * 8;
            const uint32_t w0 = s < 4 ? q0.x : q0.y;   // qlo x
            const uint32_t w1 = s < 4 ? q0.z : q0.w;   // qlo y
            const uint32_t w2 = s < 4 ? q1.x : q1.y;   // qlo z
            const uint32_t w3 = s < 4 ? q1.z : q1.w;   // qhi x
            const uint32_t w4 = s < 4 ? q2.x : q2.y;   // qhi y
            const uint32_t w5 = s < 4 ? q2.z : q2.w;   // qhi z
            const float tlx = fmaf((float)((w0 >> sh) & 0xffu), ax, bx);
            const float thx = fmaf((float)((w3 >> sh) & 0xffu), ax, bx);
            const float tly = fmaf((float)((w1 >> sh) & 0xffu), ay, by);
            const float thy = fmaf((float)((w4 >> sh) & 0xffu), ay, by);
            const float tlz = fmaf((float)((w2 >> sh) & 0xffu), az, bz);
            const float thz = fmaf((float)((w5 >> sh) & 0xffu), az, bz);
            const float tn = fmaxf(fmaxf(fminf(tlx, thx), fminf(tly, thy)), fmaxf(fminf(tlz, thz), lo_cull));
            const float tf = fminf(fminf(fmaxf(tlx, thx), fmaxf(tly, thy)), fminf(fmaxf(tlz, thz), bound));
            if (tn <= tf) hits |= 1u << s;
        }
        // valid slots: internal children or leaves with a non-zero count
        uint32_t valid = imask;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t c = ((s < 4 ? cnt_lo : cnt_hi) >> ((s & 3) * 8)) & 0xffu;
            valid |= (c != 0u ? 1u : 0u) << s;
        }
        hits &= valid;
        uint32_t leaves = hits & ~imask;
        if (leaves) {
            if (top) {
                // top-level leaves are single shapes: queued, visited in turn
                uint32_t lk = to_korder(leaves, pre.oct);
                // push far ones first so the nearest is popped first
                while (lk) {
                    const uint32_t k = 31u - __builtin_clz(lk);
                    lk &= ~(1u << k);
                    const uint32_t s = k ^ pre.oct;
                    st.push(make_uint2(reinterpret_cast<const uint32_t *>(np)[8 + s], ENT_INST << 30));
                }
            } else {
                while (leaves) {
                    const uint32_t s = __builtin_ctz(leaves);
                    leaves &= leaves - 1u;
                    const uint32_t start = reinterpret_cast<const uint32_t *>(np)[8 + s];
                    const uint32_t c = ((s < 4 ? cnt_lo : cnt_hi) >> ((s & 3) * 8)) & 0xffu;
                    for (uint32_t i = start; i < start + c; ++i) {
                        float t, u, v;
                        if (COUNT) ++tests;
                        if (!tri_test(S.tris, i, ray, tmin, tmax, t, u, v)) continue;
                        if (ANY) {
                            best.t = t;
                            best.u = u;
                            best.v = v;
                            best.prim = (int32_t)i;
                            best.shape = cur_shape;
                            if (COUNT) { atomicAdd(&counters[0], (unsigned long long)visits); atomicAdd(&counters[1], (unsigned long long)tests); }
                            return;
                        }
                        const bool take = t < best.t ||
                                          (t == best.t && best.prim >= 0 &&
                                           (cur_shape < best.shape || (cur_shape == best.shape && (int32_t)i < best.prim)));
                        if (take) {
                            best.t = t;
                            best.u = u;
                            best.v = v;
                            best.prim = (int32_t)i;
                            best.shape = cur_shape;
                            bound = fminf(cull_hi(tmax), cull_hi(best.t));
                        }
                    }
                }
            }
        }
        const uint32_t inner = to_korder(hits & imask, pre.oct);
        if (inner) {
            const uint32_t k0 = __builtin_ctz(inner);
            const uint32_t s0 = k0 ^ pre.oct;
            const uint32_t rem = inner & (inner - 1u);
            if (rem) st.push(make_uint2(child_base, (ENT_GROUP << 30) | (rem << 8) | imask));
            node = child_base + __builtin_popcount(imask & ((1u << s0) - 1u));
            have_node = true;
        } else {
            have_node = false;
        }
    }
    if (COUNT) { atomicAdd(&counters[0], (unsigned long long)visits); atomicAdd(&counters[1], (unsigned long long)tests); }
}

}  // namespace s3dg
