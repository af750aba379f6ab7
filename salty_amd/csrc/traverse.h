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
    l.ox = X;
    l.oy = Y;
    l.oz = Z;
    // an affine matrix gives W = 1 exactly, and X / 1 is X
    if (W != 1.0f) {
        l.ox = X / W;
        l.oy = Y / W;
        l.oz = Z / W;
    }
    l.dx = nx / mag;
    l.dy = ny / mag;
    l.dz = nz / mag;
    return l;
}

// Culling bounds with slack: a box is skipped only when no triangle inside it
// can produce an accepted distance (boxes are also grown on the host).
__device__ __forceinline__ float cull_hi(float t) { return t + fabsf(t) * 1e-4f + 1e-6f; }
__device__ __forceinline__ float cull_lo(float t) { return t - fabsf(t) * 1e-4f - 1e-6f; }

// Rounding bound of a slab distance, in distance units, for one axis of a
// node: a plane at origin + q * 2^e (q <= 255) gives t = fma(q, a, b) with
// a = 2^e * inv (exact), b = (origin - o) * inv and inv the approximate
// reciprocal of the direction (v_rcp_f32, 1 ulp).  With D = origin - o,
// E = q * 2^e, the computed t differs from (D + E) / d by at most
// 2.5 * 2^-23 * (|D| + |E|) / |d| (the reciprocal, the subtraction, the
// product and the fma).  slab_scale = |b| + 255 |a| bounds
// (|D| + |E|) * |inv|, and kSlabSlack = 2^-20 is 3.2 times the bound (also
// covering the rounding of the shifted b): entry distances are computed from
// b - kSlabSlack * scale, exit distances from b + kSlabSlack * scale, so no
// box holding an accepted hit is rejected, and the boxes themselves are not
// grown.  A ray leaving a flat surface almost parallel to it (a grazing
// sample off the ground plane) then no longer meets every leaf box of that
// surface, as it did with boxes grown by a fixed fraction of the scene size.
constexpr float kSlabSlack = 0x1p-20f;
__device__ __forceinline__ float slab_scale(float a, float b) { return fmaf(255.0f, fabsf(a), fabsf(b)); }

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

// Stack entries: a group of a node's hit inner children still to visit, a
// group of a top-level node's hit shapes still to enter, and a group of a
// node's hit leaves whose triangles are still to test.  A top-level node thus
// takes at most two entries, like any other node.  Entries a top-level node
// pushes carry kWorldEntry: popping one while the ray is an instance's local
// ray first returns to the world ray (in the same step, so leaving an
// instance costs no step of its own).
enum { ENT_GROUP = 0u, ENT_INST = 1u, ENT_LEAVES = 3u };
constexpr uint32_t kWorldEntry = 1u << 29;

// Traversal stack: the first LDS entries of a lane live in LDS (stride
// kBlock), the rest in a global overflow area (stride gstride).  The pointers
// carry their address spaces so that the two halves are never merged into one
// flat access (which would wait on both the LDS and the vector-memory counters).
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(1))) uint64_t glb_u64;

__device__ __forceinline__ uint64_t pack_entry(uint2 e) { return (uint64_t)e.x | ((uint64_t)e.y << 32); }
__device__ __forceinline__ uint2 unpack_entry(uint64_t v) { return make_uint2((uint32_t)v, (uint32_t)(v >> 32)); }

template <int LDS>
struct Stack {
    lds_u64 *lds;
    glb_u64 *gst;
    uint32_t gstride;
    int32_t gcap;
    int32_t *err;
    int32_t sp;
    uint32_t spills;   // pushes that went to the global area

    __device__ __forceinline__ void push(uint2 e) {
        if (sp < LDS) {
            lds[sp * kBlock] = pack_entry(e);
        } else if (sp - LDS < gcap) {
            gst[(size_t)(sp - LDS) * gstride] = pack_entry(e);
            ++spills;
        } else {
            atomicOr(err, 1);
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ uint2 top() const {
        const int32_t i = sp - 1;
        if (i < LDS) return unpack_entry(lds[i * kBlock]);
        uint64_t v = gst[(size_t)(i - LDS) * gstride];
        // keeps the two loads apart (an empty asm that names the loaded value)
        __asm__ volatile("" : "+v"(v));
        return unpack_entry(v);
    }
    __device__ __forceinline__ void set_top(uint2 e) {
        const int32_t i = sp - 1;
        if (i < LDS) lds[i * kBlock] = pack_entry(e);
        else gst[(size_t)(i - LDS) * gstride] = pack_entry(e);
    }
    // Lane `lane` of the block's stacks over the shared array `base`.
    __device__ __forceinline__ void init(uint2 *base, uint32_t lane, uint2 *global_base, uint32_t global_stride,
                                         int32_t global_cap, int32_t *error) {
        lds = (lds_u64 *)(base + lane);
        gst = (glb_u64 *)global_base;
        gstride = global_stride;
        gcap = global_cap;
        err = error;
        sp = 0;
        spills = 0;
    }
    // Adds the wave's global-area pushes to *counter (one atomic per wave);
    // every lane of the wave must call it.
    __device__ __forceinline__ void flush_spills(unsigned long long *counter) const {
        unsigned long long v = spills;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (__lane_id() == 0 && v) atomicAdd(counter, v);
    }
};

// First triangle (or shape record) of leaf slot s of a node: leaf_base plus
// the counts of the slots below s (count bytes of slots 0-3 in lo, 4-7 in hi;
// inner slots count 0).
__device__ __forceinline__ uint32_t slot_first(uint32_t leaf_base, uint32_t lo, uint32_t hi, uint32_t s) {
    const uint32_t mlo = s >= 4u ? lo : (lo & ((1u << (8u * s)) - 1u));
    const uint32_t mhi = s <= 4u ? 0u : (hi & ((1u << (8u * (s - 4u))) - 1u));
    return __builtin_amdgcn_sad_u8(mlo, 0u, __builtin_amdgcn_sad_u8(mhi, 0u, leaf_base));
}

// Bit b (b < 4) set when byte b of w is non-zero: bit 7 of each byte of t
// is set iff the byte is non-zero, and the multiply gathers bits 7, 15, 23,
// 31 into bits 28..31 (the other partial products land on distinct lower
// bits, so no carry reaches them).
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t w) {
    const uint32_t t = ((w & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w;
    return ((t & 0x80808080u) * 0x00204081u) >> 28;
}

__device__ __forceinline__ uint32_t slot_count(uint32_t lo, uint32_t hi, uint32_t s) {
    return ((s < 4u ? lo : hi) >> ((s & 3u) * 8u)) & 0xffu;
}

// Incremental traversal of one ray.  start() sets the ray up; each step()
// tests at most one pending triangle (more while most lanes of the wave have
// triangles pending), visits at most one node and pops a stack entry when the
// lane has neither, and returns true once the ray is finished (its result is
// then in `best`).  Triangle tests and node visits of different lanes thus
// share the wave's instructions instead of waiting for each other; kernels
// call step() from persistent loops that hand finished lanes new rays.
// ANY = false: closest hit in [tmin, tmax]; ANY = true: first hit found.
template <bool ANY, bool COUNT>
struct Tracer {
    DRay wray;          // world-space ray
    DRay ray;           // ray of the current tree (instance space inside an instance)
    RayPre pre;
    DHit best;
    float tmin, tmax, lo_cull, bound;
    uint32_t node;       // next node to visit (when have_node)
    uint32_t tri_cur, tri_end;   // triangles of the current leaf still to test
    uint32_t leaf_mask;          // further hit leaf slots of the node below
    uint32_t leaf_base, leaf_lo, leaf_hi;   // that node's leaf_base and count words
    int32_t cur_shape;
    bool have_node;
    bool in_local;       // `ray` is an instance's local ray
    uint32_t visits, tests;

    // Starts a ray at the root of scene `rec`.
    __device__ __forceinline__ void start(const SceneRec &rec, const DRay &r, float t0, float t1) {
        wray = r;
        ray = r;
        pre = precompute(r);
        best = no_hit();
        tmin = t0;
        tmax = t1;
        lo_cull = cull_lo(t0);
        bound = cull_hi(t1);
        node = rec.root;
        cur_shape = rec.single_shape;
        have_node = true;
        in_local = false;
        tri_cur = tri_end = 0;
        leaf_mask = 0;
        leaf_base = leaf_lo = leaf_hi = 0;
    }

    __device__ __forceinline__ void take(float t, float u, float v, int32_t prim, int32_t shape) {
        best.t = t;
        best.u = u;
        best.v = v;
        best.prim = prim;
        best.shape = shape;
        bound = fminf(cull_hi(tmax), cull_hi(t));
    }

    // Moves to the triangles of the next pending leaf slot.
    __device__ __forceinline__ void next_leaf() {
        const uint32_t s = (uint32_t)__builtin_ctz(leaf_mask);
        leaf_mask &= leaf_mask - 1u;
        tri_cur = slot_first(leaf_base, leaf_lo, leaf_hi, s);
        tri_end = tri_cur + slot_count(leaf_lo, leaf_hi, s);
    }

    // Tests triangle tri_cur.  Returns true when an any-hit search is over.
    __device__ __forceinline__ bool test_triangle(const SceneDev &S) {
        const uint32_t i = tri_cur++;
        float t, u, v;
        if (COUNT) ++tests;
        if (tri_test(S.tris, i, ray, tmin, tmax, t, u, v)) {
            if (ANY) {
                take(t, u, v, (int32_t)i, cur_shape);
                return true;
            }
            // the first triangle in shape order, then depth-first order, wins a tie
            const bool win = t < best.t ||
                             (t == best.t && best.prim >= 0 &&
                              (cur_shape < best.shape ||
                               (cur_shape == best.shape && S.tri_rank[i] < S.tri_rank[best.prim])));
            if (win) take(t, u, v, (int32_t)i, cur_shape);
        }
        if (tri_cur == tri_end && leaf_mask) next_leaf();
        return false;
    }

    // Pops one entry: a node group yields its next child, an instance entry
    // enters the instance (or tests the sphere), a leaf group becomes the
    // pending triangles; a top-level entry popped inside an instance returns
    // to the world ray first.  Returns true when an any-hit search is over.
    template <class ST>
    __device__ __forceinline__ bool pop(const SceneDev &S, ST &st) {
        const uint2 e = st.top();
        const uint32_t kind = e.y >> 30;
        if ((e.y & kWorldEntry) && in_local) {
            in_local = false;
            ray = wray;
            pre = precompute(ray);
        }
        if (kind == ENT_GROUP) {
            const uint32_t imask = e.y & 0xffu;
            uint32_t rem = (e.y >> 8) & 0xffu;
            const uint32_t s0 = (uint32_t)__builtin_ctz(rem) ^ pre.oct;
            node = e.x + (uint32_t)__builtin_popcount(imask & ((1u << s0) - 1u));
            rem &= rem - 1u;
            if (rem) st.set_top(make_uint2(e.x, (e.y & ~0xff00u) | (rem << 8)));
            else st.sp--;
            have_node = true;
        } else if (kind == ENT_LEAVES) {
            st.sp--;
            const uint4 h1 = reinterpret_cast<const uint4 *>(S.nodes + e.x)[1];
            leaf_base = h1.y;
            leaf_lo = h1.z;
            leaf_hi = h1.w;
            leaf_mask = e.y & 0xffu;
            next_leaf();
        } else if (kind == ENT_INST) {
            uint32_t rem = e.y & 0xffu;
            const uint32_t s = (uint32_t)__builtin_ctz(rem) ^ pre.oct;
            const uint32_t lslots = (e.y >> 8) & 0xffu;
            const uint32_t rec = e.x + (uint32_t)__builtin_popcount(lslots & ((1u << s) - 1u));
            rem &= rem - 1u;
            if (rem) st.set_top(make_uint2(e.x, (e.y & ~0xffu) | rem));
            else st.sp--;
            const GInst &in = S.insts[rec];
            if (in.type == INST_SPHERE) {
                const float d = sphere_test(in.center[0], in.center[1], in.center[2], in.radius, ray, tmin);
                if (d >= 0.0f && d <= tmax) {
                    // a sphere takes any tie (Sphere::IsHit rejects only t > best)
                    const bool win = ANY || d < best.t || (d == best.t && (best.prim != -2 || in.shape > best.shape));
                    if (win) {
                        take(d, 0.0f, 0.0f, -2, in.shape);
                        if (ANY) return true;
                    }
                }
            } else {
                cur_shape = in.shape;
                if (in.type == INST_TRANSFORMED) {
                    in_local = true;
                    ray = to_local(in, wray);
                    pre = precompute(ray);
                }
                node = in.blas_root;
                have_node = true;
            }
        }
        return false;
    }

    // Visits `node`: slab tests of its eight children; the hit leaves become
    // the pending triangles (or a stacked leaf group when triangles are still
    // pending), the nearest hit inner child is visited next and the others are
    // pushed as one group.
    // The node's header lines (origin, exponents, masks, child and leaf bases,
    // counts), requested before a step's triangle tests; the plane lines are
    // loaded by visit() itself.
    struct NodeLines {
        uint4 h0, h1;
    };

    __device__ __forceinline__ NodeLines fetch(const SceneDev &S) const {
        const uint4 *np = reinterpret_cast<const uint4 *>(S.nodes + node);
        NodeLines n;
        n.h0 = np[0];
        n.h1 = np[1];
        return n;
    }

    // Visits `node`: slab tests of its eight children; the hit leaves become
    // the pending triangles (or a stacked leaf group when triangles are still
    // pending), the nearest hit inner child is visited next and the others are
    // pushed as one group.
    template <class ST>
    __device__ __forceinline__ void visit(const SceneDev &S, ST &st, const NodeLines &nl) {
        const uint4 h0 = nl.h0, h1 = nl.h1;
        const uint4 *np = reinterpret_cast<const uint4 *>(S.nodes + node);
        const uint4 q0 = np[2], q1 = np[3], q2 = np[4];   // planes: lo x, lo y | lo z, hi x | hi y, hi z
        if (COUNT) ++visits;
        const float ox = __uint_as_float(h0.x), oy = __uint_as_float(h0.y), oz = __uint_as_float(h0.z);
        const uint32_t imask = h0.w >> 24;
        const float sx = __uint_as_float((h0.w & 0xffu) << 23);
        const float sy = __uint_as_float(((h0.w >> 8) & 0xffu) << 23);
        const float sz = __uint_as_float(((h0.w >> 16) & 0xffu) << 23);
        const float ax = sx * pre.ix, ay = sy * pre.iy, az = sz * pre.iz;
        const float bx = (ox - ray.ox) * pre.ix, by = (oy - ray.oy) * pre.iy, bz = (oz - ray.oz) * pre.iz;
        // entry distances lowered and exit distances raised by a bound on
        // their rounding (kSlabSlack)
        const float sx_ = slab_scale(ax, bx), sy_ = slab_scale(ay, by), sz_ = slab_scale(az, bz);
        const float bnx = fmaf(-kSlabSlack, sx_, bx), bux = fmaf(kSlabSlack, sx_, bx);
        const float bny = fmaf(-kSlabSlack, sy_, by), buy = fmaf(kSlabSlack, sy_, by);
        const float bnz = fmaf(-kSlabSlack, sz_, bz), buz = fmaf(kSlabSlack, sz_, bz);
        const uint32_t top = h1.x & kTopLevelBit;
        const uint32_t child_base = h1.x & ~kTopLevelBit;
        const uint32_t lbase = h1.y, cnt_lo = h1.z, cnt_hi = h1.w;
        // near / far planes per axis from the direction signs: for a positive
        // direction component the low plane is entered first
        const bool fx = (pre.oct & 1u) != 0, fy = (pre.oct & 2u) != 0, fz = (pre.oct & 4u) != 0;
        const uint32_t nx0 = fx ? q1.z : q0.x, nx1 = fx ? q1.w : q0.y;
        const uint32_t ux0 = fx ? q0.x : q1.z, ux1 = fx ? q0.y : q1.w;
        const uint32_t ny0 = fy ? q2.x : q0.z, ny1 = fy ? q2.y : q0.w;
        const uint32_t uy0 = fy ? q0.z : q2.x, uy1 = fy ? q0.w : q2.y;
        const uint32_t nz0 = fz ? q2.z : q1.x, nz1 = fz ? q2.w : q1.y;
        const uint32_t uz0 = fz ? q1.x : q2.z, uz1 = fz ? q1.y : q2.w;
        uint32_t hits = 0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int sh = (s & 3) * 8;
            const uint32_t wnx = s < 4 ? nx0 : nx1, wux = s < 4 ? ux0 : ux1;
            const uint32_t wny = s < 4 ? ny0 : ny1, wuy = s < 4 ? uy0 : uy1;
            const uint32_t wnz = s < 4 ? nz0 : nz1, wuz = s < 4 ? uz0 : uz1;
            const float tnx = fmaf((float)((wnx >> sh) & 0xffu), ax, bnx);
            const float tux = fmaf((float)((wux >> sh) & 0xffu), ax, bux);
            const float tny = fmaf((float)((wny >> sh) & 0xffu), ay, bny);
            const float tuy = fmaf((float)((wuy >> sh) & 0xffu), ay, buy);
            const float tnz = fmaf((float)((wnz >> sh) & 0xffu), az, bnz);
            const float tuz = fmaf((float)((wuz >> sh) & 0xffu), az, buz);
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, lo_cull));
            const float tf = fminf(fminf(tux, tuy), fminf(tuz, bound));
            if (tn <= tf) hits |= 1u << s;
        }
        // valid slots: internal children or leaves with a non-zero count
        const uint32_t valid = imask | nonzero_bytes(cnt_lo) | (nonzero_bytes(cnt_hi) << 4);
        hits &= valid;
        const uint32_t leaves = hits & ~imask;
        if (leaves) {
            if (top) {
                // the node's hit shape leaves as one entry, entered nearest
                // first; each top-level leaf slot holds one shape record, so
                // a slot's record is lbase plus the leaf slots below it
                const uint32_t lslots = valid & ~imask;
                st.push(make_uint2(lbase, (ENT_INST << 30) | kWorldEntry | (lslots << 8) | to_korder(leaves, pre.oct)));
            } else if (tri_cur == tri_end && leaf_mask == 0) {
                leaf_base = lbase;
                leaf_lo = cnt_lo;
                leaf_hi = cnt_hi;
                leaf_mask = leaves;
                next_leaf();
            } else {
                st.push(make_uint2(node, (ENT_LEAVES << 30) | leaves));
            }
        }
        const uint32_t inner = to_korder(hits & imask, pre.oct);
        if (inner) {
            const uint32_t s0 = (uint32_t)__builtin_ctz(inner) ^ pre.oct;
            const uint32_t rem = inner & (inner - 1u);
            static_assert((kTopLevelBit >> 2) == kWorldEntry, "a top-level node's group entry is a world entry");
            if (rem) st.push(make_uint2(child_base, (ENT_GROUP << 30) | (top >> 2) | (rem << 8) | imask));
            node = child_base + (uint32_t)__builtin_popcount(imask & ((1u << s0) - 1u));
            have_node = true;
        } else {
            have_node = false;
        }
    }

    // One step; true when the ray is finished.  The next node's lines are
    // requested before the pending triangles are tested, so that the two
    // memory round trips overlap.
    template <class ST>
    __device__ __forceinline__ bool step(const SceneDev &S, ST &st) {
        bool tris = tri_cur < tri_end;
        if (!have_node && !tris) {
            if (pop(S, st)) return true;
            tris = tri_cur < tri_end;
        }
        NodeLines nl;
        if (have_node) nl = fetch(S);
        // triangle tests: one per step, more while most of the wave has triangles pending
        for (;;) {
            if (tris) {
                if (test_triangle(S)) return true;
                tris = tri_cur < tri_end;
            }
            if ((uint32_t)__popcll(__ballot(tris)) < S.tri_batch) break;
        }
        if (have_node) {
            // keeps the compiler from evaluating node fields before the
            // triangle tests (which would wait for the node lines there)
            __asm__ volatile("" : "+v"(nl.h0.w), "+v"(nl.h1.x));
            visit(S, st, nl);
        }
        return !have_node && tri_cur >= tri_end && leaf_mask == 0 && st.sp == 0;
    }

    // Work counters (COUNT): visits and tests accumulate over every ray the
    // lane traces; flush() adds the wave's totals with one atomic each and
    // must be reached by all lanes of the wave.
    __device__ __forceinline__ void reset_counts() {
        visits = 0;
        tests = 0;
    }
    __device__ __forceinline__ void flush(unsigned long long *counters) const {
        if (COUNT) {
            unsigned long long v = visits, t = tests;
            for (int off = 32; off > 0; off >>= 1) {
                v += __shfl_xor(v, off);
                t += __shfl_xor(t, off);
            }
            if (__lane_id() == 0) {
                atomicAdd(&counters[0], v);
                atomicAdd(&counters[1], t);
            }
        }
    }
};

// Whole traversal of one ray (for kernels without lane refill).
template <bool ANY, bool COUNT, class ST>
__device__ __forceinline__ DHit traverse(const SceneDev &S, const SceneRec &rec, const DRay &r, float tmin,
                                         float tmax, ST &st, unsigned long long *counters) {
    Tracer<ANY, COUNT> tr;
    tr.reset_counts();
    st.sp = 0;
    tr.start(rec, r, tmin, tmax);
    while (!tr.step(S, st)) {
    }
    if (COUNT) {
        atomicAdd(&counters[0], (unsigned long long)tr.visits);
        atomicAdd(&counters[1], (unsigned long long)tr.tests);
    }
    return tr.best;
}

// Persistent-kernel work distribution with lane refill.  Each wave owns a
// chunk of queue entries [cb, ce): its first chunk is fixed by its index in
// the grid, later ones are taken from the global cursor `chunk` at a time
// (past the first chunks of all waves); whenever at least `refill` lanes are
// idle they get the next entries of the chunk (the chunk is renewed with one
// atomic when it runs short), so the cursor sees one atomic per chunk instead
// of one per refill, and a launch with little work sees none at all.
// `begin(i)` sets lane state up for queue entry i and returns false when the
// entry needs no traversal; `finish()` consumes the result.  The wave leaves
// once the cursor has passed n, its chunk is used up and no lane is active.
struct FetchParams {
    uint32_t refill;   // idle lanes that trigger a refill (1..64)
    uint32_t chunk;    // entries taken from the global cursor at a time (>= 64)
};

template <class Begin, class Step, class Finish>
__device__ __forceinline__ void refill_loop(uint32_t *cursor, uint32_t n, FetchParams fp, Begin begin, Step step,
                                            Finish finish) {
    bool active = false;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    // first chunks: wave w owns [w * chunk, (w + 1) * chunk); the cursor hands
    // out entries from waves * chunk on (64-bit: the product can pass 2^32)
    const uint64_t first_end = (uint64_t)waves * fp.chunk;
    const uint32_t dyn0 = first_end < n ? (uint32_t)first_end : n;
    uint32_t cb = (uint64_t)wave * fp.chunk < n ? wave * fp.chunk : n;   // this wave's chunk (entries < n only)
    uint32_t ce = n - cb < fp.chunk ? n : cb + fp.chunk;
    bool exhausted = dyn0 >= n;   // the global cursor has passed n
    const uint64_t below = (1ull << __lane_id()) - 1ull;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const uint32_t want = (uint32_t)__popcll(idle);
        const bool drained = exhausted && cb >= ce;
        if (!drained && want >= fp.refill) {
            const uint32_t avail = ce - cb;
            uint32_t nb = 0, ne = 0;
            if (avail < want && !exhausted) {
                uint32_t base = 0;
                if (__lane_id() == 0) base = atomicAdd(cursor, fp.chunk);
                // 64-bit: the cursor keeps counting past n by up to a chunk per wave
                const uint64_t nb64 = (uint64_t)dyn0 + __shfl(base, 0);
                nb = nb64 < n ? (uint32_t)nb64 : n;
                ne = n - nb < fp.chunk ? n : nb + fp.chunk;
                exhausted = nb64 + fp.chunk >= n;
            }
            if (!active) {
                const uint32_t r = (uint32_t)__popcll(idle & below);
                uint32_t i = n;
                if (r < avail) i = cb + r;
                else if (nb + (r - avail) < ne) i = nb + (r - avail);
                if (i < n) active = begin(i);
            }
            if (avail >= want) {
                cb += want;
            } else if (ne > nb) {
                cb = nb + (want - avail) < ne ? nb + (want - avail) : ne;
                ce = ne;
            } else {
                cb = ce;
            }
        }
        if (!active) {
            if (exhausted && cb >= ce && __ballot(active) == 0) break;
            continue;
        }
        if (step()) {
            finish();
            active = false;
        }
    }
}

}  // namespace s3dg
