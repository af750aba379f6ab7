// oracle.cpp -- CPU restatement of the reference's rendering hot path.
//
// TEST INFRASTRUCTURE ONLY: the HIP path is compared against this file, and it
// is the timed CPU baseline.  It is built with g++ -ffp-contract=off so that every
// float operation is one IEEE single-precision operation in source order, as
// in the reference's MSVC /arch:AVX build (which has no FMA).  Each function
// cites the reference file:line whose behaviour it restates (paths relative to
// the reference root).  Written from the reference's observable semantics; the
// data structures are index-based rather than the reference's virtual IShape
// objects, which changes nothing that a caller can observe.
//
// Pins (tests/test_oracle_*.py): published PCG32 and Philox4x32-10 vectors;
// closed forms and float64 recomputation for the triangle and sphere tests,
// cameras, textures, materials, tonemap operators (hand-evaluated values with
// the reference's constants) and NLM; hand-worked SAH trees (bucket choice,
// the untruncated partition predicate, first-minimum ties, std::partition's
// swap order, the BVH8 -> BVH4 fall-back, an eight-child node) plus tree
// invariants; brute force for the traversal; rigid and scaled instances
// against pre-transformed meshes.  Parity unpinned: outputs that depend on
// the C runtime's cosf/sinf/powf/expf/logf (MSVC for the reference, glibc
// here), which are compared with tolerances.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Constants: src/include/s3d_math.h:44-54
// ---------------------------------------------------------------------------
const float kHitMax = 1e12f;
const float kHitMin = 1e-3f;
const float kPi = 3.1415926535897932384626433832795f;
const float k2Pi = 6.283185307179586476925286766559f;
const float kFMax = 3.402823466e+38F;

// ---------------------------------------------------------------------------
// Vector3: src/include/s3d_math.h:505-970 (component-wise ops, Dot/Cross order)
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
static inline V3 operator*(V3 a, float b) { return V3(a.x * b, a.y * b, a.z * b); }
static inline V3 operator/(V3 a, float b) { return V3(a.x / b, a.y / b, a.z / b); }
static inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 operator/(V3 a, V3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline V3 operator+(V3 a, float b) { return V3(a.x + b, a.y + b, a.z + b); }
static inline V3 operator-(V3 a, float b) { return V3(a.x - b, a.y - b, a.z - b); }
static inline V3 operator+(float f, V3 v) { return V3(f + v.x, f + v.y, f + v.z); }
static inline V3 operator-(float f, V3 v) { return V3(f - v.x, f - v.y, f - v.z); }
static inline V3 operator*(float f, V3 v) { return v * f; }
static inline float dot(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline V3 cross(V3 a, V3 b) {
    return V3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
// UnitVector: s3d_math.h:829-838 (divide, not multiply by reciprocal)
static inline V3 unit(V3 v) {
    float mag = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return V3(v.x / mag, v.y / mag, v.z / mag);
}
// SafeUnitVector: s3d_math.h:843-855
static inline V3 safe_unit(V3 v) {
    float mag = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (mag > 0.0f) return V3(v.x / mag, v.y / mag, v.z / mag);
    return v;
}
// Reflect: s3d_math.h:892-902
static inline V3 reflect(V3 i, V3 n) {
    float d = (n.x * i.x) + (n.y * i.y) + (n.z * i.z);
    float d2 = 2.0f * d;
    return V3(i.x - (d2 * n.x), i.y - (d2 * n.y), i.z - (d2 * n.z));
}
static inline V3 vmin(V3 a, V3 b) {
    return V3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z);
}
static inline V3 vmax(V3 a, V3 b) {
    return V3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z);
}
static inline V3 load3(const float *p) { return V3(p[0], p[1], p[2]); }
static inline void store3(float *p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
// s3d::Max / s3d::Min: s3d_math.h:66-75 (NaN picks the second operand)
static inline float smax(float a, float b) { return (a > b) ? a : b; }
static inline float smin(float a, float b) { return (a < b) ? a : b; }
static inline float saturate(float v) { return smax(0.0f, smin(1.0f, v)); }
static inline float safe_sqrt(float v) { return v > FLT_EPSILON ? sqrtf(v) : 0.0f; }

// ---------------------------------------------------------------------------
// BoundingBox: s3d_math.h:2143-2264 and SurfaceArea(box) 2904-2910
// ---------------------------------------------------------------------------
struct Box {
    V3 mini, maxi, center;
    bool empty;
    Box() : mini(kFMax, kFMax, kFMax), maxi(-kFMax, -kFMax, -kFMax), center(0, 0, 0), empty(true) {}
    explicit Box(V3 p) : mini(p), maxi(p), center(p), empty(false) {}
    Box(V3 lo, V3 hi) : mini(lo), maxi(hi), center((lo + hi) * 0.5f), empty(false) {}
};
static inline Box merge(const Box &a, const Box &b) {
    if (!a.empty && !b.empty) return Box(vmin(a.mini, b.mini), vmax(a.maxi, b.maxi));
    if (a.empty && !b.empty) return b;
    return a;
}
static inline Box merge(const Box &b, V3 p) {
    if (!b.empty) return Box(vmin(b.mini, p), vmax(b.maxi, p));
    return Box(p);
}
static inline float surface_area(const Box &b) {
    float X = fabsf(b.maxi.x - b.mini.x);
    float Y = fabsf(b.maxi.y - b.mini.y);
    float Z = fabsf(b.maxi.z - b.mini.z);
    return 2.0f * ((X * Y) + (X * Z) + (Y * Z));
}

// ---------------------------------------------------------------------------
// RNG
// ---------------------------------------------------------------------------
// PCG: s3d_math.h:2769-2817 (XSH-RR output, fixed increment, SetSeed = seed+inc
// then one step).  GetAsF32 = float(u32) / float(0xffffffff) can return 1.0f.
struct PCG {
    uint64_t state;
    static const uint64_t kMul = 6364136223846793005ull;
    static const uint64_t kInc = 1442695040888963407ull;
    void seed(uint64_t s) { state = s + kInc; u32(); }
    uint32_t u32() {
        uint64_t old = state;
        state = old * kMul + kInc;
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((~rot + 1) & 31));
    }
    float f32() { return (float)u32() / (float)0xffffffffu; }
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Not in the reference:
// the counter-based key for each path's PCG stream.
static void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
// Seed of the PCG stream of one path: Philox(counter = {pixel, pass, scene, 0},
// key = seed).  Replaces the shared m_Random of src/src/s3d_pt.cpp:191,242,304.
static uint64_t path_seed(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t scene) {
    uint32_t ctr[4] = {pixel, pass, scene, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox(ctr, key, o);
    return (uint64_t)o[0] | ((uint64_t)o[1] << 32);
}

// ---------------------------------------------------------------------------
// Rays and hit records: s3d_math.h:1437-1469, s3d_shape.h:38-54
// ---------------------------------------------------------------------------
struct Ray {
    V3 pos, dir;
    float tmin, tmax;
};
struct Hit {
    float distance;
    float b1, b2;
    int shape;      // top-level shape index
    int prim;       // depth-first position of the triangle inside its mesh, -2 = sphere
    Hit() : distance(kFMax), b1(0), b2(0), shape(-1), prim(-1) {}
};

// ---------------------------------------------------------------------------
// Triangle: src/src/s3d_triangle.cpp:22-40 (ctor), 73-105 (IsHit), 107-121 (CalcParam)
// ---------------------------------------------------------------------------
struct Tri {
    V3 p[3], n[3];
    float uv[3][2];
    V3 e0, e1;   // m_Edge[0] = v1 - v0, m_Edge[1] = v2 - v0
    Box box;
    int material;
    int input_index;
};
static Tri make_tri(const s3dg_vertex *v, int material, int index) {
    Tri t;
    for (int i = 0; i < 3; ++i) {
        t.p[i] = load3(v[i].position);
        t.n[i] = load3(v[i].normal);
        t.uv[i][0] = v[i].texcoord[0];
        t.uv[i][1] = v[i].texcoord[1];
    }
    V3 lo = t.p[0], hi = t.p[0];
    lo = vmin(lo, t.p[1]); lo = vmin(lo, t.p[2]);
    hi = vmax(hi, t.p[1]); hi = vmax(hi, t.p[2]);
    t.box = Box(lo, hi);
    t.e0 = t.p[1] - t.p[0];
    t.e1 = t.p[2] - t.p[0];
    t.material = material;
    t.input_index = index;
    return t;
}
// Triangle::IsHit with the ray's [tmin, tmax] in place of [F_HIT_MIN, F_HIT_MAX].
// Returns true and writes rec when the hit is strictly closer than rec.distance.
static inline bool tri_hit(const Tri &tr, const Ray &ray, float &dist_out, float &b1, float &b2,
                           float cur) {
    V3 s1 = cross(ray.dir, tr.e1);
    float div = dot(s1, tr.e0);
    if (fabsf(div) <= FLT_EPSILON) return false;
    V3 d = ray.pos - tr.p[0];
    float beta = dot(d, s1) / div;
    if (beta <= 0.0f || beta >= 1.0f) return false;
    V3 s2 = cross(d, tr.e0);
    float gamma = dot(ray.dir, s2) / div;
    if (gamma <= 0.0f || (beta + gamma) >= 1.0f) return false;
    float dist = dot(tr.e1, s2) / div;
    if (dist < ray.tmin || dist > ray.tmax) return false;
    if (dist >= cur) return false;
    dist_out = dist; b1 = beta; b2 = gamma;
    return true;
}
// Triangle::CalcParam normal (s3d_triangle.cpp:107-121) with SafeNormalize (s3d_math.h:815-825)
static V3 tri_normal(const Tri &t, float beta, float gamma) {
    float alpha = 1.0f - beta - gamma;
    V3 n((t.n[0].x * alpha) + (t.n[1].x * beta) + (t.n[2].x * gamma),
         (t.n[0].y * alpha) + (t.n[1].y * beta) + (t.n[2].y * gamma),
         (t.n[0].z * alpha) + (t.n[1].z * beta) + (t.n[2].z * gamma));
    float mag = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    if (mag > 0.0f) { n.x /= mag; n.y /= mag; n.z /= mag; }
    return n;
}

// ---------------------------------------------------------------------------
// Binned-SAH builders.
//   BVH8: src/src/s3d_bvh8.cpp:28-110 (helpers, CreateNode), 254-342 (SplitSAH),
//         401-479 (Create).  BVH4: src/src/s3d_bvh4.cpp:94-104, 237-325, 384-423.
//   BVH2: src/src/s3d_bvh2.cpp:174-274.  Leaf: src/src/s3d_leaf.cpp:22-35.
// The prim array is an index array partitioned in place exactly where the
// reference partitions its IShape** array, so the final array is the
// depth-first leaf order.
// ---------------------------------------------------------------------------
struct Node {
    int kind;                 // 8, 4, 2 or 0 (leaf)
    std::vector<int> child;
    int begin, count;         // range of the prim array covered by this node
    Box box;                  // GetBox()
};

struct Builder {
    const std::vector<Tri> &tris;
    std::vector<int> &prims;
    std::vector<Node> &nodes;

    Box merged_box(size_t count, const int *pp) const {
        if (count == 0) return Box();
        Box b = tris[pp[0]].box;
        for (size_t i = 1; i < count; ++i) b = merge(b, tris[pp[i]].box);
        return b;
    }
    Box centroid_box(size_t count, const int *pp) const {
        if (count == 0) return Box();
        Box b(tris[pp[0]].box.center);
        for (size_t i = 1; i < count; ++i) b = merge(b, tris[pp[i]].box.center);
        return b;
    }
    static int longest_axis(const Box &b) {
        V3 v = b.maxi - b.mini;
        if (v.x > v.y && v.x > v.z) return 0;
        if (v.y > v.z) return 1;
        return 2;
    }
    static float offset_axis(const Box &b, V3 p, int axis) {
        V3 o = p - b.mini;
        if (b.maxi.x > b.mini.x) o.x /= b.maxi.x - b.mini.x;
        if (b.maxi.y > b.mini.y) o.y /= b.maxi.y - b.mini.y;
        if (b.maxi.z > b.mini.z) o.z /= b.maxi.z - b.mini.z;
        return o[axis];
    }
    // Two-ended partition: same element order as both libstdc++ and MSVC
    // std::partition on bidirectional iterators.
    template <class P> static size_t partition(int *first, int *last, P pred) {
        int *base = first;
        for (;;) {
            for (;;) {
                if (first == last) return (size_t)(first - base);
                if (pred(*first)) ++first; else break;
            }
            --last;
            for (;;) {
                if (first == last) return (size_t)(first - base);
                if (!pred(*last)) --last; else break;
            }
            std::swap(*first, *last);
            ++first;
        }
    }
    // SplitSAH shared by BVH8/BVH4 (min count 8 / 4, 16 buckets) and the body
    // of BVH2::Create (12 buckets).  Returns false where the reference does.
    bool split_sah(size_t count, int *pp, size_t &mid, size_t min_count, int buckets) {
        if (count <= min_count) return false;
        Box bound = merged_box(count, pp);
        Box centroid = centroid_box(count, pp);
        int axis = longest_axis(centroid);
        if (centroid.maxi[axis] == centroid.mini[axis]) return false;
        int bcount[16];
        Box bbox[16];
        for (int i = 0; i < buckets; ++i) { bcount[i] = 0; bbox[i] = Box(); }
        for (size_t i = 0; i < count; ++i) {
            int idx = (int)((float)buckets * offset_axis(centroid, tris[pp[i]].box.center, axis));
            if (idx == buckets) idx = buckets - 1;
            bcount[idx]++;
            bbox[idx] = merge(bbox[idx], tris[pp[i]].box);
        }
        float cost[15] = {};
        for (int i = 0; i < buckets - 1; ++i) {
            Box b0, b1;
            int c0 = 0, c1 = 0;
            for (int j = 0; j <= i; ++j) { b0 = merge(b0, bbox[j]); c0 += bcount[j]; }
            for (int j = i + 1; j < buckets; ++j) { b1 = merge(b1, bbox[j]); c1 += bcount[j]; }
            cost[i] = 1 + ((float)c0 * surface_area(b0) + (float)c1 * surface_area(b1)) / surface_area(bound);
        }
        float min_cost = cost[0];
        int split = 0;
        for (int i = 1; i < buckets - 1; ++i)
            if (cost[i] < min_cost) { min_cost = cost[i]; split = i; }
        float leaf_cost = (float)count;
        if (!(min_cost < leaf_cost)) return false;
        // The predicate compares the untruncated float 16*offset with the
        // bucket index (s3d_bvh8.cpp:321-334).
        const float fb = (float)buckets;
        mid = partition(pp, pp + count, [&](int p) {
            float idx = fb * offset_axis(centroid, tris[p].box.center, axis);
            if (idx == fb) idx = fb - 1;
            return idx <= (float)split;
        });
        return true;
    }
    int new_node(int kind, size_t begin, size_t count) {
        Node n;
        n.kind = kind;
        n.begin = (int)begin;
        n.count = (int)count;
        n.box = merged_box(count, prims.data() + begin);
        nodes.push_back(n);
        return (int)nodes.size() - 1;
    }
    int leaf(size_t begin, size_t count) { return new_node(0, begin, count); }
    // BVH4 CreateNode: always a Leaf (s3d_bvh4.cpp:94-104)
    int bvh4(size_t begin, size_t count) {
        int *pp = prims.data() + begin;
        if (count <= 4) return leaf(begin, count);
        size_t mid = 0;
        if (!split_sah(count, pp, mid, 4, 16)) return leaf(begin, count);
        size_t midL = 0, midR = 0;
        if (!split_sah(mid, pp, midL, 4, 16)) return leaf(begin, count);
        if (!split_sah(count - mid, pp + mid, midR, 4, 16)) return leaf(begin, count);
        size_t idx[5] = {0, midL, mid, mid + midR, count};
        int c[4];
        for (int i = 0; i < 4; ++i) c[i] = bvh4(begin + idx[i], idx[i + 1] - idx[i]);
        int n = new_node(4, begin, count);
        nodes[n].child.assign(c, c + 4);
        return n;
    }
    // BVH8 CreateNode: Leaf if count <= 8 else BVH4::Create (s3d_bvh8.cpp:94-110)
    int create_node8(size_t begin, size_t count) {
        if (count <= 8) return leaf(begin, count);
        return bvh4(begin, count);
    }
    int bvh8(size_t begin, size_t count) {
        int *pp = prims.data() + begin;
        if (count <= 8) return create_node8(begin, count);
        size_t mid = 0;
        if (!split_sah(count, pp, mid, 8, 16)) return create_node8(begin, count);
        size_t midL = 0, midR = 0;
        if (!split_sah(mid, pp, midL, 8, 16)) return create_node8(begin, count);
        if (!split_sah(count - mid, pp + mid, midR, 8, 16)) return create_node8(begin, count);
        size_t a[4] = {0, midL, mid, mid + midR};
        size_t cn[4] = {midL, mid - midL, midR, count - mid - midR};
        size_t m[4];
        for (int i = 0; i < 4; ++i)
            if (!split_sah(cn[i], pp + a[i], m[i], 8, 16)) return create_node8(begin, count);
        size_t idx[9] = {0, m[0], midL, midL + m[1], mid, mid + m[2], mid + midR, mid + midR + m[3], count};
        int c[8];
        for (int i = 0; i < 8; ++i) c[i] = bvh8(begin + idx[i], idx[i + 1] - idx[i]);
        int n = new_node(8, begin, count);
        nodes[n].child.assign(c, c + 8);
        return n;
    }
    int bvh2(size_t begin, size_t count) {
        int *pp = prims.data() + begin;
        if (count <= 4) return leaf(begin, count);
        size_t mid = 0;
        // BVH2::Create has no "count <= min" early out inside the split, its
        // own count <= 4 test comes first; same 12-bucket SAH otherwise.
        if (!split_sah(count, pp, mid, 0, 12)) return leaf(begin, count);
        int c0 = bvh2(begin, mid);
        int c1 = bvh2(begin + mid, count - mid);
        int n = new_node(2, begin, count);
        nodes[n].child.assign({c0, c1});
        return n;
    }
};

// Slab test against the infinite line, SSE/AVX lane semantics:
// BoundingBox4/8::IsHit (s3d_math.h:2364-2401, 2549-2586).  True division,
// _mm_min/max_ps operand order, final tmax >= tmin (false on NaN).
static inline bool slab_simd(const Box &b, const Ray &r) {
    float tmin = -kHitMax, tmax = kHitMax;
    for (int a = 0; a < 3; ++a) {
        float t0 = (b.mini[a] - r.pos[a]) / r.dir[a];
        float t1 = (b.maxi[a] - r.pos[a]) / r.dir[a];
        float n = (t0 < t1) ? t0 : t1;
        float f = (t0 > t1) ? t0 : t1;
        tmin = (tmin > n) ? tmin : n;
        tmax = (tmax < f) ? tmax : f;
    }
    return tmax >= tmin;
}
// BoundingBox::IsHit (s3d_math.h:2201-2226), used by BVH2 (s3d_bvh2.cpp:143-153)
static inline bool slab_scalar(const Box &b, const Ray &r) {
    if (b.empty) return false;
    float tmin = -kHitMax, tmax = kHitMax;
    for (int a = 0; a < 3; ++a) {
        float t0 = (b.mini[a] - r.pos[a]) / r.dir[a];
        float t1 = (b.maxi[a] - r.pos[a]) / r.dir[a];
        float n = smin(t0, t1), f = smax(t0, t1);
        tmin = smax(tmin, n);
        tmax = smin(tmax, f);
        if (tmin > tmax) return false;
    }
    return true;
}

struct Mesh {
    std::vector<Tri> tris;      // input order
    std::vector<int> prims;     // depth-first order -> input index
    std::vector<Node> nodes;
    int root = -1;
    int kind = 8;
    Box box;

    // Reference traversal: BVH8/BVH4/BVH2/Leaf::IsHit recursion
    // (s3d_bvh8.cpp:191-206, s3d_bvh4.cpp:174-189, s3d_bvh2.cpp:143-153, s3d_leaf.cpp:74-87).
    bool visit(int ni, const Ray &r, Hit &h) const {
        const Node &n = nodes[ni];
        if (n.kind == 0) {
            bool hit = false;
            for (int i = 0; i < n.count; ++i) {
                int dfs = n.begin + i;
                const Tri &t = tris[prims[dfs]];
                float d, b1, b2;
                if (tri_hit(t, r, d, b1, b2, h.distance)) {
                    h.distance = d; h.b1 = b1; h.b2 = b2; h.prim = dfs;
                    hit = true;
                }
            }
            return hit;
        }
        if (n.kind == 2) {
            if (!slab_scalar(n.box, r)) return false;
            bool hit = false;
            for (int c : n.child) hit |= visit(c, r, h);
            return hit;
        }
        bool mask[8];
        bool any = false;
        for (size_t i = 0; i < n.child.size(); ++i) {
            mask[i] = slab_simd(nodes[n.child[i]].box, r);
            any |= mask[i];
        }
        if (!any) return false;
        bool hit = false;
        for (size_t i = 0; i < n.child.size(); ++i)
            if (mask[i]) hit |= visit(n.child[i], r, h);
        return hit;
    }
    bool intersect(const Ray &r, Hit &h, bool brute) const {
        if (root < 0) return false;
        if (!brute) return visit(root, r, h);
        bool hit = false;
        for (size_t dfs = 0; dfs < prims.size(); ++dfs) {
            float d, b1, b2;
            if (tri_hit(tris[prims[dfs]], r, d, b1, b2, h.distance)) {
                h.distance = d; h.b1 = b1; h.b2 = b2; h.prim = (int)dfs;
                hit = true;
            }
        }
        return hit;
    }
};

static Mesh *build_mesh(uint32_t vertex_count, const s3dg_vertex *v, const int32_t *tri_mat,
                        int material, int kind) {
    Mesh *m = new Mesh();
    uint32_t n = vertex_count / 3;
    m->tris.reserve(n);
    for (uint32_t i = 0; i < n; ++i)
        m->tris.push_back(make_tri(v + 3 * i, tri_mat ? tri_mat[i] : material, (int)i));
    m->prims.resize(n);
    for (uint32_t i = 0; i < n; ++i) m->prims[i] = (int)i;
    m->kind = kind;
    if (n > 0) {
        Builder b{m->tris, m->prims, m->nodes};
        if (kind == 2) m->root = b.bvh2(0, n);
        else if (kind == 4) m->root = b.bvh4(0, n);
        else m->root = b.bvh8(0, n);
        m->box = m->nodes[m->root].box;
    }
    return m;
}

// ---------------------------------------------------------------------------
// Matrix (row-vector convention): Invert s3d_math.h:1776-1842 (Det, and the
// operator-precedence quirk that divides only the last product by det),
// TransformCoord 2112-2125, TransformNormal 2098-2106.
// ---------------------------------------------------------------------------
struct Mat { float m[4][4]; };
static float det(const Mat &M) {
    const float(&a)[4][4] = M.m;
#define E(r, c) a[r - 1][c - 1]
    return ((E(1,1) * E(2,2) * E(3,3) * E(4,4)) + (E(1,1) * E(2,3) * E(3,4) * E(4,2)) +
            (E(1,1) * E(2,4) * E(3,2) * E(4,3)) + (E(1,2) * E(2,1) * E(3,4) * E(4,3)) +
            (E(1,2) * E(2,3) * E(3,1) * E(4,4)) + (E(1,2) * E(2,4) * E(3,3) * E(4,1)) +
            (E(1,3) * E(2,1) * E(3,2) * E(4,4)) + (E(1,3) * E(2,2) * E(3,4) * E(4,1)) +
            (E(1,3) * E(2,4) * E(3,1) * E(4,2)) + (E(1,4) * E(2,1) * E(3,3) * E(4,2)) +
            (E(1,4) * E(2,2) * E(3,1) * E(4,3)) + (E(1,4) * E(2,3) * E(3,2) * E(4,1)) -
            (E(1,1) * E(2,2) * E(3,4) * E(4,3)) - (E(1,1) * E(2,3) * E(3,2) * E(4,4)) -
            (E(1,1) * E(2,4) * E(3,3) * E(4,2)) - (E(1,2) * E(2,1) * E(3,3) * E(4,4)) -
            (E(1,2) * E(2,3) * E(3,4) * E(4,1)) - (E(1,2) * E(2,4) * E(3,1) * E(4,3)) -
            (E(1,3) * E(2,1) * E(3,4) * E(4,2)) - (E(1,3) * E(2,2) * E(3,1) * E(4,4)) -
            (E(1,3) * E(2,4) * E(3,2) * E(4,1)) - (E(1,4) * E(2,1) * E(3,2) * E(4,3)) -
            (E(1,4) * E(2,2) * E(3,3) * E(4,1)) - (E(1,4) * E(2,3) * E(3,1) * E(4,2)));
}
static Mat invert(const Mat &V) {
    const float(&a)[4][4] = V.m;
    float d = det(V);
    Mat R;
    // Each entry: p1 + p2 + p3 - p4 - p5 - p6 / det, left to right.
#define T3(r1,c1,r2,c2,r3,c3) (E(r1,c1) * E(r2,c2) * E(r3,c3))
#define ENTRY(o, A, B, C, D, F, G) R.m[o / 10 - 1][o % 10 - 1] = A + B + C - D - F - G / d
    ENTRY(11, T3(2,2,3,3,4,4), T3(2,3,3,4,4,2), T3(2,4,3,2,4,3), T3(2,2,3,4,4,3), T3(2,3,3,2,4,4), T3(2,4,3,3,4,2));
    ENTRY(12, T3(1,2,3,4,4,3), T3(1,3,3,2,4,4), T3(1,4,3,3,4,2), T3(1,2,3,3,4,4), T3(1,3,3,4,4,2), T3(1,4,3,2,4,3));
    ENTRY(13, T3(1,2,2,3,4,4), T3(1,3,2,4,4,2), T3(1,4,2,2,4,3), T3(1,2,2,4,4,3), T3(1,3,2,2,4,4), T3(1,4,2,3,4,2));
    ENTRY(14, T3(1,2,2,4,3,3), T3(1,3,2,2,3,4), T3(1,4,2,3,3,2), T3(1,2,2,3,3,4), T3(1,3,2,4,3,2), T3(1,4,2,2,3,3));
    ENTRY(21, T3(2,1,3,4,4,3), T3(2,3,3,1,4,4), T3(2,4,3,3,4,1), T3(2,1,3,3,4,4), T3(2,3,3,4,4,1), T3(2,4,3,1,4,3));
    ENTRY(22, T3(1,1,3,3,4,4), T3(1,3,3,4,4,1), T3(1,4,3,1,4,3), T3(1,1,3,4,4,3), T3(1,3,3,1,4,4), T3(1,4,3,3,4,1));
    ENTRY(23, T3(1,1,2,4,4,3), T3(1,3,2,1,4,4), T3(1,4,2,3,4,1), T3(1,1,2,3,4,4), T3(1,3,2,4,4,1), T3(1,4,2,1,4,3));
    ENTRY(24, T3(1,1,2,3,3,4), T3(1,3,2,4,3,1), T3(1,4,2,1,3,3), T3(1,1,2,4,3,3), T3(1,3,2,1,3,4), T3(1,4,2,3,3,1));
    ENTRY(31, T3(2,1,3,2,4,4), T3(2,2,3,4,4,1), T3(2,4,3,1,4,2), T3(2,1,3,4,4,2), T3(2,2,3,1,4,4), T3(2,4,3,2,4,1));
    ENTRY(32, T3(1,1,3,4,4,2), T3(1,2,3,1,4,4), T3(1,4,3,2,4,1), T3(1,1,3,2,4,4), T3(1,2,3,4,4,1), T3(1,4,3,1,4,2));
    ENTRY(33, T3(1,1,2,2,4,4), T3(1,2,2,4,4,1), T3(1,4,2,1,4,2), T3(1,1,2,4,4,2), T3(1,2,2,1,4,4), T3(1,4,2,2,4,1));
    ENTRY(34, T3(1,1,2,4,3,2), T3(1,2,2,1,3,4), T3(1,4,2,2,3,1), T3(1,1,2,2,3,4), T3(1,2,2,4,3,1), T3(1,4,2,1,3,2));
    ENTRY(41, T3(2,1,3,3,4,2), T3(2,2,3,1,4,3), T3(2,3,3,2,4,1), T3(2,1,3,2,4,3), T3(2,2,3,3,4,1), T3(2,3,3,1,4,2));
    ENTRY(42, T3(1,1,3,2,4,3), T3(1,2,3,3,4,1), T3(1,3,3,1,4,2), T3(1,1,3,3,4,2), T3(1,2,3,1,4,3), T3(1,3,3,2,4,1));
    ENTRY(43, T3(1,1,2,3,4,2), T3(1,2,2,1,4,3), T3(1,3,2,2,4,1), T3(1,1,2,2,4,3), T3(1,2,2,3,4,1), T3(1,3,2,1,4,2));
    ENTRY(44, T3(1,1,2,2,3,3), T3(1,2,2,3,3,1), T3(1,3,2,1,3,2), T3(1,1,2,3,3,2), T3(1,2,2,1,3,3), T3(1,3,2,2,3,1));
#undef ENTRY
#undef T3
#undef E
    return R;
}
static V3 transform_coord(V3 c, const Mat &M) {
    const float(&m)[4][4] = M.m;
    float X = (c.x * m[0][0]) + (c.y * m[1][0]) + (c.z * m[2][0]) + m[3][0];
    float Y = (c.x * m[0][1]) + (c.y * m[1][1]) + (c.z * m[2][1]) + m[3][1];
    float Z = (c.x * m[0][2]) + (c.y * m[1][2]) + (c.z * m[2][2]) + m[3][2];
    float W = (c.x * m[0][3]) + (c.y * m[1][3]) + (c.z * m[2][3]) + m[3][3];
    return V3(X / W, Y / W, Z / W);
}
static V3 transform_normal(V3 n, const Mat &M) {
    const float(&m)[4][4] = M.m;
    return V3((n.x * m[0][0]) + (n.y * m[1][0]) + (n.z * m[2][0]),
              (n.x * m[0][1]) + (n.y * m[1][1]) + (n.z * m[2][1]),
              (n.x * m[0][2]) + (n.y * m[1][2]) + (n.z * m[2][2]));
}

// ---------------------------------------------------------------------------
// Materials: Lambert (src/src/s3d_lambert.cpp:23-30 threshold, 64-86 Shade),
// Mirror (src/src/s3d_mirror.cpp:59-71), Glass (src/src/s3d_glass.cpp:60-128),
// Phong (src/src/s3d_phong.cpp:23-31, 65-88), Plastic (src/src/s3d_plastic.cpp:23-37, 71-142).
// Random draws that the reference passes as two arguments of one call
// (s3d_lambert.cpp:75 and similar) are taken left to right.
// ---------------------------------------------------------------------------
struct Material {
    int type;
    V3 diffuse, specular, emissive;
    float ior, power;
    float thr0, thr1;
    int texture;   // TexturedMaterial wrapper (src/src/s3d_texturedmaterial.cpp): -1 none
};
static Material make_material(const s3dg_material &m) {
    Material r;
    r.type = m.type;
    r.texture = m.texture - 1;
    r.diffuse = load3(m.diffuse);
    r.specular = load3(m.specular);
    r.emissive = load3(m.emissive);
    r.ior = m.ior;
    r.power = m.power;
    V3 base = (m.type == S3DG_MAT_LAMBERT || m.type == S3DG_MAT_PLASTIC) ? r.diffuse : r.specular;
    float t = base.x;
    t = smax(t, base.y);
    t = smax(t, base.z);
    t = smax(t, 0.01f);
    r.thr0 = t;
    // Plastic: m_Threshold[1] is built from m_Threshold[0] (s3d_plastic.cpp:33-37)
    float t1 = r.specular.x;
    t1 = smax(t, r.specular.y);
    t1 = smax(t, r.specular.z);
    t1 = smax(t, 0.01f);
    r.thr1 = t1;
    return r;
}
static bool has_delta(const Material &m) {
    return m.type == S3DG_MAT_MIRROR || m.type == S3DG_MAT_GLASS;
}
static V3 base_color(const Material &m) {
    return (m.type == S3DG_MAT_LAMBERT || m.type == S3DG_MAT_PLASTIC) ? m.diffuse : m.specular;
}
// TangentSpace: s3d_math.h:2875-2886 (Duff et al. 2017, Listing 3)
static void tangent_space(V3 N, V3 &T, V3 &B) {
    float s = (N.z >= 0.0f) ? 1.0f : -1.0f;
    float a = -1.0f / (s + N.z);
    float b = N.x * N.y * a;
    T = V3(1.0f + s * N.x * N.x * a, s * b, -s * N.x);
    B = V3(b, s + N.y * N.y * a, -N.y);
}
// SampleLambert / SamplePhong: s3d_math.h:2968-2982
static V3 sample_lambert(float u, float v) {
    float r = sqrtf(1.0f - u * u);
    float p = k2Pi * v;
    return V3(r * cosf(p), r * sinf(p), u);
}
static V3 sample_phong(float u, float v, float shininess) {
    float p = k2Pi * u;
    float ct = powf(1.0f - v, 1.0f / (shininess + 1.0f));
    float st = sqrtf(1.0f - ct * ct);
    return V3(st * cosf(p), st * sinf(p), ct);
}
static V3 shade(const Material &m, V3 input, V3 normal, PCG &rng, V3 &output, bool &dice) {
    switch (m.type) {
    case S3DG_MAT_LAMBERT: {
        V3 N = (dot(normal, input) < 0.0f) ? normal : -normal;
        V3 T, B;
        tangent_space(N, T, B);
        float u = rng.f32();
        float v = rng.f32();
        V3 s = sample_lambert(u, v);
        output = safe_unit(T * s.x + B * s.y + N * s.z);
        dice = (rng.f32() >= m.thr0);
        return m.diffuse;
    }
    case S3DG_MAT_MIRROR: {
        V3 nm = (dot(normal, input) < 0.0f) ? normal : -normal;
        output = safe_unit(reflect(input, nm));
        dice = false;
        return m.specular;
    }
    case S3DG_MAT_GLASS: {
        V3 nm = (dot(normal, input) < 0.0f) ? normal : -normal;
        V3 refl = safe_unit(reflect(input, nm));
        bool into = dot(normal, nm) > 0.0f;
        dice = false;
        const float nc = 1.0f, nt = m.ior;
        const float nnt = into ? (nc / nt) : (nt / nc);
        const float ddn = dot(input, nm);
        const float cos2t = 1.0f - nnt * nnt * (1.0f - ddn * ddn);
        if (cos2t < 0.0f) { output = refl; return m.specular; }
        V3 refr = safe_unit(input * nnt - normal * (into ? 1.0f : -1.0f) * (ddn * nnt + safe_sqrt(cos2t)));
        const float a = nt - nc, b = nt + nc;
        const float R0 = (a * a) / (b * b);
        const float c = 1.0f - (into ? -ddn : dot(refr, normal));
        float c2 = c * c;
        const float Re = R0 + (1.0f - R0) * (c2 * c2 * c);
        const float Tr = 1.0f - Re;
        const float P = 0.25f + 0.5f * Re;
        float prob = 0.5f;
        if (rng.f32() < P) { output = refl; return m.specular * Re / (P * prob); }
        output = refr;
        return m.specular * Tr / ((1.0f - P) * prob);
    }
    case S3DG_MAT_PHONG: {
        float u = rng.f32();
        float v = rng.f32();
        V3 s = sample_phong(u, v, m.power);
        V3 n = (dot(input, normal) < 0.0f) ? normal : -normal;
        V3 w = reflect(input, n);
        {
            float mag = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
            if (mag > 0.0f) { w.x /= mag; w.y /= mag; w.z /= mag; }
        }
        V3 T, B;
        tangent_space(w, T, B);
        V3 dir = safe_unit(T * s.x + B * s.y + w * s.z);
        float cosine = fabsf(dot(dir, n));
        output = dir;
        dice = (rng.f32() >= m.thr0);
        return m.specular * cosine;
    }
    case S3DG_MAT_PLASTIC:
    default: {
        float cosine = dot(normal, input);
        V3 nm = (cosine < 0.0f) ? normal : -normal;
        if (cosine < 0.0f) cosine = -cosine;
        float temp1 = 1.0f - cosine;
        const float R0 = 0.5f;
        float R = R0 + (1.0f - R0) * temp1 * temp1 * temp1 * temp1 * temp1;
        float P = (R + 0.5f) / 2.0f;
        if (rng.f32() <= P) {
            V3 T, B;
            tangent_space(nm, T, B);
            float u = rng.f32();
            float v = rng.f32();
            V3 s = sample_lambert(u, v);
            output = unit(T * s.x + B * s.y + nm * s.z);
            V3 result = m.diffuse * R / P;
            dice = (rng.f32() >= m.thr0);
            return result;
        }
        float u = rng.f32();
        float v = rng.f32();
        V3 s = sample_phong(u, v, m.power);
        V3 w = reflect(input, nm);
        w = unit(w);
        V3 T, B;
        tangent_space(w, T, B);
        V3 dir = unit(T * s.x + B * s.y + w * s.z);
        float dots = dot(dir, nm);
        output = dir;
        dice = (rng.f32() >= m.thr1);
        return m.specular * dots * (1.0f - R) / (1.0f - P);
    }
    }
}

// ---------------------------------------------------------------------------
// Scene: Scene::Intersect brute force over top-level shapes in list order
// (src/include/s3d_scene.h:63-75); Instance::IsHit (s3d_instance.cpp:65-72);
// Sphere::IsHit / CalcParam / Sample (s3d_sphere.cpp:66-115); cameras
// (s3d_camera.h:63-100, 171-225).
// ---------------------------------------------------------------------------
struct Shape {
    int type;
    int mesh;
    int material;
    V3 center;
    float radius;
    Mat inv;
};
struct Light {
    int type, shape, first, count;
    V3 corner, e1, e2;
    float pdf;
};
// Texture (src/include/s3d_texture.h:106-111 GetColor, src/src/s3d_texture.cpp:925-972).
struct Texture {
    int w = 0, h = 0;
    std::vector<V3> px;
    V3 get(int x, int y) const { return px[(size_t)w * y + x]; }
    // SampleColor(const Vector2&): bilinear over the corners of uv * (size - 1),
    // wrapped by abs(i % size).  The float -> s32 conversion of NaN gives 0 and
    // the range is held to +-2^30 (undefined in the reference).
    V3 sample(float u, float v) const {
        float fx = u * (float)(w - 1);
        float fy = v * (float)(h - 1);
        auto to_i = [](float f) -> int {
            f = floorf(f);
            if (f != f) return 0;
            return (int)smin(smax(f, -1073741824.0f), 1073741824.0f);
        };
        int x0 = to_i(fx), y0 = to_i(fy);
        int x1 = x0 + 1, y1 = y0 + 1;
        float x_w0 = (float)x1 - fx, x_w1 = fx - (float)x0;
        float y_w0 = (float)y1 - fy, y_w1 = fy - (float)y0;
        x0 = abs(x0 % w); x1 = abs(x1 % w);
        y0 = abs(y0 % h); y1 = abs(y1 % h);
        return x_w0 * (y_w0 * get(x0, y0) + y_w1 * get(x0, y1)) + x_w1 * (y_w0 * get(x1, y0) + y_w1 * get(x1, y1));
    }
    // SampleColor(const Vector3&): latitude-longitude map; IsZero is |x| <= FLT_EPSILON
    // (s3d_math.h:125-126).
    V3 sample_dir(V3 dir) const {
        float u = 0.0f;
        const float theta = acosf(dir.y);
        float v = theta * 0.31830988618379067153776752674503f;            // F_1DIVPI
        if (!(fabsf(dir.x) <= FLT_EPSILON) && !(fabsf(dir.z) <= FLT_EPSILON)) {
            const float phi = atan2f(dir.z, dir.x);
            u = (dir.z < 0.0f) ? (phi + k2Pi) : phi;
            u *= 0.15915494309189533576888376337251f;                    // F_1DIV2PI
        }
        return sample(u, v);
    }
};

struct Scene {
    std::vector<Mesh *> meshes;
    std::vector<Shape> shapes;
    std::vector<Material> materials;
    std::vector<Light> lights;
    std::vector<Texture> textures;
    int env_tex = -1;   // IBL texture, -1: constant environment
    V3 env;
    int cam_type;
    V3 cam_pos, cam_dir, CX, CY, CZ;
    float lens_radius, focal_distance;
    ~Scene() { for (Mesh *m : meshes) delete m; }
};

static bool sphere_hit(const Shape &s, const Ray &r, Hit &h) {
    const V3 po = s.center - r.pos;
    const float b = dot(po, r.dir);
    const float D4 = b * b - dot(po, po) + s.radius * s.radius;
    if (D4 < 0.0f) return false;
    const float sq = sqrtf(D4);
    const float t1 = b - sq, t2 = b + sq;
    if (t1 < r.tmin && t2 < r.tmin) return false;
    float dist = (t1 > r.tmin) ? t1 : t2;
    if (dist > h.distance) return false;
    h.distance = dist;
    h.b1 = 0.0f; h.b2 = 0.0f;
    h.prim = -2;
    return true;
}

static bool scene_intersect(const Scene &sc, const Ray &ray, Hit &h) {
    bool flag = false;
    for (size_t i = 0; i < sc.shapes.size(); ++i) {
        const Shape &s = sc.shapes[i];
        bool hit = false;
        if (s.type == S3DG_SHAPE_SPHERE) {
            hit = sphere_hit(s, ray, h);
        } else if (s.type == S3DG_SHAPE_MESH) {
            hit = sc.meshes[s.mesh]->intersect(ray, h, false);
        } else {
            Ray lr = ray;
            lr.pos = transform_coord(ray.pos, s.inv);
            lr.dir = unit(transform_normal(ray.dir, s.inv));
            hit = sc.meshes[s.mesh]->intersect(lr, h, false);
        }
        if (hit) h.shape = (int)i;
        flag |= hit;
    }
    return flag;
}

static void camera_ray(const Scene &sc, float x, float y, PCG *rng, V3 &org, V3 &dir) {
    V3 pos = (sc.CX * x) + (sc.CY * y) + sc.CZ;
    dir = unit(pos - sc.cam_pos);
    org = sc.cam_pos;
    if (sc.cam_type == S3DG_CAMERA_THINLENS && sc.lens_radius > 0.0f && rng) {
        // SampleLens (s3d_camera.h:262-267) draws from the path's stream here.
        float theta = k2Pi * rng->f32();
        float rr = sc.lens_radius * safe_sqrt(rng->f32());
        V3 diff(rr * cosf(theta), rr * sinf(theta), 0.0f);
        float hit_dist = sc.focal_distance / fabsf(dir.z);
        V3 focus = sc.cam_pos + dir * hit_dist;
        org = sc.cam_pos + diff;
        dir = unit(focus - org);
    }
}

// Shading point data of a hit (CalcParam + material)
static void hit_surface(const Scene &sc, const Ray &ray, const Hit &h, V3 &pos, V3 &normal,
                        const Material *&mat, float *uv = nullptr) {
    pos = ray.pos + ray.dir * h.distance;
    const Shape &s = sc.shapes[h.shape];
    if (h.prim == -2) {
        normal = unit(pos - s.center);
        mat = &sc.materials[s.material];
        if (uv) {
            // Sphere::CalcParam texcoord (s3d_sphere.cpp:94-104)
            const float theta = acosf(normal.y);
            float phi = atan2f(normal.x, normal.z);
            if (phi < 0.0f) phi += k2Pi;
            uv[0] = phi * 0.15915494309189533576888376337251f;
            uv[1] = (kPi - theta) * 0.31830988618379067153776752674503f;
        }
    } else {
        const Mesh *m = sc.meshes[s.mesh];
        const Tri &t = m->tris[m->prims[h.prim]];
        normal = tri_normal(t, h.b1, h.b2);   // object space for instances (no CalcParam of Instance)
        if (uv) {
            // Triangle::CalcParam texcoord (s3d_triangle.cpp:118-120)
            const float alpha = 1.0f - h.b1 - h.b2;
            uv[0] = t.uv[0][0] * alpha + t.uv[1][0] * h.b1 + t.uv[2][0] * h.b2;
            uv[1] = t.uv[0][1] * alpha + t.uv[1][1] * h.b1 + t.uv[2][1] * h.b2;
        }
        mat = &sc.materials[t.material];
        // instance material override (an extension of the reference's Instance)
        if (s.type == S3DG_SHAPE_INSTANCE && s.material >= 0) mat = &sc.materials[s.material];
    }
}

// PathTracer::NextEventEstimation (src/src/s3d_pt.cpp:251-292) with Scene::GetLight
// (s3d_scene.h:84-90).  QUAD lights are an extension (the
// reference samples spheres only).
static V3 nee(const Scene &sc, V3 position, V3 normal, const float *texcoord, const Material &mat, PCG &rng,
              uint64_t &shadow_rays) {
    uint32_t li = rng.u32() % (uint32_t)sc.lights.size();
    const Light &L = sc.lights[li];
    V3 light_pos;
    float light_pdf;
    if (L.type == S3DG_LIGHT_SHAPE) {
        const Shape &s = sc.shapes[L.shape];
        const float r1 = k2Pi * rng.f32();
        const float r2 = 1.0f - 2.0f * rng.f32();
        const float r3 = sqrtf(1.0f - r2 * r2);
        light_pos = s.center + (s.radius + 1e-1f) * safe_unit(V3(r3 * cosf(r1), r3 * sinf(r1), r2));
        light_pdf = 1.0f / (4.0f * kPi * s.radius * s.radius);
    } else {
        float u = rng.f32();
        float v = rng.f32();
        light_pos = L.corner + L.e1 * u + L.e2 * v;
        light_pdf = L.pdf;
    }
    V3 light_dir = light_pos - position;
    float light_dist2 = dot(light_dir, light_dir);
    light_dir = safe_unit(light_dir);
    Ray sr;
    sr.pos = position; sr.dir = light_dir; sr.tmin = kHitMin; sr.tmax = kHitMax;
    Hit sh;
    shadow_rays++;
    if (!scene_intersect(sc, sr, sh)) return V3(0, 0, 0);
    bool is_light;
    if (L.type == S3DG_LIGHT_SHAPE) is_light = (sh.shape == L.shape && sh.prim == -2);
    else {
        is_light = false;
        if (sh.shape == L.shape && sh.prim >= 0) {
            const Mesh *m = sc.meshes[sc.shapes[L.shape].mesh];
            int in = m->prims[sh.prim];
            is_light = in >= L.first && in < L.first + L.count;
        }
    }
    if (!is_light) return V3(0, 0, 0);
    V3 p, light_normal;
    const Material *lm;
    hit_surface(sc, sr, sh, p, light_normal, lm);
    float cos_shadow = fabsf(dot(normal, light_dir));
    float cos_light = fabsf(dot(light_normal, -light_dir));
    // GetBaseColor(texcoord): TexturedMaterial multiplies by the texture
    V3 base = base_color(mat);
    if (mat.texture >= 0) base = base * sc.textures[mat.texture].sample(texcoord[0], texcoord[1]);
    V3 fs = base / kPi;
    float G = (cos_shadow * cos_light) / light_dist2;
    float brdf_pdf = fabsf(sr.dir.z / kPi) * cos_light / light_dist2;
    float mis = light_pdf / (brdf_pdf + light_pdf);
    return lm->emissive * fs * mis * (G / light_pdf);
}

// PathTracer::Radiance (src/src/s3d_pt.cpp:182-246).  The miss term is
// SampleIBL(dir) = lat-long SampleColor(dir) * 10 (s3d_scene.h:80-82), or the
// constant environment when the scene has no IBL texture.
static V3 radiance(const Scene &sc, V3 org, V3 dir, PCG &rng, int max_depth, int rr_start,
                   uint64_t &ext_rays, uint64_t &shadow_rays) {
    V3 W(1, 1, 1), L(0, 0, 0);
    Ray ray;
    ray.pos = org; ray.dir = dir; ray.tmin = kHitMin; ray.tmax = kHitMax;
    for (int depth = 0; depth < max_depth; ++depth) {
        Hit h;
        ext_rays++;
        if (!scene_intersect(sc, ray, h)) {
            if (sc.env_tex >= 0) L = L + W * (sc.textures[sc.env_tex].sample_dir(ray.dir) * V3(10.0f, 10.0f, 10.0f));
            else L = L + W * sc.env;
            break;
        }
        V3 pos, normal;
        const Material *mat;
        float uv[2];
        hit_surface(sc, ray, h, pos, normal, mat, uv);
        L = L + W * mat->emissive;
        if (!has_delta(*mat) && !sc.lights.empty())
            L = L + W * nee(sc, pos, normal, uv, *mat, rng, shadow_rays);
        V3 out;
        bool dice = false;
        // TexturedMaterial::Shade = inner Shade times the texture
        V3 wt = shade(*mat, ray.dir, normal, rng, out, dice);
        if (mat->texture >= 0) wt = wt * sc.textures[mat->texture].sample(uv[0], uv[1]);
        W = W * wt;
        if (dice && depth >= rr_start) break;
        if (W.x < FLT_EPSILON && W.y < FLT_EPSILON && W.z < FLT_EPSILON) break;
        ray.pos = pos;
        ray.dir = out;
    }
    return L;
}

// ---------------------------------------------------------------------------
// Tonemapping: src/src/s3d_tonemapper.cpp:20-33 (RGBToY), 38-95
// (ComputeLogarithmicAverage), 100-130 (Uncharted2Tonemap, ACESFilm),
// 237-353 (operators); capture: src/src/s3d_pt.cpp:143-171.
// ---------------------------------------------------------------------------
static float rgb_to_y(V3 c) { return dot(V3(0.29900f, 0.58700f, 0.11400f), c); }
static void log_average(int w, int h, const V3 *px, float eps, float &ave, float &maxl) {
    bool init = false;
    ave = 0.0f;
    maxl = 0.0f;
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            float Lw = rgb_to_y(px[i * w + j]);
            if (Lw != Lw) Lw = eps;
            if (Lw < FLT_EPSILON) Lw = eps;
            if (!init) { maxl = Lw; init = true; } else maxl = smax(maxl, Lw);
            ave += logf(eps + Lw);
        }
    if (ave != ave) ave = eps;
    ave /= (float)(w * h);
    ave = expf(ave);
}
static float u2(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.01f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
static float aces1(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f, f = 0.665406f, g = 12.0f;
    return saturate((x * (a * x * f / g + b)) / (x * f / g * (c * x * f + d) + e));
}

}  // namespace orc

using namespace orc;

// ===========================================================================
// C interface
// ===========================================================================
extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}
void orc_pcg_stream(uint64_t seed, int32_t count, uint32_t *out) {
    PCG p;
    p.seed(seed);
    for (int i = 0; i < count; ++i) out[i] = p.u32();
}
// pcg32_srandom_r of the published PCG library (O'Neill 2014): inc = (seq<<1)|1,
// state = 0, step, state += initstate, step.  Uses the same step/output as PCG
// above so that the published demo vectors pin them.
void orc_pcg_srandom_stream(uint64_t initstate, uint64_t initseq, int32_t count, uint32_t *out) {
    uint64_t inc = (initseq << 1u) | 1u;
    uint64_t state = 0;
    auto step = [&]() {
        uint64_t old = state;
        state = old * PCG::kMul + inc;
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((~rot + 1) & 31));
    };
    step();
    state += initstate;
    step();
    for (int i = 0; i < count; ++i) out[i] = step();
}
float orc_pcg_f32(uint32_t x) { return (float)x / (float)0xffffffffu; }
uint64_t orc_path_seed(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t scene) {
    return path_seed(seed, pixel, pass, scene);
}

int orc_triangle_hit(const float v0[3], const float v1[3], const float v2[3], const s3dg_ray *ray,
                     s3dg_hit *hit) {
    s3dg_vertex v[3];
    memset(v, 0, sizeof(v));
    for (int i = 0; i < 3; ++i) {
        v[0].position[i] = v0[i];
        v[1].position[i] = v1[i];
        v[2].position[i] = v2[i];
    }
    Tri t = make_tri(v, 0, 0);
    Ray r;
    r.pos = load3(ray->org); r.dir = load3(ray->dir); r.tmin = ray->tmin; r.tmax = ray->tmax;
    float d, b1, b2;
    if (tri_hit(t, r, d, b1, b2, kFMax)) {
        hit->prim = 0; hit->t = d; hit->u = b1; hit->v = b2;
        return 1;
    }
    hit->prim = -1; hit->t = kFMax; hit->u = 0; hit->v = 0;
    return 0;
}
void orc_tangent_space(const float n[3], float t[3], float b[3]) {
    V3 T, B;
    tangent_space(load3(n), T, B);
    store3(t, T);
    store3(b, B);
}
void orc_sample_lambert(float u, float v, float out[3]) { store3(out, sample_lambert(u, v)); }
void orc_sample_phong(float u, float v, float p, float out[3]) { store3(out, sample_phong(u, v, p)); }
int orc_shade(const s3dg_material *m, const float input[3], const float normal[3], uint64_t *state,
              float out_dir[3], float weight[3]) {
    Material mat = make_material(*m);
    PCG p;
    p.state = *state;
    V3 o;
    bool dice = false;
    V3 w = shade(mat, load3(input), load3(normal), p, o, dice);
    *state = p.state;
    store3(out_dir, o);
    store3(weight, w);
    return dice ? 1 : 0;
}
void orc_invert(const float m[16], float out[16]) {
    Mat M;
    memcpy(M.m, m, sizeof(M.m));
    Mat R = invert(M);
    memcpy(out, R.m, sizeof(R.m));
}

void *orc_mesh_create(uint32_t vertex_count, const s3dg_vertex *vertices, int32_t kind) {
    if (vertex_count % 3 != 0 || (kind != 2 && kind != 4 && kind != 8)) return nullptr;
    return build_mesh(vertex_count, vertices, nullptr, 0, kind);
}
void orc_mesh_destroy(void *m) { delete (Mesh *)m; }
int orc_mesh_order(const void *mp, int32_t *order) {
    const Mesh *m = (const Mesh *)mp;
    for (size_t i = 0; i < m->prims.size(); ++i) order[i] = m->prims[i];
    return (int)m->prims.size();
}
int orc_mesh_tree(const void *mp, int32_t *kind, int32_t *cb, int32_t *cc, int32_t *children,
                  int32_t *lb, int32_t *lc, float *boxes) {
    const Mesh *m = (const Mesh *)mp;
    int nn = (int)m->nodes.size();
    if (!kind) return nn;
    int slot = 0;
    for (int i = 0; i < nn; ++i) {
        const Node &n = m->nodes[i];
        kind[i] = n.kind;
        cb[i] = slot;
        cc[i] = (int)n.child.size();
        for (int c : n.child) children[slot++] = c;
        lb[i] = n.begin;
        lc[i] = n.count;
        if (boxes) {
            store3(boxes + 6 * i, n.box.mini);
            store3(boxes + 6 * i + 3, n.box.maxi);
        }
    }
    return nn;
}
static void set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
int orc_mesh_trace(const void *mp, const s3dg_ray *rays, uint32_t count, int32_t method,
                   s3dg_hit *hits, int32_t threads) {
    const Mesh *m = (const Mesh *)mp;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < (long long)count; ++i) {
        Ray r;
        r.pos = load3(rays[i].org); r.dir = load3(rays[i].dir);
        r.tmin = rays[i].tmin; r.tmax = rays[i].tmax;
        Hit h;
        if (m->intersect(r, h, method == 1)) {
            hits[i].prim = m->prims[h.prim];
            hits[i].t = h.distance; hits[i].u = h.b1; hits[i].v = h.b2;
        } else {
            hits[i].prim = -1; hits[i].t = kFMax; hits[i].u = 0; hits[i].v = 0;
        }
    }
    return 0;
}

void *orc_scene_create(const s3dg_scene_desc *d) {
    Scene *sc = new Scene();
    for (uint32_t i = 0; i < d->material_count; ++i) sc->materials.push_back(make_material(d->materials[i]));
    for (uint32_t i = 0; i < d->mesh_count; ++i) {
        const s3dg_mesh_desc &md = d->meshes[i];
        int kind = md.bvh_kind == 2 || md.bvh_kind == 4 ? md.bvh_kind : 8;
        sc->meshes.push_back(build_mesh(md.vertex_count, md.vertices, md.triangle_material, md.material, kind));
    }
    for (uint32_t i = 0; i < d->shape_count; ++i) {
        const s3dg_shape_desc &s = d->shapes[i];
        Shape sh;
        sh.type = s.type; sh.mesh = s.mesh; sh.material = s.material;
        sh.center = load3(s.center); sh.radius = s.radius;
        Mat W;
        memcpy(W.m, s.world, sizeof(W.m));
        sh.inv = invert(W);
        sc->shapes.push_back(sh);
    }
    for (uint32_t i = 0; i < d->light_count; ++i) {
        const s3dg_light_desc &l = d->lights[i];
        Light L;
        L.type = l.type; L.shape = l.shape; L.first = l.first_triangle; L.count = l.triangle_count;
        L.corner = load3(l.corner); L.e1 = load3(l.edge1); L.e2 = load3(l.edge2);
        V3 c = cross(L.e1, L.e2);
        L.pdf = 1.0f / sqrtf(dot(c, c));
        sc->lights.push_back(L);
    }
    sc->env = load3(d->environment);
    for (uint32_t i = 0; i < d->texture_count; ++i) {
        Texture t;
        t.w = d->textures[i].width;
        t.h = d->textures[i].height;
        const float *rgb = d->textures[i].rgb;
        for (size_t k = 0; k < (size_t)t.w * t.h; ++k) t.px.push_back(V3(rgb[3 * k], rgb[3 * k + 1], rgb[3 * k + 2]));
        sc->textures.push_back(std::move(t));
    }
    sc->env_tex = d->environment_texture - 1;
    const s3dg_camera_desc &c = d->camera;
    sc->cam_type = c.type;
    sc->cam_pos = load3(c.position);
    V3 target = load3(c.target), up = load3(c.upward);
    sc->cam_dir = unit(target - sc->cam_pos);
    sc->CX = unit(cross(sc->cam_dir, up)) * c.fov * c.aspect;
    sc->CY = unit(cross(sc->cam_dir, sc->CX)) * c.fov;
    sc->CZ = sc->cam_pos + (sc->cam_dir * c.near_clip);
    sc->lens_radius = c.lens_radius;
    V3 tp = target - sc->cam_pos;
    sc->focal_distance = sqrtf((tp.x * tp.x) + (tp.y * tp.y) + (tp.z * tp.z));
    return sc;
}
void orc_scene_destroy(void *s) { delete (Scene *)s; }

int orc_scene_trace(const void *sp, const s3dg_ray *rays, uint32_t count, s3dg_hit *hits,
                    int32_t *shapes, int32_t threads) {
    const Scene *sc = (const Scene *)sp;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < (long long)count; ++i) {
        Ray r;
        r.pos = load3(rays[i].org); r.dir = load3(rays[i].dir);
        r.tmin = rays[i].tmin; r.tmax = rays[i].tmax;
        Hit h;
        if (scene_intersect(*sc, r, h)) {
            int prim = h.prim;
            if (prim >= 0) prim = sc->meshes[sc->shapes[h.shape].mesh]->prims[prim];
            hits[i].prim = prim; hits[i].t = h.distance; hits[i].u = h.b1; hits[i].v = h.b2;
            if (shapes) shapes[i] = h.shape;
        } else {
            hits[i].prim = -1; hits[i].t = kFMax; hits[i].u = 0; hits[i].v = 0;
            if (shapes) shapes[i] = -1;
        }
    }
    return 0;
}

void orc_camera_ray(const void *sp, float x, float y, float org[3], float dir[3]) {
    const Scene *sc = (const Scene *)sp;
    V3 o, d;
    camera_ray(*sc, x, y, nullptr, o, d);
    store3(org, o);
    store3(dir, d);
}

// The camera ray with ThinLensCamera's lens sample drawn from a PCG state
// (advanced in place), as a path's raygen does.
void orc_camera_ray_lens(const void *sp, float x, float y, uint64_t *state, float org[3], float dir[3]) {
    const Scene *sc = (const Scene *)sp;
    PCG rng;
    rng.state = *state;
    V3 o, d;
    camera_ray(*sc, x, y, &rng, o, d);
    *state = rng.state;
    store3(org, o);
    store3(dir, d);
}

// Texture::SampleColor by texcoord and by direction on a w x h RGB texel array.
static Texture make_texture(int32_t w, int32_t h, const float *rgb) {
    Texture t;
    t.w = w;
    t.h = h;
    t.px.resize((size_t)w * h);
    for (size_t i = 0; i < t.px.size(); ++i) t.px[i] = load3(rgb + 3 * i);
    return t;
}
void orc_texture_sample(int32_t w, int32_t h, const float *rgb, float u, float v, float out[3]) {
    store3(out, make_texture(w, h, rgb).sample(u, v));
}
void orc_texture_sample_dir(int32_t w, int32_t h, const float *rgb, const float dir[3], float out[3]) {
    store3(out, make_texture(w, h, rgb).sample_dir(load3(dir)));
}

// PathTracer::TracePath (src/src/s3d_pt.cpp:297-353): pixel-centre camera rays
// ((0.5+x)/W - 0.5, (0.5+y)/H - 0.5), one sample per pass added to a sum buffer.
int orc_render(const void *sp, const s3dg_render_config *cfg, int32_t row_begin, int32_t row_end,
               float *accum, uint64_t counts[3], int32_t threads) {
    const Scene *sc = (const Scene *)sp;
    const int W = cfg->width, H = cfg->height;
    uint64_t ext = 0, shd = 0, paths = 0;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic) reduction(+ : ext, shd, paths)
    for (int y = row_begin; y < row_end; ++y) {
        for (int x = 0; x < W; ++x) {
            const int idx = y * W + x;
            V3 sum = load3(accum + 3 * idx);
            for (int s = 0; s < cfg->spp; ++s) {
                PCG rng;
                rng.seed(path_seed(cfg->seed, (uint32_t)idx, (uint32_t)(cfg->first_pass + s), 0u));
                V3 o, d;
                camera_ray(*sc, (0.5f + (float)x) / (float)W - 0.5f,
                           (0.5f + (float)y) / (float)H - 0.5f, &rng, o, d);
                V3 L = radiance(*sc, o, d, rng, cfg->max_bounce_count, cfg->rr_start_depth, ext, shd);
                sum = sum + L;
                paths++;
            }
            store3(accum + 3 * idx, sum);
        }
    }
    if (counts) { counts[0] = ext; counts[1] = shd; counts[2] = paths; }
    return 0;
}

int orc_render_guides(const void *sp, const s3dg_render_config *cfg, float *albedo, float *normal,
                      float *depth) {
    const Scene *sc = (const Scene *)sp;
    const int W = cfg->width, H = cfg->height;
#pragma omp parallel for schedule(dynamic)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int idx = y * W + x;
            PCG rng;
            rng.seed(path_seed(cfg->seed, (uint32_t)idx, (uint32_t)cfg->first_pass, 0u));
            V3 o, d;
            camera_ray(*sc, (0.5f + (float)x) / (float)W - 0.5f, (0.5f + (float)y) / (float)H - 0.5f,
                       &rng, o, d);
            Ray r;
            r.pos = o; r.dir = d; r.tmin = kHitMin; r.tmax = kHitMax;
            Hit h;
            V3 a(0, 0, 0), n(0, 0, 0);
            float t = 0.0f;
            if (scene_intersect(*sc, r, h)) {
                V3 p;
                const Material *m;
                float uv[2];
                hit_surface(*sc, r, h, p, n, m, uv);
                a = base_color(*m);
                if (m->texture >= 0) a = a * sc->textures[m->texture].sample(uv[0], uv[1]);
                t = h.distance;
            }
            store3(albedo + 3 * idx, a);
            store3(normal + 3 * idx, n);
            depth[idx] = t;
        }
    return 0;
}

int orc_tonemap(int32_t type, int32_t w, int32_t h, const float *in, float *out, float *ave_out,
                float *max_out) {
    const V3 *px = (const V3 *)in;
    float ave, maxl;
    log_average(w, h, px, 0.00001f, ave, maxl);
    if (ave_out) *ave_out = ave;
    if (max_out) *max_out = maxl;
    const int n = w * h;
    if (type == S3DG_TONEMAP_ACES) {
        float coeff = 0.27f / ave;
        for (int i = 0; i < n; ++i) {
            V3 t = px[i] * coeff;
            store3(out + 3 * i, V3(aces1(t.x), aces1(t.y), aces1(t.z)));
        }
    } else if (type == S3DG_TONEMAP_UNCHARTED2) {
        float coeff = 0.18f / ave;
        const float white = u2(11.2f);
        for (int i = 0; i < n; ++i) {
            V3 t = px[i] * coeff;
            V3 c(u2(2.0f * t.x), u2(2.0f * t.y), u2(2.0f * t.z));
            store3(out + 3 * i, c / white);
        }
    } else {
        float coeff = 0.18f / ave;
        float m2 = maxl * coeff;
        m2 *= m2;
        for (int i = 0; i < n; ++i) {
            V3 L = px[i] * coeff;
            V3 Ld = L * (1.0f + (L / m2)) / (1.0f + L);
            store3(out + 3 * i, Ld);
        }
    }
    return 0;
}

int orc_capture(int32_t w, int32_t h, const float *in, uint8_t *out) {
    const int n = w * h;
    for (int i = 0; i < n; ++i) {
        float c[3] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
        for (int k = 0; k < 3; ++k) {
            float r = c[k];
            if (r > 1.0f) r = 1.0f;
            if (r < 0.0f) r = 0.0f;
            r = (r <= 0.0031308f) ? 12.92f * r : powf(1.055f * r, 1.0f / 2.4f) - 0.055f;
            out[3 * i + k] = (uint8_t)(r * 255.0f + 0.5f);
        }
    }
    return 0;
}

// FilterNLM (src/src/s3d_denoiser.cpp:19-95) with the patch and search radii
// as parameters (the reference fixes kKernel=5, kSupport=13).  Templates are
// walked sx-outer / sy-inner; |a-b|^2 of a Color4 sums (xx+yy)+(zz+ww) like
// _mm_dp_ps(.., 0xff).
int orc_nlm(int32_t width, int32_t height, float coeff, int32_t rs, int32_t rp, const float *src,
            float *dst, int32_t threads) {
    const float param_h = smax(0.0001f, coeff);
    const float sigma = smax(0.0001f, coeff);
    const float frac_h2 = 1.0f / (param_h * param_h);
    const float sigma2 = sigma * sigma;
    auto at = [&](int x, int y) -> const float * {
        int sx = x < 0 ? 0 : (x > width - 1 ? width - 1 : x);
        int sy = y < 0 ? 0 : (y > height - 1 ? height - 1 : y);
        return src + 4 * (sy * width + sx);
    };
    set_threads(threads);
#pragma omp parallel for schedule(dynamic)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            float sum[4] = {0, 0, 0, 0};
            float sw = 0.0f;
            for (int sx = x - rs; sx <= x + rs; ++sx)
                for (int sy = y - rs; sy <= y + rs; ++sy) {
                    float accum = 0.0f;
                    for (int px = -rp; px <= rp; ++px)
                        for (int py = -rp; py <= rp; ++py) {
                            const float *a = at(x + px, y + py);
                            const float *b = at(sx + px, sy + py);
                            float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2], d3 = a[3] - b[3];
                            accum += ((d0 * d0) + (d1 * d1)) + ((d2 * d2) + (d3 * d3));
                        }
                    float arg = -smax(accum - 2.0f * sigma2, 0.0f) * frac_h2;
                    float wgt = expf(arg);
                    sw += wgt;
                    const float *s = at(sx, sy);
                    for (int k = 0; k < 4; ++k) sum[k] += s[k] * wgt;
                }
            float *o = dst + 4 * (y * width + x);
            for (int k = 0; k < 4; ++k) o[k] = sum[k] / sw;
        }
    }
    return 0;
}

}  // extern "C"
