// shading.h -- gfx950 device code for the path loop: counter-keyed PCG streams,
// camera rays, surface parameters and the reference materials.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "gpu_types.h"

namespace s3dg {

constexpr float kPi = 3.1415926535897932384626433832795f;     // F_PI
constexpr float k2Pi = 6.283185307179586476925286766559f;     // F_2PI

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return v3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
// Vector3::UnitVector / SafeUnitVector (src/include/s3d_math.h:829-855)
__device__ __forceinline__ V3 unit3(V3 v) {
    const float m = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return v3(v.x / m, v.y / m, v.z / m);
}
__device__ __forceinline__ V3 safe_unit3(V3 v) {
    const float m = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (m > 0.0f) return v3(v.x / m, v.y / m, v.z / m);
    return v;
}
// Vector3::Reflect (s3d_math.h:892-902)
__device__ __forceinline__ V3 reflect3(V3 i, V3 n) {
    const float d = (n.x * i.x) + (n.y * i.y) + (n.z * i.z);
    const float d2 = 2.0f * d;
    return v3(i.x - (d2 * n.x), i.y - (d2 * n.y), i.z - (d2 * n.z));
}
__device__ __forceinline__ V3 ld3(const float *p) { return v3(p[0], p[1], p[2]); }

// PCG (s3d_math.h:2769-2817): XSH-RR output of a 64-bit LCG with the fixed
// increment; SetSeed adds the increment and steps once.
struct Pcg {
    uint64_t state;
    __device__ __forceinline__ void seed(uint64_t s) {
        state = s + 1442695040888963407ull;
        (void)u32();
    }
    __device__ __forceinline__ uint32_t u32() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + 1442695040888963407ull;
        const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        const uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    // GetAsF32: float(u32) / float(0xffffffff); may return exactly 1.0f
    __device__ __forceinline__ float f32() { return (float)u32() / 4294967296.0f; }
};

// Philox4x32-10 (Salmon et al., SC'11) keyed by the render seed; the counter
// names the path: {pixel, pass, scene, 0}.  Its first two words seed the
// path's PCG stream.
__device__ __forceinline__ uint64_t path_seed(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t scene) {
    uint32_t c0 = pixel, c1 = pass, c2 = scene, c3 = 0u;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return (uint64_t)c0 | ((uint64_t)c1 << 32);
}

// TangentSpace (s3d_math.h:2875-2886; Duff et al. 2017, Listing 3)
__device__ __forceinline__ void tangent_space(V3 N, V3 &T, V3 &B) {
    const float s = (N.z >= 0.0f) ? 1.0f : -1.0f;
    const float a = -1.0f / (s + N.z);
    const float b = N.x * N.y * a;
    T = v3(1.0f + s * N.x * N.x * a, s * b, -s * N.x);
    B = v3(b, s + N.y * N.y * a, -N.y);
}
// SampleLambert / SamplePhong (s3d_math.h:2968-2982)
__device__ __forceinline__ V3 sample_lambert(float u, float v) {
    const float r = sqrtf(1.0f - u * u);
    const float p = k2Pi * v;
    return v3(r * cosf(p), r * sinf(p), u);
}
__device__ __forceinline__ V3 sample_phong(float u, float v, float shininess) {
    const float p = k2Pi * u;
    const float ct = powf(1.0f - v, 1.0f / (shininess + 1.0f));
    const float st = sqrtf(1.0f - ct * ct);
    return v3(st * cosf(p), st * sinf(p), ct);
}
__device__ __forceinline__ float safe_sqrtf(float v) { return v > FLT_EPSILON ? sqrtf(v) : 0.0f; }

__device__ __forceinline__ V3 base_color(const GMat &m) {
    return (m.type == S3DG_MAT_LAMBERT || m.type == S3DG_MAT_PLASTIC) ? ld3(m.diffuse) : ld3(m.specular);
}

constexpr float k1DivPi = 0.31830988618379067153776752674503f;    // F_1DIVPI
constexpr float k1Div2Pi = 0.15915494309189533576888376337251f;   // F_1DIV2PI

// static_cast<s32>(floor(x)) with NaN taken to 0 and the range held to
// +-2^30, so that the corner arithmetic below cannot overflow.
__device__ __forceinline__ int32_t floor_i32(float x) {
    const float f = floorf(x);
    if (!(f == f)) return 0;
    return (int32_t)fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
}

// Texture::SampleColor(const Vector2&) (src/src/s3d_texture.cpp:925-952):
// bilinear weights from the corners around uv * (size - 1), the corners then
// wrapped by abs(i % size) (a truncating remainder: negative coordinates
// reflect rather than repeat).
__device__ __forceinline__ V3 texel(const float *texels, const GTexture &t, int32_t x, int32_t y) {
    const float *p = texels + t.offset + 3ull * ((uint64_t)(uint32_t)t.width * (uint32_t)y + (uint32_t)x);
    return v3(p[0], p[1], p[2]);
}
__device__ __forceinline__ V3 tex_sample(const float *texels, const GTexture &t, float u, float v) {
    const float fx = u * (float)(t.width - 1);
    const float fy = v * (float)(t.height - 1);
    int32_t x0 = floor_i32(fx), y0 = floor_i32(fy);
    int32_t x1 = x0 + 1, y1 = y0 + 1;
    const float xw0 = (float)x1 - fx, xw1 = fx - (float)x0;
    const float yw0 = (float)y1 - fy, yw1 = fy - (float)y0;
    x0 = abs(x0 % t.width);
    x1 = abs(x1 % t.width);
    y0 = abs(y0 % t.height);
    y1 = abs(y1 % t.height);
    return (texel(texels, t, x0, y0) * yw0 + texel(texels, t, x0, y1) * yw1) * xw0 +
           (texel(texels, t, x1, y0) * yw0 + texel(texels, t, x1, y1) * yw1) * xw1;
}

// Texture::SampleColor(const Vector3&) (src/src/s3d_texture.cpp:957-972):
// latitude-longitude lookup; u stays 0 when x or z is within FLT_EPSILON of 0.
__device__ __forceinline__ V3 tex_sample_dir(const float *texels, const GTexture &t, V3 d) {
    float u = 0.0f;
    const float v = acosf(d.y) * k1DivPi;
    if (!(fabsf(d.x) <= FLT_EPSILON) && !(fabsf(d.z) <= FLT_EPSILON)) {
        const float phi = atan2f(d.z, d.x);
        u = ((d.z < 0.0f) ? (phi + k2Pi) : phi) * k1Div2Pi;
    }
    return tex_sample(texels, t, u, v);
}

// Texture coordinates of a hit (the texcoord half of CalcParam): barycentric
// blend of a triangle's vertex coordinates (src/src/s3d_triangle.cpp:107-121),
// or a sphere's longitude / latitude (src/src/s3d_sphere.cpp:94-104).
__device__ __forceinline__ void tri_texcoord(const GTriUV &w, float beta, float gamma, float &u, float &v) {
    const float alpha = 1.0f - beta - gamma;
    u = ((w.uv[0] * alpha) + (w.uv[2] * beta)) + (w.uv[4] * gamma);
    v = ((w.uv[1] * alpha) + (w.uv[3] * beta)) + (w.uv[5] * gamma);
}
__device__ __forceinline__ void sphere_texcoord(V3 n, float &u, float &v) {
    const float theta = acosf(n.y);
    float phi = atan2f(n.x, n.z);
    if (phi < 0.0f) phi += k2Pi;
    u = phi * k1Div2Pi;
    v = (kPi - theta) * k1DivPi;
}

// IMaterial::Shade of Lambert (src/src/s3d_lambert.cpp:64-86), Mirror
// (s3d_mirror.cpp:59-71), Glass (s3d_glass.cpp:60-128), Phong (s3d_phong.cpp:65-88)
// and Plastic (s3d_plastic.cpp:71-142).  Two draws passed as arguments of one
// call in the reference are taken left to right.
__device__ __forceinline__ V3 shade(const GMat &m, V3 input, V3 normal, Pcg &rng, V3 &output, bool &dice) {
    const V3 diffuse = ld3(m.diffuse), specular = ld3(m.specular);
    if (m.type == S3DG_MAT_LAMBERT) {
        const V3 N = (dot3(normal, input) < 0.0f) ? normal : -normal;
        V3 T, B;
        tangent_space(N, T, B);
        const float u = rng.f32();
        const float v = rng.f32();
        const V3 s = sample_lambert(u, v);
        output = safe_unit3(T * s.x + B * s.y + N * s.z);
        dice = (rng.f32() >= m.thr0);
        return diffuse;
    }
    if (m.type == S3DG_MAT_MIRROR) {
        const V3 nm = (dot3(normal, input) < 0.0f) ? normal : -normal;
        output = safe_unit3(reflect3(input, nm));
        dice = false;
        return specular;
    }
    if (m.type == S3DG_MAT_GLASS) {
        const V3 nm = (dot3(normal, input) < 0.0f) ? normal : -normal;
        const V3 refl = safe_unit3(reflect3(input, nm));
        const bool into = dot3(normal, nm) > 0.0f;
        dice = false;
        const float nc = 1.0f, nt = m.ior;
        const float nnt = into ? (nc / nt) : (nt / nc);
        const float ddn = dot3(input, nm);
        const float cos2t = 1.0f - nnt * nnt * (1.0f - ddn * ddn);
        if (cos2t < 0.0f) {
            output = refl;
            return specular;
        }
        const V3 refr = safe_unit3(input * nnt - normal * (into ? 1.0f : -1.0f) * (ddn * nnt + safe_sqrtf(cos2t)));
        const float a = nt - nc, b = nt + nc;
        const float R0 = (a * a) / (b * b);
        const float c = 1.0f - (into ? -ddn : dot3(refr, normal));
        const float c2 = c * c;
        const float Re = R0 + (1.0f - R0) * (c2 * c2 * c);
        const float Tr = 1.0f - Re;
        const float P = 0.25f + 0.5f * Re;
        const float prob = 0.5f;
        if (rng.f32() < P) {
            output = refl;
            return specular * Re / (P * prob);
        }
        output = refr;
        return specular * Tr / ((1.0f - P) * prob);
    }
    if (m.type == S3DG_MAT_PHONG) {
        const float u = rng.f32();
        const float v = rng.f32();
        const V3 s = sample_phong(u, v, m.power);
        const V3 n = (dot3(input, normal) < 0.0f) ? normal : -normal;
        V3 w = reflect3(input, n);
        {
            const float mag = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
            if (mag > 0.0f) w = v3(w.x / mag, w.y / mag, w.z / mag);
        }
        V3 T, B;
        tangent_space(w, T, B);
        const V3 dir = safe_unit3(T * s.x + B * s.y + w * s.z);
        const float cosine = fabsf(dot3(dir, n));
        output = dir;
        dice = (rng.f32() >= m.thr0);
        return specular * cosine;
    }
    // Plastic
    float cosine = dot3(normal, input);
    const V3 nm = (cosine < 0.0f) ? normal : -normal;
    if (cosine < 0.0f) cosine = -cosine;
    const float temp1 = 1.0f - cosine;
    const float R0 = 0.5f;
    const float R = R0 + (1.0f - R0) * temp1 * temp1 * temp1 * temp1 * temp1;
    const float P = (R + 0.5f) / 2.0f;
    if (rng.f32() <= P) {
        V3 T, B;
        tangent_space(nm, T, B);
        const float u = rng.f32();
        const float v = rng.f32();
        const V3 s = sample_lambert(u, v);
        output = unit3(T * s.x + B * s.y + nm * s.z);
        const V3 result = diffuse * R / P;
        dice = (rng.f32() >= m.thr0);
        return result;
    }
    const float u = rng.f32();
    const float v = rng.f32();
    const V3 s = sample_phong(u, v, m.power);
    V3 w = unit3(reflect3(input, nm));
    V3 T, B;
    tangent_space(w, T, B);
    const V3 dir = unit3(T * s.x + B * s.y + w * s.z);
    const float dots = dot3(dir, nm);
    output = dir;
    dice = (rng.f32() >= m.thr1);
    return specular * dots * (1.0f - R) / (1.0f - P);
}

// Camera ray through (x, y) in [-0.5, 0.5]^2: PinholeCamera::GetRay
// (src/include/s3d_camera.h:95-100); ThinLensCamera::GetRay (207-225) draws
// its lens sample from the path's stream.
__device__ __forceinline__ void camera_ray(const GCamera &c, float x, float y, Pcg &rng, V3 &org, V3 &dir) {
    const V3 pos = (ld3(c.cx) * x) + (ld3(c.cy) * y) + ld3(c.cz);
    const V3 cp = ld3(c.pos);
    dir = unit3(pos - cp);
    org = cp;
    if (c.type == S3DG_CAMERA_THINLENS && c.lens_radius > 0.0f) {
        const float theta = k2Pi * rng.f32();
        const float r = c.lens_radius * safe_sqrtf(rng.f32());
        const V3 diff = v3(r * cosf(theta), r * sinf(theta), 0.0f);
        const float hit_dist = c.focal_distance / fabsf(dir.z);
        const V3 focus = cp + dir * hit_dist;
        org = cp + diff;
        dir = unit3(focus - org);
    }
}

// Sphere::Sample (src/src/s3d_sphere.cpp:106-115): a point uniform on the
// sphere of radius + 0.1 about the centre, and the pdf 1 / (4 pi r^2).
__device__ __forceinline__ V3 sphere_sample_point(V3 center, float radius, Pcg &rng) {
    const float r1 = k2Pi * rng.f32();
    const float r2 = 1.0f - 2.0f * rng.f32();
    const float r3 = sqrtf(1.0f - r2 * r2);
    return center + safe_unit3(v3(r3 * cosf(r1), r3 * sinf(r1), r2)) * (radius + 1e-1f);
}
__device__ __forceinline__ float sphere_sample_pdf(float radius) { return 1.0f / (4.0f * kPi * radius * radius); }

// Triangle::CalcParam normal (src/src/s3d_triangle.cpp:107-121) + SafeNormalize
__device__ __forceinline__ V3 tri_normal(const GTriShade &s, float beta, float gamma) {
    const float alpha = 1.0f - beta - gamma;
    V3 n = v3((s.n0[0] * alpha) + (s.n1[0] * beta) + (s.n2[0] * gamma),
              (s.n0[1] * alpha) + (s.n1[1] * beta) + (s.n2[1] * gamma),
              (s.n0[2] * alpha) + (s.n1[2] * beta) + (s.n2[2] * gamma));
    const float mag = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    if (mag > 0.0f) n = v3(n.x / mag, n.y / mag, n.z / mag);
    return n;
}

}  // namespace s3dg
