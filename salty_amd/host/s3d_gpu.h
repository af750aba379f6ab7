// s3d_gpu.h -- C++ drop-in host layer over the C ABI (include/s3dg.h).
//
// Keeps the reference's host interfaces, with their signatures, so that scene,
// shape, material, camera, BVH, post-processing and render-config code written
// against them compiles against this header:
//   IReference            src/include/s3d_reference.h:19-25
//   Vertex / HitRecord    src/include/s3d_shape.h:27-54
//   IShape                src/include/s3d_shape.h:60-68 (IsHit, GetBox, GetCenter,
//                         CalcParam, Sample)
//   Triangle / Leaf       Triangle::Create (s3d_triangle.h:40), Leaf::Create (s3d_leaf.h:41)
//   BVH8 / BVH4 / BVH2    ::Create(size_t, IShape**) (s3d_bvh8.h:41, s3d_bvh4.h, s3d_bvh2.h),
//                         reordering the array in place as the reference does
//   Mesh / Instance / Sphere  Mesh::Create (s3d_mesh.h:42-47), Instance::Create
//                         (s3d_instance.h:40), Sphere::Create (s3d_sphere.h:40)
//   IMaterial / ShadingArg    src/include/s3d_material.h:21-42 (Shade, GetEmissive,
//                         GetBaseColor, HasDelta); MaterialFactory (s3d_materialfactory.h:22-49)
//   ICamera               src/include/s3d_camera.h:19-25; PinholeCamera / ThinLensCamera
//   PCG                   src/include/s3d_math.h:2769-2817
//   Texture               src/include/s3d_texture.h (Init from BMP / HDR files)
//   Scene                 src/include/s3d_scene.h:22-109 (GetRay, Intersect, SampleIBL,
//                         GetLight, Update)
//   PathTracer::Config    src/include/s3d_pt.h:33-40, PathTracer::Run
//   ToneMapper::Map       src/include/s3d_tonemapper.h:33
//   FilterNLM             src/include/s3d_denoiser.h:16
//
// Every geometric and shading call runs on the GPU through the C ABI (no host
// restatement of the hot path): a shape traced on its own (Triangle,
// Mesh, BVH, Instance, Sphere) keeps a device scene of itself, built on first
// use, and IsHit is one s3dg_trace of that scene; Shade, GetBaseColor, GetRay,
// Sample, CalcParam and SampleIBL are s3dg_shade, s3dg_base_color,
// s3dg_camera_rays, s3dg_sphere_sample, s3dg_scene_calc_param and
// s3dg_scene_environment.  These per-call paths carry one host round trip
// each; whole frames go through Scene + PathTracer (s3dg_render), and batches
// of rays through the batched Scene::Intersect overload or s3dg_trace on
// Scene::GetDeviceScene().
//
// All device work uses one process-wide context (DefaultContext), and the
// calls are serialised by one mutex, so the const methods stay reentrant as
// the reference's are.  Objects are reference counted like the reference's
// (count 1 on creation, Release deletes at 0, containers AddRef their
// children); factories return nullptr on failure instead of throwing.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/s3dg.h"

namespace s3d {

typedef int32_t s32;
typedef uint32_t u32;
typedef uint64_t u64;
typedef float f32;

const f32 F_MAX = 3.402823466e+38F;   // s3d_math.h:46
const f32 F_HIT_MAX = 1e12f;          // s3d_math.h:44
const f32 F_HIT_MIN = 1e-3f;          // s3d_math.h:45

// ---------------------------------------------------------------------------
// Math types (the subset of s3d_math.h a caller of these interfaces uses)
// ---------------------------------------------------------------------------
struct Vector2 {
    f32 x, y;
    Vector2() : x(0), y(0) {}
    Vector2(f32 a, f32 b) : x(a), y(b) {}
};

struct Vector3 {
    f32 x, y, z;
    Vector3() : x(0), y(0), z(0) {}
    Vector3(f32 a, f32 b, f32 c) : x(a), y(b), z(c) {}
    Vector3 operator+(const Vector3 &v) const { return Vector3(x + v.x, y + v.y, z + v.z); }
    Vector3 operator-(const Vector3 &v) const { return Vector3(x - v.x, y - v.y, z - v.z); }
    Vector3 operator-() const { return Vector3(-x, -y, -z); }
    Vector3 operator*(const Vector3 &v) const { return Vector3(x * v.x, y * v.y, z * v.z); }
    Vector3 operator*(f32 s) const { return Vector3(x * s, y * s, z * s); }
    Vector3 operator/(f32 s) const { return Vector3(x / s, y / s, z / s); }
    f32 Length() const { return std::sqrt(x * x + y * y + z * z); }
    static f32 Dot(const Vector3 &a, const Vector3 &b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
    static Vector3 Cross(const Vector3 &a, const Vector3 &b) {
        return Vector3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
    }
    static Vector3 UnitVector(const Vector3 &v) {
        const f32 m = v.Length();
        return Vector3(v.x / m, v.y / m, v.z / m);
    }
    static Vector3 Min(const Vector3 &a, const Vector3 &b) {
        return Vector3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z);
    }
    static Vector3 Max(const Vector3 &a, const Vector3 &b) {
        return Vector3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z);
    }
};

struct Vector4 {
    f32 x, y, z, w;
};
typedef Vector3 Color3;
typedef Vector4 Color4;

// Matrix in the reference's row-vector convention (_41.._43 = translation).
struct Matrix {
    f32 m[4][4];
    static Matrix Identity() {
        Matrix r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[i][j] = i == j ? 1.0f : 0.0f;
        return r;
    }
    static Matrix Translate(f32 x, f32 y, f32 z) {
        Matrix r = Identity();
        r.m[3][0] = x;
        r.m[3][1] = y;
        r.m[3][2] = z;
        return r;
    }
    // Vector3::Transform (s3d_math.h:2082-2089)
    Vector3 Transform(const Vector3 &p) const {
        return Vector3((p.x * m[0][0]) + (p.y * m[1][0]) + (p.z * m[2][0]) + m[3][0],
                       (p.x * m[0][1]) + (p.y * m[1][1]) + (p.z * m[2][1]) + m[3][1],
                       (p.x * m[0][2]) + (p.y * m[1][2]) + (p.z * m[2][2]) + m[3][2]);
    }
};

// BoundingBox (s3d_math.h:2145-2272): an empty box is (+F_MAX, -F_MAX).
struct BoundingBox {
    Vector3 mini, maxi, center;
    bool empty;
    BoundingBox() : mini(F_MAX, F_MAX, F_MAX), maxi(-F_MAX, -F_MAX, -F_MAX), center(), empty(true) {}
    explicit BoundingBox(const Vector3 &p) : mini(p), maxi(p), center(p), empty(false) {}
    BoundingBox(const Vector3 &lo, const Vector3 &hi) : mini(lo), maxi(hi), center((lo + hi) * 0.5f), empty(false) {}
    static BoundingBox Merge(const BoundingBox &a, const BoundingBox &b) {
        if (!a.empty && !b.empty) return BoundingBox(Vector3::Min(a.mini, b.mini), Vector3::Max(a.maxi, b.maxi));
        return a.empty ? b : a;
    }
    // BoundingBox::Transform: the two corners only, as the reference does
    static BoundingBox Transform(const BoundingBox &b, const Matrix &m) {
        return BoundingBox(m.Transform(b.mini), m.Transform(b.maxi));
    }
};

struct Ray {
    Vector3 pos;
    Vector3 dir;
    Ray() {}
    Ray(const Vector3 &p, const Vector3 &d) : pos(p), dir(d) {}
};

// RaySet (s3d_math.h:1437-1469) holds the ray broadcast into SSE/AVX lanes for
// the reference's box tests; here the ray itself is all a shape reads.
struct RaySet {
    Ray ray;
};
inline RaySet MakeRaySet(const Vector3 &pos, const Vector3 &dir) {
    RaySet r;
    r.ray = Ray(pos, dir);
    return r;
}

// PCG (s3d_math.h:2769-2817): PCG32 XSH-RR with the reference's seeding
// (state = seed + increment, then one step).
class PCG {
public:
    PCG() { SetSeed(123456789); }
    explicit PCG(u64 seed) { SetSeed(seed); }
    void SetSeed(u64 seed) {
        m_State = seed + kIncrement;
        GetAsU32();
    }
    u32 GetAsU32() {
        const u64 old = m_State;
        m_State = old * kMultiplier + kIncrement;
        const u32 xorshifted = (u32)(((old >> 18) ^ old) >> 27);
        const u32 rot = (u32)(old >> 59);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1) & 31));
    }
    f32 GetAsF32() { return (f32)GetAsU32() / (f32)0xffffffffu; }
    // extension: the raw state, which the device draws from
    u64 GetState() const { return m_State; }
    void SetState(u64 s) { m_State = s; }

private:
    static const u64 kMultiplier = 6364136223846793005ull;
    static const u64 kIncrement = 1442695040888963407ull;
    u64 m_State = 0x4d595df4d0f33173ull;
};

// Vertex: same 32-byte layout as s3dg_vertex.
struct Vertex {
    Vector3 Position;
    Vector3 Normal;
    Vector2 TexCoord;
};
static_assert(sizeof(Vertex) == sizeof(s3dg_vertex), "Vertex layout");

// IReference (s3d_reference.h:19-25)
struct IReference {
    virtual ~IReference() {}
    void AddRef() { ++m_Count; }
    void Release() {
        if (--m_Count == 0) delete this;
    }
    u32 GetCount() const { return m_Count; }

private:
    std::atomic<u32> m_Count{1};
};

template <class T>
inline void SafeRelease(T *&p) {
    if (p) {
        p->Release();
        p = nullptr;
    }
}

struct IShape;
struct IMaterial;

// HitRecord (s3d_shape.h:38-54).  shape and triangle are extensions: the
// top-level shape index of a Scene::Intersect hit and the triangle's index in
// its mesh's input order (-2: a sphere).
struct HitRecord {
    f32 distance = F_MAX;
    Vector2 barycentric;
    const IShape *pShape = nullptr;
    const IMaterial *pMaterial = nullptr;
    s32 shape = -1;
    s32 triangle = -1;
};

// IShape (s3d_shape.h:60-68)
struct IShape : IReference {
    virtual ~IShape() {}
    virtual bool IsHit(const RaySet &, HitRecord &) const = 0;
    virtual BoundingBox GetBox() const = 0;
    virtual Vector3 GetCenter() const = 0;
    virtual void CalcParam(const Vector3 &, const Vector2 &, Vector3 *, Vector2 *) const {}
    virtual void Sample(PCG &, Vector3 *, float *) {}
};

// ShadingArg / IMaterial (s3d_material.h:21-42)
struct ShadingArg {
    Vector3 input;
    Vector3 output;
    Vector3 normal;
    Vector2 texcoord;
    PCG random;
    bool dice = false;
};

struct IMaterial : IReference {
    virtual ~IMaterial() {}
    virtual Color3 Shade(ShadingArg &) const = 0;
    virtual Color3 GetEmissive() const = 0;
    virtual Color3 GetBaseColor(const Vector2 &texcoord) const = 0;
    virtual bool HasDelta() const = 0;
};

// ICamera (s3d_camera.h:19-25)
struct ICamera {
    virtual ~ICamera() {}
    virtual Ray GetRay(const f32 x, const f32 y) = 0;
};

// ---------------------------------------------------------------------------
// The process-wide device context and the lock that serialises its calls.
// ---------------------------------------------------------------------------
namespace detail {
struct ContextHolder {
    s3dg_context *ctx = nullptr;
    ContextHolder() {
        if (s3dg_create(&ctx, -1) != S3DG_OK) ctx = nullptr;
    }
    ~ContextHolder() {
        if (ctx) s3dg_destroy(ctx);
    }
};
inline ContextHolder &Holder() {
    static ContextHolder h;
    return h;
}
inline std::recursive_mutex &Lock() {
    static std::recursive_mutex m;
    return m;
}
}  // namespace detail

// The context every call of this layer uses (device 0 unless the HIP device
// was set before first use); nullptr without a usable gfx950 device.
inline s3dg_context *DefaultContext() { return detail::Holder().ctx; }

// ---------------------------------------------------------------------------
// Textures (s3d::Texture): RGB texels in memory, loaded from the reference's
// BMP and HDR formats by the library's loaders (TGA is not supported here).
// ---------------------------------------------------------------------------
enum TEXTURE_FILE_FORMAT { TFF_BMP = 0, TFF_TGA, TFF_HDR };

class Texture {
public:
    bool Init(const char *path, TEXTURE_FILE_FORMAT format) {
        Term();
        int32_t w = 0, h = 0;
        float *rgb = nullptr;
        int rc = S3DG_ERR_UNSUPPORTED;
        if (format == TFF_BMP) rc = s3dg_image_load_bmp(path, &w, &h, &rgb);
        else if (format == TFF_HDR) rc = s3dg_image_load_hdr(path, &w, &h, &rgb);
        if (rc != S3DG_OK) return false;
        m_Width = w;
        m_Height = h;
        m_Texels.assign(rgb, rgb + 3 * (size_t)w * h);
        s3dg_image_free(rgb);
        return true;
    }
    void Term() {
        m_Width = m_Height = 0;
        m_Texels.clear();
    }
    bool IsValid() const { return m_Width > 0 && m_Height > 0; }
    s32 GetWidth() const { return m_Width; }
    s32 GetHeight() const { return m_Height; }
    const float *GetTexels() const { return m_Texels.data(); }
    s3dg_texture_desc Desc() const { return s3dg_texture_desc{m_Width, m_Height, m_Texels.data()}; }

private:
    s32 m_Width = 0, m_Height = 0;
    std::vector<float> m_Texels;
};

// ---------------------------------------------------------------------------
// Materials: the reference's five (with TexturedMaterial when a texture is
// given), shaded on the device by the renderer's own code.
// ---------------------------------------------------------------------------
class Material : public IMaterial {
public:
    s3dg_material desc{};
    const Texture *texture = nullptr;   // TexturedMaterial wrapper when set

    Color3 Shade(ShadingArg &arg) const override {
        s3dg_shade_arg a = make_arg(arg);
        a.random = arg.random.GetState();
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        s3dg_context *ctx = DefaultContext();
        s3dg_texture_desc td{};
        if (texture) td = texture->Desc();
        if (!ctx || s3dg_shade(ctx, &desc, texture ? &td : nullptr, 1, &a) != S3DG_OK) {
            arg.dice = true;   // the path cannot go on
            return Color3();
        }
        arg.output = Vector3(a.output[0], a.output[1], a.output[2]);
        arg.dice = a.dice != 0;
        arg.random.SetState(a.random);
        return Color3(a.weight[0], a.weight[1], a.weight[2]);
    }
    Color3 GetEmissive() const override { return Color3(desc.emissive[0], desc.emissive[1], desc.emissive[2]); }
    Color3 GetBaseColor(const Vector2 &texcoord) const override {
        ShadingArg arg;
        arg.texcoord = texcoord;
        s3dg_shade_arg a = make_arg(arg);
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        s3dg_context *ctx = DefaultContext();
        s3dg_texture_desc td{};
        if (texture) td = texture->Desc();
        if (!ctx || s3dg_base_color(ctx, &desc, texture ? &td : nullptr, 1, &a) != S3DG_OK) return Color3();
        return Color3(a.weight[0], a.weight[1], a.weight[2]);
    }
    bool HasDelta() const override { return desc.type == S3DG_MAT_MIRROR || desc.type == S3DG_MAT_GLASS; }

private:
    static s3dg_shade_arg make_arg(const ShadingArg &arg) {
        s3dg_shade_arg a{};
        a.input[0] = arg.input.x; a.input[1] = arg.input.y; a.input[2] = arg.input.z;
        a.normal[0] = arg.normal.x; a.normal[1] = arg.normal.y; a.normal[2] = arg.normal.z;
        a.texcoord[0] = arg.texcoord.x; a.texcoord[1] = arg.texcoord.y;
        return a;
    }
};

// MaterialFactory::Create* with the reference's signatures: a texture wraps
// the material in a TexturedMaterial (an invalid texture leaves it plain; the
// reference would sample an empty texture).
struct MaterialFactory {
    static IMaterial *Make(int type, const Color3 &diffuse, const Color3 &specular, f32 ior, f32 power,
                           const Texture *pTexture, const Color3 &emissive) {
        Material *m = new (std::nothrow) Material();
        if (!m) return nullptr;
        if (pTexture && pTexture->IsValid()) m->texture = pTexture;
        m->desc.type = type;
        m->desc.diffuse[0] = diffuse.x; m->desc.diffuse[1] = diffuse.y; m->desc.diffuse[2] = diffuse.z;
        m->desc.specular[0] = specular.x; m->desc.specular[1] = specular.y; m->desc.specular[2] = specular.z;
        m->desc.emissive[0] = emissive.x; m->desc.emissive[1] = emissive.y; m->desc.emissive[2] = emissive.z;
        m->desc.ior = ior;
        m->desc.power = power;
        return m;
    }
    static IMaterial *CreateLambert(const Color3 &diffuse, const Texture *pTexture = nullptr,
                                    const Color3 &emissive = Color3()) {
        return Make(S3DG_MAT_LAMBERT, diffuse, Color3(), 1.0f, 1.0f, pTexture, emissive);
    }
    static IMaterial *CreatePhong(const Color3 &specular, f32 power, const Texture *pTexture = nullptr,
                                  const Color3 &emissive = Color3()) {
        return Make(S3DG_MAT_PHONG, Color3(), specular, 1.0f, power, pTexture, emissive);
    }
    static IMaterial *CreateMirror(const Color3 &specular, const Texture *pTexture = nullptr,
                                   const Color3 &emissive = Color3()) {
        return Make(S3DG_MAT_MIRROR, Color3(), specular, 1.0f, 1.0f, pTexture, emissive);
    }
    static IMaterial *CreateGlass(const Color3 &specular, f32 ior, const Texture *pTexture = nullptr,
                                  const Color3 &emissive = Color3()) {
        return Make(S3DG_MAT_GLASS, Color3(), specular, ior, 1.0f, pTexture, emissive);
    }
    static IMaterial *CreatePlastic(const Color3 &diffuse, const Color3 &specular, f32 power,
                                    const Texture *pTexture = nullptr, const Color3 &emissive = Color3()) {
        return Make(S3DG_MAT_PLASTIC, diffuse, specular, 1.0f, power, pTexture, emissive);
    }
};

// ---------------------------------------------------------------------------
// Cameras (PinholeCamera / ThinLensCamera, s3d_camera.h:31-268)
// ---------------------------------------------------------------------------
namespace detail {
inline bool CameraRay(const s3dg_camera_desc &d, f32 x, f32 y, u64 *state, Ray &out) {
    const float xy[2] = {x, y};
    float od[6] = {0, 0, 0, 0, 0, 0};
    std::lock_guard<std::recursive_mutex> lock(Lock());
    s3dg_context *ctx = DefaultContext();
    if (!ctx || s3dg_camera_rays(ctx, &d, 1, xy, state, od) != S3DG_OK) return false;
    out = Ray(Vector3(od[0], od[1], od[2]), Vector3(od[3], od[4], od[5]));
    return true;
}
inline void FillCamera(s3dg_camera_desc &d, const Vector3 &position, const Vector3 &target, const Vector3 &upward,
                       f32 fov, f32 aspect, f32 nearClip) {
    d.position[0] = position.x; d.position[1] = position.y; d.position[2] = position.z;
    d.target[0] = target.x; d.target[1] = target.y; d.target[2] = target.z;
    d.upward[0] = upward.x; d.upward[1] = upward.y; d.upward[2] = upward.z;
    d.fov = fov;
    d.aspect = aspect;
    d.near_clip = nearClip;
}
}  // namespace detail

// A camera the library renders with: its ABI description.
struct DeviceCamera : ICamera {
    s3dg_camera_desc desc{};
};

class PinholeCamera : public DeviceCamera {
public:
    void Update(const Vector3 position, const Vector3 target, const Vector3 upward, const f32 fov,
                const f32 aspectRatio, const f32 nearClip) {
        desc = s3dg_camera_desc{};
        desc.type = S3DG_CAMERA_PINHOLE;
        detail::FillCamera(desc, position, target, upward, fov, aspectRatio, nearClip);
    }
    Ray GetRay(const f32 x, const f32 y) override {
        Ray r;
        detail::CameraRay(desc, x, y, nullptr, r);
        return r;
    }
};

// The lens sample of GetRay comes from the camera's own PCG stream (the
// reference keeps its own xorshift generator there), so GetRay is not thread
// safe, as in the reference.
class ThinLensCamera : public DeviceCamera {
public:
    void Update(const Vector3 position, const Vector3 target, const Vector3 upward, const f32 fov,
                const f32 aspectRatio, const f32 nearClip, const f32 lensRadius) {
        desc = s3dg_camera_desc{};
        desc.type = S3DG_CAMERA_THINLENS;
        detail::FillCamera(desc, position, target, upward, fov, aspectRatio, nearClip);
        desc.lens_radius = lensRadius;
    }
    Ray GetRay(const f32 x, const f32 y) override {
        Ray r;
        u64 s = m_Random.GetState();
        if (detail::CameraRay(desc, x, y, &s, r)) m_Random.SetState(s);
        return r;
    }
    PCG &GetRandom() { return m_Random; }

private:
    PCG m_Random;
};

// ---------------------------------------------------------------------------
// Shapes traced on their own: each keeps a device scene of itself.
// ---------------------------------------------------------------------------
namespace detail {

// The ABI description of an IMaterial for a device scene: a Material's own,
// or a placeholder Lambert for materials defined elsewhere (tracing reads no
// material).
inline s3dg_material MaterialDesc(const IMaterial *m) {
    const Material *x = dynamic_cast<const Material *>(m);
    if (x) {
        s3dg_material d = x->desc;
        d.texture = 0;
        return d;
    }
    s3dg_material d{};
    d.type = S3DG_MAT_LAMBERT;
    d.ior = 1.0f;
    d.power = 1.0f;
    return d;
}

// A device scene built on first use and destroyed with its owner.
class LocalScene {
public:
    LocalScene() = default;
    LocalScene(const LocalScene &) = delete;
    LocalScene &operator=(const LocalScene &) = delete;
    ~LocalScene() { Reset(); }
    void Reset() {
        std::lock_guard<std::recursive_mutex> lock(Lock());
        if (m_Scene) s3dg_scene_destroy(m_Scene);
        m_Scene = nullptr;
    }
    s3dg_scene *Get() const { return m_Scene; }
    // Creates the scene from `d` unless it exists already.
    bool Ensure(const s3dg_scene_desc &d) {
        std::lock_guard<std::recursive_mutex> lock(Lock());
        if (m_Scene) return true;
        s3dg_context *ctx = DefaultContext();
        return ctx && s3dg_scene_create(ctx, &d, &m_Scene) == S3DG_OK;
    }
    // One ray, closest hit in [tmin, tmax]: false on a miss or a device error.
    bool Trace(const Ray &r, f32 tmin, f32 tmax, s3dg_hit &h, int32_t &shape) const {
        std::lock_guard<std::recursive_mutex> lock(Lock());
        if (!m_Scene) return false;
        s3dg_ray ray{{r.pos.x, r.pos.y, r.pos.z}, tmin, {r.dir.x, r.dir.y, r.dir.z}, tmax};
        if (s3dg_trace(m_Scene, &ray, 1, S3DG_TRACE_CLOSEST, &h, &shape) != S3DG_OK) return false;
        return h.prim != -1;
    }
    bool CalcParam(s32 shape, s32 triangle, const Vector3 &pos, const Vector2 &bary, Vector3 *n, Vector2 *uv) const {
        std::lock_guard<std::recursive_mutex> lock(Lock());
        if (!m_Scene) return false;
        s3dg_param_query q{shape, triangle, {pos.x, pos.y, pos.z}, {bary.x, bary.y}};
        float nn[3], tt[2];
        if (s3dg_scene_calc_param(m_Scene, 1, &q, nn, tt) != S3DG_OK) return false;
        if (n) *n = Vector3(nn[0], nn[1], nn[2]);
        if (uv) *uv = Vector2(tt[0], tt[1]);
        return true;
    }

private:
    s3dg_scene *m_Scene = nullptr;
};

// Scene description of one mesh of `count` triangles (vertices 3 per
// triangle, material per triangle) as shape 0.
inline bool BuildMeshScene(LocalScene &ls, const std::vector<Vertex> &v, const std::vector<s3dg_material> &mats,
                           const std::vector<int32_t> &tri_mat, int bvh_kind, const float *world) {
    s3dg_mesh_desc md{};
    md.vertex_count = (uint32_t)v.size();
    md.vertices = reinterpret_cast<const s3dg_vertex *>(v.data());
    md.triangle_material = tri_mat.data();
    md.bvh_kind = bvh_kind;
    s3dg_shape_desc sd{};
    sd.type = world ? S3DG_SHAPE_INSTANCE : S3DG_SHAPE_MESH;
    sd.mesh = 0;
    sd.material = -1;
    if (world) memcpy(sd.world, world, sizeof(sd.world));
    s3dg_scene_desc d{};
    d.mesh_count = 1;
    d.meshes = &md;
    d.shape_count = 1;
    d.shapes = &sd;
    d.material_count = (uint32_t)mats.size();
    d.materials = mats.data();
    d.camera.type = S3DG_CAMERA_PINHOLE;
    d.camera.position[2] = 1.0f;
    d.camera.upward[1] = 1.0f;
    d.camera.fov = 1.0f;
    d.camera.aspect = 1.0f;
    d.camera.near_clip = 1.0f;
    return ls.Ensure(d);
}

// Upper end of a triangle's accepted distance: it rejects t >= the record's
// distance and t > F_HIT_MAX (Triangle::IsHit, s3d_triangle.cpp:95-99).
inline f32 TriangleTmax(f32 distance) {
    const f32 below = std::nextafter(distance, 0.0f);
    return below < F_HIT_MAX ? below : F_HIT_MAX;
}

}  // namespace detail

class Mesh;
class Instance;

// Triangle (s3d_triangle.h / .cpp).  A triangle of a Mesh is a view of it
// (what a Mesh hit reports as pShape); its CalcParam then uses the mesh's
// device scene.
class Triangle : public IShape {
public:
    // Triangle::Create(Vertex*, IMaterial*) (s3d_triangle.cpp:136-137)
    static IShape *Create(const Vertex *pVertices, IMaterial *pMaterial) {
        if (!pVertices || !pMaterial) return nullptr;
        Triangle *t = new (std::nothrow) Triangle(pVertices, pMaterial);
        return t;
    }
    bool IsHit(const RaySet &raySet, HitRecord &record) const override {
        if (!m_Local.Get()) {
            std::vector<Vertex> v(m_Vertex, m_Vertex + 3);
            if (!detail::BuildMeshScene(m_Local, v, {detail::MaterialDesc(m_pMaterial)}, {0}, S3DG_BVH8, nullptr))
                return false;
        }
        s3dg_hit h;
        int32_t shape = -1;
        if (!m_Local.Trace(raySet.ray, F_HIT_MIN, detail::TriangleTmax(record.distance), h, shape)) return false;
        record.distance = h.t;
        record.barycentric = Vector2(h.u, h.v);
        record.pShape = this;
        record.pMaterial = m_pMaterial;
        record.triangle = 0;
        return true;
    }
    BoundingBox GetBox() const override { return m_Box; }
    Vector3 GetCenter() const override { return m_Box.center; }
    void CalcParam(const Vector3 &pos, const Vector2 &barycentric, Vector3 *pNormal,
                   Vector2 *pTexCoord) const override;
    const Vertex *GetVertices() const { return m_Vertex; }
    IMaterial *GetMaterial() const { return m_pMaterial; }
    ~Triangle() override { SafeRelease(m_pMaterial); }

private:
    friend class Mesh;
    Triangle(const Vertex *v, IMaterial *m) : m_pMaterial(m) {
        for (int i = 0; i < 3; ++i) m_Vertex[i] = v[i];
        Vector3 lo = m_Vertex[0].Position, hi = m_Vertex[0].Position;
        lo = Vector3::Min(lo, m_Vertex[1].Position);
        lo = Vector3::Min(lo, m_Vertex[2].Position);
        hi = Vector3::Max(hi, m_Vertex[1].Position);
        hi = Vector3::Max(hi, m_Vertex[2].Position);
        m_Box = BoundingBox(lo, hi);
        if (m_pMaterial) m_pMaterial->AddRef();
    }
    Vertex m_Vertex[3];
    IMaterial *m_pMaterial = nullptr;
    BoundingBox m_Box;
    const Mesh *m_pOwner = nullptr;   // set for a view of a mesh triangle
    u32 m_Index = 0;
    mutable detail::LocalScene m_Local;
};

// Mesh::Create (s3d_mesh.h:42-47): a triangle list with one material, or an
// SMD file with per-triangle materials and textures.  IsHit reports the hit
// triangle (a view of it) as pShape, as the reference's Mesh, whose BVH
// holds Triangle objects, does.
class Mesh : public IShape {
public:
    std::vector<Vertex> vertices;
    IMaterial *material = nullptr;                 // one material for every triangle, or
    std::vector<IMaterial *> materials;            // per-triangle materials (SMD files)
    std::vector<int32_t> triangle_material;
    std::vector<Texture *> textures;               // owned, as Mesh::m_Textures
    int bvh_kind = S3DG_BVH8;

    ~Mesh() override {
        for (Triangle *t : m_Views)
            if (t) t->Release();
        SafeRelease(material);
        for (IMaterial *m : materials) m->Release();
        for (Texture *t : textures) delete t;
    }
    // Mesh::Create(const char* filename) (s3d_mesh.cpp:250-465): an SMD file
    // with its materials and BMP textures.  A texture that fails to load leaves
    // its material untextured (the reference logs a warning); a triangle
    // without a material makes the mesh fail to build.
    static IShape *Create(const char *filename) {
        s3dg_smd *smd = nullptr;
        if (s3dg_smd_load(filename, &smd) != S3DG_OK) return nullptr;
        Mesh *m = new (std::nothrow) Mesh();
        if (!m) {
            s3dg_smd_free(smd);
            return nullptr;
        }
        for (uint32_t i = 0; i < smd->texture_count; ++i) {
            Texture *t = new Texture();
            if (!t->Init(smd->texture_paths[i], TFF_BMP))
                fprintf(stderr, "Warning : Texture Load Failed. filename = %s\n", smd->texture_paths[i]);
            m->textures.push_back(t);
        }
        for (uint32_t i = 0; i < smd->material_count; ++i) {
            const s3dg_material &d = smd->materials[i];
            Material *x = new (std::nothrow) Material();
            if (!x) break;
            x->desc = d;
            x->desc.texture = 0;
            if (d.texture > 0 && m->textures[d.texture - 1]->IsValid()) x->texture = m->textures[d.texture - 1];
            m->materials.push_back(x);
        }
        m->vertices.resize(smd->vertex_count);
        for (uint32_t i = 0; i < smd->vertex_count; ++i) {
            const s3dg_vertex &v = smd->vertices[i];
            m->vertices[i] = Vertex{Vector3(v.position[0], v.position[1], v.position[2]),
                                    Vector3(v.normal[0], v.normal[1], v.normal[2]),
                                    Vector2(v.texcoord[0], v.texcoord[1])};
        }
        m->triangle_material.assign(smd->triangle_material, smd->triangle_material + smd->vertex_count / 3);
        const bool ok = m->materials.size() == smd->material_count;
        s3dg_smd_free(smd);
        if (!ok) {
            m->Release();
            return nullptr;
        }
        m->finish();
        return m;
    }
    // Mesh::Create(u32 vertexCount, Vertex*, IMaterial*) (s3d_mesh.cpp:536-551)
    static IShape *Create(u32 vertexCount, const Vertex *pVertices, IMaterial *pMaterial) {
        if (vertexCount == 0 || vertexCount % 3 != 0 || !pVertices || !pMaterial) return nullptr;
        Mesh *m = new (std::nothrow) Mesh();
        if (!m) return nullptr;
        m->vertices.assign(pVertices, pVertices + vertexCount);
        m->material = pMaterial;
        pMaterial->AddRef();
        m->finish();
        return m;
    }

    bool IsHit(const RaySet &raySet, HitRecord &record) const override {
        if (!Device()) return false;
        s3dg_hit h;
        int32_t shape = -1;
        if (!m_Local.Trace(raySet.ray, F_HIT_MIN, detail::TriangleTmax(record.distance), h, shape)) return false;
        const Triangle *t = View((u32)h.prim);
        record.distance = h.t;
        record.barycentric = Vector2(h.u, h.v);
        record.pShape = t;
        record.pMaterial = TriangleMaterial((u32)h.prim);
        record.triangle = h.prim;
        return true;
    }
    BoundingBox GetBox() const override { return m_Box; }
    Vector3 GetCenter() const override { return m_Box.center; }
    u32 GetTriangleCount() const { return (u32)(vertices.size() / 3); }
    IMaterial *TriangleMaterial(u32 i) const {
        if (material) return material;
        const int32_t k = i < triangle_material.size() ? triangle_material[i] : -1;
        return k >= 0 && (size_t)k < materials.size() ? materials[k] : nullptr;
    }
    // The view of triangle i (created on first use, owned by the mesh).
    const Triangle *View(u32 i) const {
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        if (m_Views.size() != GetTriangleCount()) m_Views.assign(GetTriangleCount(), nullptr);
        if (!m_Views[i]) {
            Triangle *t = new (std::nothrow) Triangle(&vertices[3 * (size_t)i], TriangleMaterial(i));
            if (!t) return nullptr;
            t->m_pOwner = this;
            t->m_Index = i;
            m_Views[i] = t;
        }
        return m_Views[i];
    }
    // The mesh's own device scene (shape 0), built on first use.
    s3dg_scene *Device() const {
        if (m_Local.Get()) return m_Local.Get();
        std::vector<s3dg_material> mats;
        std::vector<int32_t> tri(GetTriangleCount(), 0);
        if (material) {
            mats.push_back(detail::MaterialDesc(material));
        } else {
            for (IMaterial *m : materials) mats.push_back(detail::MaterialDesc(m));
            for (size_t i = 0; i < tri.size(); ++i) {
                const int32_t k = i < triangle_material.size() ? triangle_material[i] : -1;
                if (k < 0 || (size_t)k >= mats.size()) return nullptr;
                tri[i] = k;
            }
        }
        detail::BuildMeshScene(m_Local, vertices, mats, tri, bvh_kind, nullptr);
        return m_Local.Get();
    }
    bool CalcParamOf(u32 i, const Vector3 &pos, const Vector2 &bary, Vector3 *n, Vector2 *uv) const {
        return Device() && m_Local.CalcParam(0, (s32)i, pos, bary, n, uv);
    }

private:
    void finish() {
        BoundingBox b;
        for (const Vertex &v : vertices) b = BoundingBox::Merge(b, BoundingBox(v.Position));
        m_Box = b;
    }
    BoundingBox m_Box;
    mutable std::vector<Triangle *> m_Views;
    mutable detail::LocalScene m_Local;
};

// Triangle::CalcParam (s3d_triangle.cpp:107-121): the barycentric blend of the
// vertex normals (SafeNormalize) and texture coordinates, on the device.
inline void Triangle::CalcParam(const Vector3 &pos, const Vector2 &barycentric, Vector3 *pNormal,
                                Vector2 *pTexCoord) const {
    if (m_pOwner) {
        m_pOwner->CalcParamOf(m_Index, pos, barycentric, pNormal, pTexCoord);
        return;
    }
    if (!m_Local.Get()) {
        std::vector<Vertex> v(m_Vertex, m_Vertex + 3);
        if (!detail::BuildMeshScene(m_Local, v, {detail::MaterialDesc(m_pMaterial)}, {0}, S3DG_BVH8, nullptr))
            return;
    }
    m_Local.CalcParam(0, 0, pos, barycentric, pNormal, pTexCoord);
}

// Leaf::Create (s3d_leaf.cpp:22-35, 74-87): its shapes tested in stored order.
class Leaf : public IShape {
public:
    static IShape *Create(size_t count, IShape **ppShapes) {
        if (count == 0 || !ppShapes) return nullptr;
        Leaf *l = new (std::nothrow) Leaf();
        if (!l) return nullptr;
        for (size_t i = 0; i < count; ++i) {
            ppShapes[i]->AddRef();
            l->m_pShapes.push_back(ppShapes[i]);
        }
        l->m_Box = ppShapes[0]->GetBox();
        for (size_t i = 1; i < count; ++i) l->m_Box = BoundingBox::Merge(l->m_Box, ppShapes[i]->GetBox());
        return l;
    }
    bool IsHit(const RaySet &raySet, HitRecord &record) const override {
        bool hit = false;
        for (IShape *s : m_pShapes) hit |= s->IsHit(raySet, record);
        return hit;
    }
    BoundingBox GetBox() const override { return m_Box; }
    Vector3 GetCenter() const override { return m_Box.center; }
    ~Leaf() override {
        for (IShape *s : m_pShapes) s->Release();
    }

private:
    std::vector<IShape *> m_pShapes;
    BoundingBox m_Box;
};

// The tree BVH8/BVH4/BVH2::Create build over Triangle objects.  The array is
// reordered in place into the reference build's depth-first order; IsHit
// returns the reference traversal's hit (the smallest distance, ties to the
// first triangle in that order) as one trace of a device scene holding the
// triangles, and reports the Triangle object as pShape.
class BVHTriangles : public IShape {
public:
    static IShape *Create(size_t count, IShape **ppShapes, int kind) {
        if (count == 0 || !ppShapes || count > 0x7fffffff / 3) return nullptr;
        std::vector<Triangle *> tris(count);
        std::vector<Vertex> v(3 * count);
        for (size_t i = 0; i < count; ++i) {
            tris[i] = dynamic_cast<Triangle *>(ppShapes[i]);
            if (!tris[i]) return nullptr;   // only triangles are supported
            for (int k = 0; k < 3; ++k) v[3 * i + k] = tris[i]->GetVertices()[k];
        }
        std::vector<int32_t> order(count);
        uint32_t nodes = 0;
        if (s3dg_build_order((uint32_t)(3 * count), reinterpret_cast<const s3dg_vertex *>(v.data()), kind,
                             order.data(), &nodes) != S3DG_OK)
            return nullptr;
        BVHTriangles *b = new (std::nothrow) BVHTriangles();
        if (!b) return nullptr;
        b->m_Kind = kind;
        // the device scene keeps the input order: its tie order is then the
        // reference build's depth-first order over that input
        for (size_t i = 0; i < count; ++i) {
            tris[i]->AddRef();
            b->m_Tris.push_back(tris[i]);
            b->m_Box = BoundingBox::Merge(b->m_Box, tris[i]->GetBox());
        }
        for (size_t i = 0; i < count; ++i) ppShapes[i] = tris[order[i]];
        return b;
    }
    bool IsHit(const RaySet &raySet, HitRecord &record) const override {
        if (!Device()) return false;
        s3dg_hit h;
        int32_t shape = -1;
        if (!m_Local.Trace(raySet.ray, F_HIT_MIN, detail::TriangleTmax(record.distance), h, shape)) return false;
        const Triangle *t = m_Tris[(size_t)h.prim];
        record.distance = h.t;
        record.barycentric = Vector2(h.u, h.v);
        record.pShape = t;
        record.pMaterial = t->GetMaterial();
        record.triangle = h.prim;
        return true;
    }
    BoundingBox GetBox() const override { return m_Box; }
    Vector3 GetCenter() const override { return m_Box.center; }
    ~BVHTriangles() override {
        for (Triangle *t : m_Tris) t->Release();
    }

private:
    s3dg_scene *Device() const {
        if (m_Local.Get()) return m_Local.Get();
        std::vector<Vertex> v(3 * m_Tris.size());
        std::vector<s3dg_material> mats;
        std::vector<int32_t> tri(m_Tris.size());
        std::vector<const IMaterial *> seen;
        for (size_t i = 0; i < m_Tris.size(); ++i) {
            for (int k = 0; k < 3; ++k) v[3 * i + k] = m_Tris[i]->GetVertices()[k];
            const IMaterial *m = m_Tris[i]->GetMaterial();
            size_t j = 0;
            while (j < seen.size() && seen[j] != m) ++j;
            if (j == seen.size()) {
                seen.push_back(m);
                mats.push_back(detail::MaterialDesc(m));
            }
            tri[i] = (int32_t)j;
        }
        detail::BuildMeshScene(m_Local, v, mats, tri, m_Kind, nullptr);
        return m_Local.Get();
    }
    std::vector<Triangle *> m_Tris;   // in the caller's input order
    BoundingBox m_Box;
    int m_Kind = S3DG_BVH8;
    mutable detail::LocalScene m_Local;
};

// BVH8::Create / BVH4::Create / BVH2::Create (s3d_bvh8.cpp:401-479,
// s3d_bvh4.cpp:370-423, s3d_bvh2.cpp:174-274) over Triangle objects.
struct BVH8 {
    static IShape *Create(size_t count, IShape **ppShapes) { return BVHTriangles::Create(count, ppShapes, S3DG_BVH8); }
};
struct BVH4 {
    static IShape *Create(size_t count, IShape **ppShapes) { return BVHTriangles::Create(count, ppShapes, S3DG_BVH4); }
};
struct BVH2 {
    static IShape *Create(size_t count, IShape **ppShapes) { return BVHTriangles::Create(count, ppShapes, S3DG_BVH2); }
};

// Instance (s3d_instance.cpp): a mesh under a world matrix; the ray moves to
// the mesh's space and the hit reports the mesh's triangle as pShape (the
// reference hands back the inner triangle too).
class Instance : public IShape {
public:
    IShape *shape = nullptr;
    Matrix world;
    IMaterial *override_material = nullptr;   // extension: replaces the triangles' materials
    u32 version = 0;   // bumped by UpdateMatrix; a built Scene catches up on its next use

    // Instance::Create(IShape*, const Matrix&) (s3d_instance.cpp:105-106)
    static IShape *Create(IShape *pShape, const Matrix &world) {
        if (!pShape || !dynamic_cast<Mesh *>(pShape)) return nullptr;
        Instance *i = new (std::nothrow) Instance();
        if (!i) return nullptr;
        i->shape = pShape;
        pShape->AddRef();
        i->world = world;
        i->m_Box = BoundingBox::Transform(pShape->GetBox(), world);
        // the constructor's centre is (maxi + maxi) * 0.5 (s3d_instance.cpp:31)
        i->m_Center = (i->m_Box.maxi + i->m_Box.maxi) * 0.5f;
        return i;
    }
    // Instance::UpdateMatrix (s3d_instance.cpp:95-101)
    void UpdateMatrix(const Matrix &m) {
        world = m;
        m_Box = BoundingBox::Transform(shape->GetBox(), m);
        m_Center = m.Transform(shape->GetCenter());
        ++version;
        m_Local.Reset();
    }
    bool IsHit(const RaySet &raySet, HitRecord &record) const override {
        const Mesh *mesh = static_cast<const Mesh *>(shape);
        if (!m_Local.Get()) {
            std::vector<s3dg_material> mats;
            std::vector<int32_t> tri(mesh->GetTriangleCount());
            for (u32 i = 0; i < mesh->GetTriangleCount(); ++i) {
                mats.push_back(detail::MaterialDesc(mesh->TriangleMaterial(i)));
                tri[i] = (int32_t)i;
            }
            float w[16];
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) w[4 * r + c] = world.m[r][c];
            if (!detail::BuildMeshScene(m_Local, mesh->vertices, mats, tri, mesh->bvh_kind, w)) return false;
        }
        s3dg_hit h;
        int32_t s = -1;
        if (!m_Local.Trace(raySet.ray, F_HIT_MIN, detail::TriangleTmax(record.distance), h, s)) return false;
        record.distance = h.t;
        record.barycentric = Vector2(h.u, h.v);
        record.pShape = mesh->View((u32)h.prim);
        record.pMaterial = override_material ? override_material : mesh->TriangleMaterial((u32)h.prim);
        record.triangle = h.prim;
        return true;
    }
    BoundingBox GetBox() const override { return m_Box; }
    Vector3 GetCenter() const override { return m_Center; }
    ~Instance() override {
        SafeRelease(shape);
        SafeRelease(override_material);
    }

private:
    BoundingBox m_Box;
    Vector3 m_Center;
    mutable detail::LocalScene m_Local;
};

// Sphere (s3d_sphere.cpp): IsHit keeps a hit at exactly the record's
// distance (a later sphere wins a tie), Sample draws a light point.
class Sphere : public IShape {
public:
    f32 radius = 0.0f;
    Vector3 center;
    IMaterial *material = nullptr;
    // Sphere::Create(f32, const Vector3&, IMaterial*) (s3d_sphere.cpp:120-121)
    static IShape *Create(f32 radius, const Vector3 &center, IMaterial *pMaterial) {
        if (!(radius > 0.0f) || !pMaterial) return nullptr;
        Sphere *s = new (std::nothrow) Sphere();
        if (!s) return nullptr;
        s->radius = radius;
        s->center = center;
        s->material = pMaterial;
        pMaterial->AddRef();
        return s;
    }
    bool IsHit(const RaySet &raySet, HitRecord &record) const override {
        if (!Device()) return false;
        s3dg_hit h;
        int32_t s = -1;
        if (!m_Local.Trace(raySet.ray, F_HIT_MIN, record.distance, h, s)) return false;
        record.distance = h.t;
        record.barycentric = Vector2();
        record.pShape = this;
        record.pMaterial = material;
        record.triangle = -2;
        return true;
    }
    BoundingBox GetBox() const override {
        return BoundingBox(Vector3(center.x - radius, center.y - radius, center.z - radius),
                           Vector3(center.x + radius, center.y + radius, center.z + radius));
    }
    Vector3 GetCenter() const override { return center; }
    void CalcParam(const Vector3 &pos, const Vector2 &barycentric, Vector3 *pNormal,
                   Vector2 *pTexCoord) const override {
        if (Device()) m_Local.CalcParam(0, -2, pos, barycentric, pNormal, pTexCoord);
    }
    void Sample(PCG &random, Vector3 *pPosition, float *pdf) override {
        const float c[3] = {center.x, center.y, center.z};
        u64 s = random.GetState();
        float p[3], q = 0.0f;
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        s3dg_context *ctx = DefaultContext();
        if (!ctx || s3dg_sphere_sample(ctx, c, radius, 1, &s, p, &q) != S3DG_OK) return;
        random.SetState(s);
        if (pPosition) *pPosition = Vector3(p[0], p[1], p[2]);
        if (pdf) *pdf = q;
    }
    ~Sphere() override { SafeRelease(material); }

private:
    s3dg_scene *Device() const {
        if (m_Local.Get()) return m_Local.Get();
        s3dg_material md = detail::MaterialDesc(material);
        s3dg_shape_desc sd{};
        sd.type = S3DG_SHAPE_SPHERE;
        sd.mesh = -1;
        sd.material = 0;
        sd.center[0] = center.x; sd.center[1] = center.y; sd.center[2] = center.z;
        sd.radius = radius;
        s3dg_scene_desc d{};
        d.shape_count = 1;
        d.shapes = &sd;
        d.material_count = 1;
        d.materials = &md;
        d.camera.type = S3DG_CAMERA_PINHOLE;
        d.camera.position[2] = 1.0f;
        d.camera.upward[1] = 1.0f;
        d.camera.fov = 1.0f;
        d.camera.aspect = 1.0f;
        d.camera.near_clip = 1.0f;
        m_Local.Ensure(d);
        return m_Local.Get();
    }
    mutable detail::LocalScene m_Local;
};

// Legacy name of the camera a Scene renders with by default.
typedef PinholeCamera Camera;

// ---------------------------------------------------------------------------
// Scene: shape list intersected in order, light list, camera, environment.
// Build() uploads it; Intersect() is the single-ray path (one s3dg_trace).
// ---------------------------------------------------------------------------
class Scene {
public:
    Scene() : m_pCamera(&m_DefaultCamera) {}
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    virtual ~Scene() { Reset(); }

    void AddShape(IShape *s) {
        s->AddRef();
        m_Shapes.push_back(s);
        Invalidate();
    }
    // A sphere already added as a shape becomes a light (Scene::m_pLightList).
    void AddLight(IShape *s) {
        s->AddRef();
        m_pLightList.push_back(s);
        Invalidate();
    }
    // Quad light made of `count` triangles of a Mesh shape (an extension).
    void AddQuadLight(IShape *mesh, s32 firstTriangle, s32 count, const Vector3 &corner, const Vector3 &edge1,
                      const Vector3 &edge2) {
        QuadLight q{mesh, firstTriangle, count, corner, edge1, edge2};
        m_Quads.push_back(q);
        Invalidate();
    }
    // The camera the scene renders with (Scene::m_pCamera): a PinholeCamera or
    // a ThinLensCamera of this header, owned by the caller.
    void SetCamera(DeviceCamera *camera) {
        m_pCamera = camera ? camera : &m_DefaultCamera;
        Invalidate();
    }
    PinholeCamera &GetCamera() {
        m_pCamera = &m_DefaultCamera;
        Invalidate();
        return m_DefaultCamera;
    }
    void SetEnvironment(const Color3 &c) { m_Env = c; Invalidate(); }
    // m_IBL: a lat-long texture whose SampleColor(dir) * 10 a missed ray sees
    // (Scene::SampleIBL); nullptr returns to the constant environment.
    void SetIBL(const Texture *t) { m_IBL = (t && t->IsValid()) ? t : nullptr; Invalidate(); }

    // Scene::GetRay (s3d_scene.h:56-58)
    Ray GetRay(f32 x, f32 y) { return m_pCamera->GetRay(x, y); }

    // Scene::Intersect (s3d_scene.h:63-75) with the reference's
    // [F_HIT_MIN, F_HIT_MAX] and the record's distance as the far end: one
    // s3dg_trace of one ray (one host round trip).
    bool Intersect(const RaySet &raySet, HitRecord &record) {
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        if (!Build()) return false;
        const s3dg_ray r = TraceRay(raySet, record);
        s3dg_hit h;
        int32_t shape = -1;
        if (s3dg_trace(m_Scene, &r, 1, S3DG_TRACE_CLOSEST, &h, &shape) != S3DG_OK || h.prim == -1) return false;
        Fill(h, shape, record);
        return true;
    }
    bool Intersect(const Ray &ray, HitRecord &record) { return Intersect(MakeRaySet(ray.pos, ray.dir), record); }

    // Many rays in one s3dg_trace (no counterpart in the reference, whose
    // Intersect takes one ray; for callers with batches, which would
    // otherwise pay a host round trip per ray): records[i] ends as
    // Intersect(raySets[i], records[i]) would leave it.  Returns the number
    // of hits, or 0 with the records untouched when the scene cannot be
    // built or traced.
    size_t Intersect(const RaySet *raySets, HitRecord *records, size_t count) {
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        if (!count || !Build() || count > 0xffffffffu) return 0;
        std::vector<s3dg_ray> r(count);
        std::vector<s3dg_hit> h(count);
        std::vector<int32_t> shape(count, -1);
        for (size_t i = 0; i < count; ++i) r[i] = TraceRay(raySets[i], records[i]);
        if (s3dg_trace(m_Scene, r.data(), (uint32_t)count, S3DG_TRACE_CLOSEST, h.data(), shape.data()) != S3DG_OK)
            return 0;
        size_t hits = 0;
        for (size_t i = 0; i < count; ++i) {
            if (h[i].prim == -1) continue;
            Fill(h[i], shape[i], records[i]);
            ++hits;
        }
        return hits;
    }

private:
    static s3dg_ray TraceRay(const RaySet &raySet, const HitRecord &record) {
        const Ray &ray = raySet.ray;
        return s3dg_ray{{ray.pos.x, ray.pos.y, ray.pos.z}, F_HIT_MIN, {ray.dir.x, ray.dir.y, ray.dir.z},
                        detail::TriangleTmax(record.distance)};
    }
    // The record of a hit: distance, barycentrics, the hit shape (the
    // triangle's view for meshes) and its material.
    void Fill(const s3dg_hit &h, int32_t shape, HitRecord &record) const {
        record.distance = h.t;
        record.barycentric = Vector2(h.u, h.v);
        record.shape = shape;
        record.triangle = h.prim;
        IShape *s = m_Shapes[(size_t)shape];
        if (Sphere *sp = dynamic_cast<Sphere *>(s)) {
            record.pShape = sp;
            record.pMaterial = sp->material;
        } else if (Instance *in = dynamic_cast<Instance *>(s)) {
            const Mesh *m = static_cast<const Mesh *>(in->shape);
            record.pShape = m->View((u32)h.prim);
            record.pMaterial = in->override_material ? in->override_material : m->TriangleMaterial((u32)h.prim);
        } else {
            const Mesh *m = static_cast<const Mesh *>(s);
            record.pShape = m->View((u32)h.prim);
            record.pMaterial = m->TriangleMaterial((u32)h.prim);
        }
    }

public:

    // Scene::SampleIBL (s3d_scene.h:80-82)
    Color3 SampleIBL(const Vector3 &dir) {
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        const float d[3] = {dir.x, dir.y, dir.z};
        float c[3] = {0, 0, 0};
        if (!Build() || s3dg_scene_environment(m_Scene, 1, d, c) != S3DG_OK) return Color3();
        return Color3(c[0], c[1], c[2]);
    }
    // Scene::GetLight (s3d_scene.h:84-90): a light picked by u32 % count
    // (nullptr with no lights, where the reference divides by zero).
    IShape *GetLight(PCG &rand) {
        if (m_pLightList.empty()) return nullptr;
        const u32 idx = rand.GetAsU32();
        return m_pLightList[idx % m_pLightList.size()];
    }

    // Uploads the scene (builds every mesh tree with the reference SAH on the host).
    bool Build() {
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        if (m_Scene) return SyncInstances();
        s3dg_context *ctx = DefaultContext();
        if (!ctx) return false;
        std::vector<IMaterial *> mats;
        bool foreign = false;   // a material this layer cannot upload
        auto mat_index = [&](IMaterial *m) -> int32_t {
            if (!dynamic_cast<Material *>(m)) foreign = true;
            for (size_t i = 0; i < mats.size(); ++i)
                if (mats[i] == m) return (int32_t)i;
            mats.push_back(m);
            return (int32_t)mats.size() - 1;
        };
        std::vector<Mesh *> meshes;
        auto mesh_index = [&](Mesh *m) -> int32_t {
            for (size_t i = 0; i < meshes.size(); ++i)
                if (meshes[i] == m) return (int32_t)i;
            meshes.push_back(m);
            return (int32_t)meshes.size() - 1;
        };
        std::vector<s3dg_shape_desc> shapes(m_Shapes.size());
        for (size_t i = 0; i < m_Shapes.size(); ++i) {
            s3dg_shape_desc &d = shapes[i];
            d = s3dg_shape_desc{};
            d.material = -1;
            IShape *s = m_Shapes[i];
            if (Mesh *mesh = dynamic_cast<Mesh *>(s)) {
                d.type = S3DG_SHAPE_MESH;
                d.mesh = mesh_index(mesh);
            } else if (Instance *in = dynamic_cast<Instance *>(s)) {
                d.type = S3DG_SHAPE_INSTANCE;
                d.mesh = mesh_index(static_cast<Mesh *>(in->shape));
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) d.world[4 * r + c] = in->world.m[r][c];
                if (in->override_material) d.material = mat_index(in->override_material);
            } else if (Sphere *sp = dynamic_cast<Sphere *>(s)) {
                d.type = S3DG_SHAPE_SPHERE;
                d.material = mat_index(sp->material);
                d.center[0] = sp->center.x; d.center[1] = sp->center.y; d.center[2] = sp->center.z;
                d.radius = sp->radius;
            } else {
                return false;   // a scene holds meshes, instances and spheres
            }
        }
        std::vector<s3dg_mesh_desc> md(meshes.size());
        std::vector<std::vector<int32_t>> tri_mats(meshes.size());
        for (size_t i = 0; i < meshes.size(); ++i) {
            md[i] = s3dg_mesh_desc{};
            md[i].vertex_count = (uint32_t)meshes[i]->vertices.size();
            md[i].vertices = reinterpret_cast<const s3dg_vertex *>(meshes[i]->vertices.data());
            md[i].bvh_kind = meshes[i]->bvh_kind;
            if (meshes[i]->material) {
                md[i].material = mat_index(meshes[i]->material);
            } else {
                // per-triangle materials; a triangle without one is rejected
                for (int32_t t : meshes[i]->triangle_material)
                    tri_mats[i].push_back(t >= 0 && (size_t)t < meshes[i]->materials.size()
                                              ? mat_index(meshes[i]->materials[t]) : -1);
                md[i].triangle_material = tri_mats[i].data();
            }
        }
        if (foreign) return false;
        std::vector<s3dg_light_desc> lights;
        for (IShape *l : m_pLightList) {
            s3dg_light_desc ld{};
            ld.type = S3DG_LIGHT_SHAPE;
            ld.shape = shape_index(l);
            lights.push_back(ld);
        }
        for (const QuadLight &q : m_Quads) {
            s3dg_light_desc ld{};
            ld.type = S3DG_LIGHT_QUAD;
            ld.shape = shape_index(q.mesh);
            ld.first_triangle = q.first;
            ld.triangle_count = q.count;
            ld.corner[0] = q.corner.x; ld.corner[1] = q.corner.y; ld.corner[2] = q.corner.z;
            ld.edge1[0] = q.e1.x; ld.edge1[1] = q.e1.y; ld.edge1[2] = q.e1.z;
            ld.edge2[0] = q.e2.x; ld.edge2[1] = q.e2.y; ld.edge2[2] = q.e2.z;
            lights.push_back(ld);
        }
        std::vector<const Texture *> texs;
        auto tex_slot = [&](const Texture *t) -> int32_t {
            if (!t) return 0;
            for (size_t i = 0; i < texs.size(); ++i)
                if (texs[i] == t) return (int32_t)i + 1;
            texs.push_back(t);
            return (int32_t)texs.size();
        };
        std::vector<s3dg_material> mdesc;
        for (IMaterial *m : mats) {
            const Material *x = static_cast<const Material *>(m);
            mdesc.push_back(x->desc);
            mdesc.back().texture = tex_slot(x->texture);
        }
        const int32_t ibl = tex_slot(m_IBL);
        std::vector<s3dg_texture_desc> tdesc;
        for (const Texture *t : texs) tdesc.push_back(t->Desc());
        s3dg_scene_desc sd{};
        sd.mesh_count = (uint32_t)md.size();
        sd.meshes = md.data();
        sd.shape_count = (uint32_t)shapes.size();
        sd.shapes = shapes.data();
        sd.material_count = (uint32_t)mdesc.size();
        sd.materials = mdesc.data();
        sd.light_count = (uint32_t)lights.size();
        sd.lights = lights.data();
        sd.camera = m_pCamera->desc;
        sd.environment[0] = m_Env.x; sd.environment[1] = m_Env.y; sd.environment[2] = m_Env.z;
        sd.texture_count = (uint32_t)tdesc.size();
        sd.textures = tdesc.data();
        sd.environment_texture = ibl;
        m_Versions.assign(m_Shapes.size(), 0);
        for (size_t i = 0; i < m_Shapes.size(); ++i)
            if (Instance *in = dynamic_cast<Instance *>(m_Shapes[i])) m_Versions[i] = in->version;
        return s3dg_scene_create(ctx, &sd, &m_Scene) == S3DG_OK;
    }

    s3dg_context *GetContext() { return DefaultContext(); }
    s3dg_scene *GetDeviceScene() { return Build() ? m_Scene : nullptr; }

    // Virtual like Scene::Update (s3d_scene.h:92): a subclass moves its
    // instances here with Instance::UpdateMatrix.
    virtual void Update(float) {}

protected:
    // Scene's protected members (s3d_scene.h:97-103) that a subclass fills
    DeviceCamera *m_pCamera;
    std::vector<IShape *> m_pLightList, m_Shapes;

private:
    struct QuadLight {
        IShape *mesh;
        s32 first, count;
        Vector3 corner, e1, e2;
    };
    std::vector<QuadLight> m_Quads;
    PinholeCamera m_DefaultCamera;
    Color3 m_Env;
    const Texture *m_IBL = nullptr;
    s3dg_scene *m_Scene = nullptr;

    std::vector<u32> m_Versions;   // instance versions the device scene holds

    // Sends the matrices of instances moved since the last use (one top-level
    // rebuild for all of them).
    bool SyncInstances() {
        std::vector<uint32_t> ids;
        std::vector<float> worlds;
        for (size_t i = 0; i < m_Shapes.size(); ++i) {
            const Instance *in = dynamic_cast<const Instance *>(m_Shapes[i]);
            if (!in || in->version == m_Versions[i]) continue;
            ids.push_back((uint32_t)i);
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) worlds.push_back(in->world.m[r][c]);
        }
        if (ids.empty()) return true;
        if (s3dg_scene_update_instances(m_Scene, (uint32_t)ids.size(), ids.data(), worlds.data()) != S3DG_OK)
            return false;
        for (uint32_t i : ids) m_Versions[i] = static_cast<const Instance *>(m_Shapes[i])->version;
        return true;
    }

    int32_t shape_index(IShape *s) const {
        for (size_t i = 0; i < m_Shapes.size(); ++i)
            if (m_Shapes[i] == s) return (int32_t)i;
        return -1;
    }
    void Invalidate() {
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        if (m_Scene) {
            s3dg_scene_destroy(m_Scene);
            m_Scene = nullptr;
        }
    }
    void Reset() {
        Invalidate();
        for (IShape *s : m_Shapes) s->Release();
        for (IShape *s : m_pLightList) s->Release();
        m_Shapes.clear();
        m_pLightList.clear();
        m_Quads.clear();
    }
};

// ---------------------------------------------------------------------------
// Tonemapping and NLM, with the reference's signatures
// ---------------------------------------------------------------------------
enum TONE_MAPPING_TYPE {
    TONE_MAPPING_REINHARD = S3DG_TONEMAP_REINHARD,
    TONE_MAPPING_UNCHARTED2_FILMIC = S3DG_TONEMAP_UNCHARTED2,
    TONE_MAPPING_ACES_FILMIC = S3DG_TONEMAP_ACES,
};

struct ToneMapper {
    // ToneMapper::Map(type, w, h, const Color3*, Color3*) (s3d_tonemapper.h:33)
    static void Map(TONE_MAPPING_TYPE type, const s32 width, const s32 height, const Color3 *pPixels,
                    Color3 *pResult) {
        Map(type, width, height, pPixels, pResult, nullptr);
    }
    // extension: also the 8-bit sRGB image of PathTracer::Capture; false on a device error
    static bool Map(TONE_MAPPING_TYPE type, s32 width, s32 height, const Color3 *pPixels, Color3 *pResult,
                    uint8_t *pRgb8) {
        std::lock_guard<std::recursive_mutex> lock(detail::Lock());
        s3dg_context *ctx = DefaultContext();
        return ctx && s3dg_tonemap(ctx, (int32_t)type, width, height, reinterpret_cast<const float *>(pPixels),
                                   reinterpret_cast<float *>(pResult), pRgb8) == S3DG_OK;
    }
};

// FilterNLM(w, h, coeff, src, dst) (s3d_denoiser.h:16): search radius 6,
// patch radius 2, as the reference.
inline void FilterNLM(int width, int height, float coeff, const Color4 *src, Color4 *dst) {
    std::lock_guard<std::recursive_mutex> lock(detail::Lock());
    s3dg_context *ctx = DefaultContext();
    if (ctx)
        s3dg_nlm(ctx, width, height, coeff, 6, 2, reinterpret_cast<const float *>(src),
                 reinterpret_cast<float *>(dst));
}
// extension: other radii; false on a device error
inline bool FilterNLM(int width, int height, float coeff, const Color4 *src, Color4 *dst, int searchRadius,
                      int patchRadius) {
    std::lock_guard<std::recursive_mutex> lock(detail::Lock());
    s3dg_context *ctx = DefaultContext();
    return ctx && s3dg_nlm(ctx, width, height, coeff, searchRadius, patchRadius,
                           reinterpret_cast<const float *>(src), reinterpret_cast<float *>(dst)) == S3DG_OK;
}

// ---------------------------------------------------------------------------
// PathTracer: progressive, time-bounded rendering like PathTracer::Run /
// TracePath (src/src/s3d_pt.cpp:77-132, 297-353): passes of one sample per
// pixel are added to a sum buffer until MaxRenderingSec has elapsed (checked
// between device launches of `PassesPerLaunch` passes) or MaxPasses is reached.
// ---------------------------------------------------------------------------
class PathTracer {
public:
    struct Config {
        s32 Width = 960;
        s32 Height = 540;
        s32 MaxBounceCount = 32;
        f32 MaxRenderingSec = 59.8f;
        s32 CpuCoreCount = 0;          // kept for source compatibility; unused
        s32 MaxPasses = 1 << 30;       // extension: stop after this many passes
        s32 PassesPerLaunch = 4;       // extension: passes per device call
        u64 Seed = 0;                  // extension: RNG key
        s32 RussianRouletteStart = 0;  // extension: 0 = every depth (reference)
    };

    bool Run(const Config &config, Scene &scene) {
        s3dg_scene *ds = scene.GetDeviceScene();
        if (!ds) return false;
        m_Config = config;
        m_Sum.assign((size_t)config.Width * config.Height, Color3());
        m_Passes = 0;
        m_Rays = 0;
        const auto t0 = std::chrono::steady_clock::now();
        while (m_Passes < config.MaxPasses) {
            s3dg_render_config rc{};
            rc.width = config.Width;
            rc.height = config.Height;
            rc.max_bounce_count = config.MaxBounceCount;
            rc.spp = std::min(config.PassesPerLaunch, config.MaxPasses - m_Passes);
            rc.seed = config.Seed;
            rc.rr_start_depth = config.RussianRouletteStart;
            rc.first_pass = m_Passes;
            s3dg_render_stats st{};
            {
                std::lock_guard<std::recursive_mutex> lock(detail::Lock());
                if (s3dg_render(ds, &rc, reinterpret_cast<float *>(m_Sum.data()), &st) != S3DG_OK) return false;
            }
            m_Passes += rc.spp;
            m_Rays += st.extension_rays + st.shadow_rays;
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (sec >= config.MaxRenderingSec) break;
        }
        m_Seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return true;
    }

    // Capture: tonemap (ACES, as s3d_pt.cpp:43) and the 8-bit sRGB encode; writes a binary PPM.
    bool Capture(Scene &, const char *filename, TONE_MAPPING_TYPE type = TONE_MAPPING_ACES_FILMIC) {
        const size_t n = m_Sum.size();
        std::vector<Color3> mapped(n);
        std::vector<uint8_t> rgb(n * 3);
        if (!ToneMapper::Map(type, m_Config.Width, m_Config.Height, m_Sum.data(), mapped.data(), rgb.data()))
            return false;
        FILE *f = fopen(filename, "wb");
        if (!f) return false;
        fprintf(f, "P6\n%d %d\n255\n", m_Config.Width, m_Config.Height);
        const bool ok = fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
        fclose(f);
        return ok;
    }

    const std::vector<Color3> &GetSum() const { return m_Sum; }
    s32 GetPassCount() const { return m_Passes; }
    u64 GetRayCount() const { return m_Rays; }
    double GetSeconds() const { return m_Seconds; }

private:
    Config m_Config;
    std::vector<Color3> m_Sum;
    s32 m_Passes = 0;
    u64 m_Rays = 0;
    double m_Seconds = 0.0;
};

}  // namespace s3d
