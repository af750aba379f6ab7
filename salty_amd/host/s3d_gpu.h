// s3d_gpu.h -- C++ drop-in host layer over the C ABI (include/s3dg.h).
//
// Mirrors the reference's host interfaces so that scene, camera, material, BVH
// and render-config code written against them ports with renames only:
//   IReference            src/include/s3d_reference.h:19-25
//   Vertex / HitRecord    src/include/s3d_shape.h:27-54
//   IShape factories      Mesh::Create (s3d_mesh.h:42-47), Instance::Create
//                         (s3d_instance.h:40), Sphere::Create (s3d_sphere.h:40)
//   MaterialFactory       src/include/s3d_materialfactory.h:22-49
//   Texture               src/include/s3d_texture.h (Init from BMP / HDR files)
//   Mesh::Create(file)    src/src/s3d_mesh.cpp:250-465 (SMD meshes)
//   Scene                 src/include/s3d_scene.h:22-109 (Intersect, GetRay)
//   PathTracer::Config    src/include/s3d_pt.h:33-40, PathTracer::Run
//   ToneMapper::Map       src/include/s3d_tonemapper.h:20-33
//   FilterNLM             src/include/s3d_denoiser.h:16
// Objects are reference counted like the reference's (count 1 on creation,
// Release deletes at 0, containers AddRef their children); factories return
// nullptr on failure instead of throwing.  All device work goes through the
// C ABI; nothing here touches HIP directly.
#pragma once
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <new>
#include <string>
#include <vector>

#include "../../include/s3dg.h"

namespace s3d {

typedef int32_t s32;
typedef uint32_t u32;
typedef uint64_t u64;
typedef float f32;

struct Vector2 {
    f32 x, y;
    Vector2() : x(0), y(0) {}
    Vector2(f32 a, f32 b) : x(a), y(b) {}
};
struct Vector3 {
    f32 x, y, z;
    Vector3() : x(0), y(0), z(0) {}
    Vector3(f32 a, f32 b, f32 c) : x(a), y(b), z(c) {}
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
};

// Vertex: same 32-byte layout as s3dg_vertex.
struct Vertex {
    Vector3 Position;
    Vector3 Normal;
    Vector2 TexCoord;
};
static_assert(sizeof(Vertex) == sizeof(s3dg_vertex), "Vertex layout");

struct Ray {
    Vector3 pos;
    Vector3 dir;
};

class Scene;

// HitRecord with the shape pointer replaced by (shape, triangle) indices.
struct HitRecord {
    f32 distance = 3.402823466e+38F;
    Vector2 barycentric;
    s32 shape = -1;      // top-level shape index in the scene
    s32 triangle = -1;   // triangle index in the mesh's input order, -2 = sphere
};

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

// ---------------------------------------------------------------------------
// Textures (s3d::Texture): RGB texels in memory, loaded from the reference's
// BMP and HDR formats by the library's loaders (TGA is not supported here).
// Like the reference's, a Texture is a plain object owned by its creator;
// materials keep a pointer to it.
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

private:
    s32 m_Width = 0, m_Height = 0;
    std::vector<float> m_Texels;
};

// ---------------------------------------------------------------------------
// Materials
// ---------------------------------------------------------------------------
struct IMaterial : IReference {
    s3dg_material desc{};
    const Texture *texture = nullptr;   // TexturedMaterial wrapper when set
    Color3 GetEmissive() const { return Color3(desc.emissive[0], desc.emissive[1], desc.emissive[2]); }
    bool HasDelta() const { return desc.type == S3DG_MAT_MIRROR || desc.type == S3DG_MAT_GLASS; }
};

// MaterialFactory::Create* with the reference's signatures: a texture wraps
// the material in a TexturedMaterial (an invalid texture leaves it plain; the
// reference would sample an empty texture).
struct MaterialFactory {
    static IMaterial *Make(int type, const Color3 &diffuse, const Color3 &specular, f32 ior, f32 power,
                           const Texture *pTexture, const Color3 &emissive) {
        IMaterial *m = new (std::nothrow) IMaterial();
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
// Shapes
// ---------------------------------------------------------------------------
struct IShape : IReference {
    enum Kind { MESH, INSTANCE, SPHERE } kind;
    explicit IShape(Kind k) : kind(k) {}
};

struct Mesh : IShape {
    std::vector<Vertex> vertices;
    IMaterial *material = nullptr;                 // one material for every triangle, or
    std::vector<IMaterial *> materials;            // per-triangle materials (SMD files)
    std::vector<int32_t> triangle_material;
    std::vector<Texture *> textures;               // owned, as Mesh::m_Textures
    int bvh_kind = S3DG_BVH8;
    Mesh() : IShape(MESH) {}
    ~Mesh() override {
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
            IMaterial *x = new (std::nothrow) IMaterial();
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
        return m;
    }
};

struct Instance : IShape {
    IShape *shape = nullptr;
    Matrix world;
    IMaterial *override_material = nullptr;
    u32 version = 0;   // bumped by UpdateMatrix; a built Scene catches up on its next use
    // Instance::UpdateMatrix (s3d_instance.cpp:95-101)
    void UpdateMatrix(const Matrix &m) {
        world = m;
        ++version;
    }
    Instance() : IShape(INSTANCE) {}
    ~Instance() override {
        SafeRelease(shape);
        SafeRelease(override_material);
    }
    // Instance::Create(IShape*, const Matrix&) (s3d_instance.cpp:105-106)
    static IShape *Create(IShape *pShape, const Matrix &world) {
        if (!pShape || pShape->kind != MESH) return nullptr;
        Instance *i = new (std::nothrow) Instance();
        if (!i) return nullptr;
        i->shape = pShape;
        pShape->AddRef();
        i->world = world;
        return i;
    }
};

struct Sphere : IShape {
    f32 radius = 0.0f;
    Vector3 center;
    IMaterial *material = nullptr;
    Sphere() : IShape(SPHERE) {}
    ~Sphere() override { SafeRelease(material); }
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
};

// ---------------------------------------------------------------------------
// Camera (PinholeCamera / ThinLensCamera::Update arguments)
// ---------------------------------------------------------------------------
struct Camera {
    s3dg_camera_desc desc{};
    void Update(const Vector3 &position, const Vector3 &target, const Vector3 &upward, f32 fov, f32 aspectRatio,
                f32 nearClip, f32 lensRadius = 0.0f) {
        desc.type = lensRadius > 0.0f ? S3DG_CAMERA_THINLENS : S3DG_CAMERA_PINHOLE;
        desc.position[0] = position.x; desc.position[1] = position.y; desc.position[2] = position.z;
        desc.target[0] = target.x; desc.target[1] = target.y; desc.target[2] = target.z;
        desc.upward[0] = upward.x; desc.upward[1] = upward.y; desc.upward[2] = upward.z;
        desc.fov = fov;
        desc.aspect = aspectRatio;
        desc.near_clip = nearClip;
        desc.lens_radius = lensRadius;
    }
};

// ---------------------------------------------------------------------------
// Scene: shape list intersected in order, light list, camera, environment.
// Build() uploads it; Intersect() is the single-ray slow path (one s3dg_trace).
// ---------------------------------------------------------------------------
class Scene {
public:
    Scene() = default;
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    virtual ~Scene() {
        Reset();
        if (m_Context) s3dg_destroy(m_Context);
    }

    void AddShape(IShape *s) {
        s->AddRef();
        m_Shapes.push_back(s);
        Invalidate();
    }
    // A sphere already added as a shape becomes a light (Scene::m_pLightList).
    void AddLight(IShape *s) {
        s->AddRef();
        m_Lights.push_back(s);
        Invalidate();
    }
    // Quad light made of `count` triangles of a MESH shape (an extension).
    void AddQuadLight(IShape *mesh, s32 firstTriangle, s32 count, const Vector3 &corner, const Vector3 &edge1,
                      const Vector3 &edge2) {
        QuadLight q{mesh, firstTriangle, count, corner, edge1, edge2};
        m_Quads.push_back(q);
        Invalidate();
    }
    Camera &GetCamera() { return m_Camera; }
    void SetEnvironment(const Color3 &c) { m_Env = c; Invalidate(); }
    // m_IBL: a lat-long texture whose SampleColor(dir) * 10 a missed ray sees
    // (Scene::SampleIBL); nullptr returns to the constant environment.
    void SetIBL(const Texture *t) { m_IBL = (t && t->IsValid()) ? t : nullptr; Invalidate(); }

    // Uploads the scene (builds every mesh tree with the reference SAH on the host).
    bool Build(int device = 0) {
        if (m_Scene) return SyncInstances();
        if (!m_Context && s3dg_create(&m_Context, device) != S3DG_OK) return false;
        std::vector<IMaterial *> mats;
        auto mat_index = [&](IMaterial *m) -> int32_t {
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
            if (s->kind == IShape::MESH) {
                d.type = S3DG_SHAPE_MESH;
                d.mesh = mesh_index(static_cast<Mesh *>(s));
            } else if (s->kind == IShape::INSTANCE) {
                Instance *in = static_cast<Instance *>(s);
                d.type = S3DG_SHAPE_INSTANCE;
                d.mesh = mesh_index(static_cast<Mesh *>(in->shape));
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) d.world[4 * r + c] = in->world.m[r][c];
                if (in->override_material) d.material = mat_index(in->override_material);
            } else {
                Sphere *sp = static_cast<Sphere *>(s);
                d.type = S3DG_SHAPE_SPHERE;
                d.material = mat_index(sp->material);
                d.center[0] = sp->center.x; d.center[1] = sp->center.y; d.center[2] = sp->center.z;
                d.radius = sp->radius;
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
        std::vector<s3dg_light_desc> lights;
        for (IShape *l : m_Lights) {
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
            mdesc.push_back(m->desc);
            mdesc.back().texture = tex_slot(m->texture);
        }
        const int32_t ibl = tex_slot(m_IBL);
        std::vector<s3dg_texture_desc> tdesc;
        for (const Texture *t : texs) tdesc.push_back(s3dg_texture_desc{t->GetWidth(), t->GetHeight(), t->GetTexels()});
        s3dg_scene_desc sd{};
        sd.mesh_count = (uint32_t)md.size();
        sd.meshes = md.data();
        sd.shape_count = (uint32_t)shapes.size();
        sd.shapes = shapes.data();
        sd.material_count = (uint32_t)mdesc.size();
        sd.materials = mdesc.data();
        sd.light_count = (uint32_t)lights.size();
        sd.lights = lights.data();
        sd.camera = m_Camera.desc;
        sd.environment[0] = m_Env.x; sd.environment[1] = m_Env.y; sd.environment[2] = m_Env.z;
        sd.texture_count = (uint32_t)tdesc.size();
        sd.textures = tdesc.data();
        sd.environment_texture = ibl;
        m_Versions.assign(m_Shapes.size(), 0);
        for (size_t i = 0; i < m_Shapes.size(); ++i)
            if (m_Shapes[i]->kind == IShape::INSTANCE) m_Versions[i] = static_cast<Instance *>(m_Shapes[i])->version;
        return s3dg_scene_create(m_Context, &sd, &m_Scene) == S3DG_OK;
    }

    // Scene::Intersect (s3d_scene.h:63-75) for one ray with the reference's
    // [F_HIT_MIN, F_HIT_MAX].  Batch callers should use s3dg_trace directly.
    bool Intersect(const Ray &ray, HitRecord &record) {
        if (!Build()) return false;
        s3dg_ray r{{ray.pos.x, ray.pos.y, ray.pos.z}, 1e-3f, {ray.dir.x, ray.dir.y, ray.dir.z}, 1e12f};
        s3dg_hit h;
        int32_t shape = -1;
        if (s3dg_trace(m_Scene, &r, 1, S3DG_TRACE_CLOSEST, &h, &shape) != S3DG_OK || h.prim == -1) return false;
        if (!(h.t < record.distance)) return false;
        record.distance = h.t;
        record.barycentric = Vector2(h.u, h.v);
        record.shape = shape;
        record.triangle = h.prim;
        return true;
    }

    s3dg_context *GetContext() { return m_Context; }
    s3dg_scene *GetDeviceScene() { return Build() ? m_Scene : nullptr; }

    // Virtual like Scene::Update (s3d_scene.h:92): a subclass moves its
    // instances here with Instance::UpdateMatrix.
    virtual void Update(float) {}

private:
    struct QuadLight {
        IShape *mesh;
        s32 first, count;
        Vector3 corner, e1, e2;
    };
    std::vector<IShape *> m_Shapes, m_Lights;
    std::vector<QuadLight> m_Quads;
    Camera m_Camera;
    Color3 m_Env;
    const Texture *m_IBL = nullptr;
    s3dg_context *m_Context = nullptr;
    s3dg_scene *m_Scene = nullptr;

    std::vector<u32> m_Versions;   // instance versions the device scene holds

    // Sends the matrices of instances moved since the last use (one top-level
    // rebuild for all of them).
    bool SyncInstances() {
        std::vector<uint32_t> ids;
        std::vector<float> worlds;
        for (size_t i = 0; i < m_Shapes.size(); ++i) {
            if (m_Shapes[i]->kind != IShape::INSTANCE) continue;
            const Instance *in = static_cast<const Instance *>(m_Shapes[i]);
            if (in->version == m_Versions[i]) continue;
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
        if (m_Scene) {
            s3dg_scene_destroy(m_Scene);
            m_Scene = nullptr;
        }
    }
    void Reset() {
        Invalidate();
        for (IShape *s : m_Shapes) s->Release();
        for (IShape *s : m_Lights) s->Release();
        m_Shapes.clear();
        m_Lights.clear();
        m_Quads.clear();
    }
};

// ---------------------------------------------------------------------------
// Tonemapping and NLM
// ---------------------------------------------------------------------------
enum TONE_MAPPING_TYPE {
    TONE_MAPPING_REINHARD = S3DG_TONEMAP_REINHARD,
    TONE_MAPPING_UNCHARTED2_FILMIC = S3DG_TONEMAP_UNCHARTED2,
    TONE_MAPPING_ACES_FILMIC = S3DG_TONEMAP_ACES,
};

struct ToneMapper {
    // ToneMapper::Map(type, w, h, const Color3*, Color3*) on the device.
    static bool Map(s3dg_context *ctx, TONE_MAPPING_TYPE type, s32 width, s32 height, const Color3 *pPixels,
                    Color3 *pResult, uint8_t *pRgb8 = nullptr) {
        return s3dg_tonemap(ctx, (int32_t)type, width, height, reinterpret_cast<const float *>(pPixels),
                            reinterpret_cast<float *>(pResult), pRgb8) == S3DG_OK;
    }
};

// FilterNLM(w, h, coeff, src, dst) with the reference's radii (search 6, patch 2).
inline bool FilterNLM(s3dg_context *ctx, int width, int height, float coeff, const Color4 *src, Color4 *dst,
                      int searchRadius = 6, int patchRadius = 2) {
    return s3dg_nlm(ctx, width, height, coeff, searchRadius, patchRadius, reinterpret_cast<const float *>(src),
                    reinterpret_cast<float *>(dst)) == S3DG_OK;
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
            if (s3dg_render(ds, &rc, reinterpret_cast<float *>(m_Sum.data()), &st) != S3DG_OK) return false;
            m_Passes += rc.spp;
            m_Rays += st.extension_rays + st.shadow_rays;
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (sec >= config.MaxRenderingSec) break;
        }
        m_Seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return true;
    }

    // Capture: tonemap (ACES, as s3d_pt.cpp:43) and the 8-bit sRGB encode; writes a binary PPM.
    bool Capture(Scene &scene, const char *filename, TONE_MAPPING_TYPE type = TONE_MAPPING_ACES_FILMIC) {
        const size_t n = m_Sum.size();
        std::vector<Color3> mapped(n);
        std::vector<uint8_t> rgb(n * 3);
        if (!ToneMapper::Map(scene.GetContext(), type, m_Config.Width, m_Config.Height, m_Sum.data(), mapped.data(),
                             rgb.data()))
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
