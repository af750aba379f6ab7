// iface_demo.cpp -- exercises the drop-in layer (s3d_gpu.h) through the
// reference's own interfaces only, the way the reference's code calls them:
//   IShape::IsHit / GetBox / GetCenter / CalcParam / Sample   (s3d_shape.h:60-67)
//   Triangle::Create, Leaf::Create, BVH8/BVH4/BVH2::Create, Sphere::Create,
//   Mesh::Create, Instance::Create
//   IMaterial::Shade / GetEmissive / GetBaseColor / HasDelta  (s3d_material.h:35-42)
//   ICamera::GetRay (PinholeCamera, ThinLensCamera)            (s3d_camera.h:19-24)
//   a Scene subclass filling m_Shapes / m_pLightList / m_pCamera as TestScene
//   does; Scene::GetRay / Intersect / SampleIBL / GetLight     (s3d_scene.h:56-90)
//   ToneMapper::Map                                           (s3d_tonemapper.h:33)
//   FilterNLM                                                 (s3d_denoiser.h:16)
// Inputs come from a seeded PCG; every line printed holds the inputs and the
// outputs of one call ("%.9g" floats round-trip exactly), for a test to check
// against the oracle.
#include <cmath>
#include <cstdio>
#include <vector>

#include "s3d_gpu.h"

using namespace s3d;

namespace {

PCG g_rng(2024);

f32 uni(f32 lo, f32 hi) { return lo + (hi - lo) * g_rng.GetAsF32(); }
Vector3 uni3(f32 lo, f32 hi) { return Vector3(uni(lo, hi), uni(lo, hi), uni(lo, hi)); }
Vector3 unit_dir() {
    for (;;) {
        const Vector3 v = uni3(-1.0f, 1.0f);
        const f32 l = v.Length();
        if (l > 0.1f && l <= 1.0f) return v / l;
    }
}
void p3(const Vector3 &v) { printf(" %.9g %.9g %.9g", v.x, v.y, v.z); }
void p2(const Vector2 &v) { printf(" %.9g %.9g", v.x, v.y); }

Vertex vtx(const Vector3 &p, const Vector3 &n, const Vector2 &uv) {
    Vertex v;
    v.Position = p;
    v.Normal = n;
    v.TexCoord = uv;
    return v;
}

// A Scene as TestScene builds one: shapes, lights and camera put in the
// protected members (s3d_testScene.cpp).
class DemoScene : public Scene {
public:
    DemoScene(IShape *mesh, IShape *lamp, PinholeCamera *cam) {
        mesh->AddRef();
        lamp->AddRef();
        m_Shapes.push_back(mesh);
        m_Shapes.push_back(lamp);
        lamp->AddRef();
        m_pLightList.push_back(lamp);
        m_pCamera = cam;
    }
};

}  // namespace

int main() {
    IMaterial *white = MaterialFactory::CreateLambert(Color3(0.7f, 0.6f, 0.5f));

    // ---- Triangle: IsHit, GetBox, GetCenter, CalcParam ----------------------
    Vertex tv[3] = {vtx(Vector3(-1, -1, 0), Vector3(0, 0, 1), Vector2(0, 0)),
                    vtx(Vector3(1.5f, -0.5f, 0.25f), Vector3(0.6f, 0, 0.8f), Vector2(1, 0)),
                    vtx(Vector3(-0.25f, 1.25f, -0.5f), Vector3(0, 0.6f, 0.8f), Vector2(0, 1))};
    IShape *tri = Triangle::Create(tv, white);
    {
        const BoundingBox b = tri->GetBox();
        printf("tri_box");
        p3(b.mini);
        p3(b.maxi);
        p3(tri->GetCenter());
        printf("\n");
    }
    for (int i = 0; i < 200; ++i) {
        const Vector3 o = uni3(-2.0f, 2.0f) + Vector3(0, 0, 3.0f);
        const Vector3 d = Vector3::UnitVector(uni3(-1.2f, 1.2f) - o);
        HitRecord rec;
        const bool hit = tri->IsHit(MakeRaySet(o, d), rec);
        printf("tri_hit");
        p3(o);
        p3(d);
        printf(" %d %.9g", hit ? 1 : 0, hit ? rec.distance : 0.0f);
        p2(rec.barycentric);
        printf(" %d %d", rec.pShape == tri ? 1 : 0, rec.pMaterial == white ? 1 : 0);
        if (hit) {
            Vector3 n;
            Vector2 uv;
            rec.pShape->CalcParam(o + d * rec.distance, rec.barycentric, &n, &uv);
            p3(n);
        } else {
            printf(" 0 0 0");
        }
        // a record already holding a nearer hit is left alone
        HitRecord near;
        near.distance = hit ? rec.distance : 1.0f;
        printf(" %d\n", tri->IsHit(MakeRaySet(o, d), near) ? 1 : 0);
    }

    // ---- BVH8 / BVH4 / BVH2 / Leaf over Triangle objects --------------------
    const int n = 300;
    std::vector<IShape *> tris(n);
    std::vector<Vertex> soup;
    for (int i = 0; i < n; ++i) {
        const Vector3 c = uni3(0.0f, 1.0f);
        Vertex v[3];
        for (int k = 0; k < 3; ++k) {
            v[k] = vtx(c + unit_dir() * uni(0.02f, 0.12f), Vector3(0, 1, 0), Vector2());
            soup.push_back(v[k]);
        }
        tris[i] = Triangle::Create(v, white);
    }
    printf("soup %d", n);
    for (const Vertex &v : soup) p3(v.Position);
    printf("\n");
    const char *names[3] = {"bvh8", "bvh4", "bvh2"};
    IShape *trees[3];
    for (int k = 0; k < 3; ++k) {
        std::vector<IShape *> arr = tris;
        trees[k] = k == 0 ? BVH8::Create(n, arr.data()) : k == 1 ? BVH4::Create(n, arr.data())
                                                                 : BVH2::Create(n, arr.data());
        // the array as the build reordered it, as input indices
        printf("%s_order", names[k]);
        for (IShape *s : arr) {
            int idx = -1;
            for (int i = 0; i < n; ++i)
                if (tris[i] == s) idx = i;
            printf(" %d", idx);
        }
        printf("\n");
    }
    IShape *leaf = Leaf::Create(8, tris.data());
    for (int i = 0; i < 300; ++i) {
        const Vector3 o = uni3(-0.2f, 1.2f);
        const Vector3 d = unit_dir();
        printf("soup_hit");
        p3(o);
        p3(d);
        for (IShape *s : {trees[0], trees[1], trees[2], leaf}) {
            HitRecord rec;
            const bool hit = s->IsHit(MakeRaySet(o, d), rec);
            int idx = -1;
            for (int j = 0; j < n; ++j)
                if (tris[j] == rec.pShape) idx = j;
            printf(" %d %.9g", hit ? idx : -1, hit ? rec.distance : 0.0f);
        }
        printf("\n");
    }

    // ---- Sphere: IsHit, CalcParam, Sample ------------------------------------
    IMaterial *glow = MaterialFactory::CreateLambert(Color3(0, 0, 0), nullptr, Color3(10, 10, 10));
    IShape *ball = Sphere::Create(0.75f, Vector3(0.2f, -0.1f, 0.3f), glow);
    for (int i = 0; i < 100; ++i) {
        const Vector3 o = uni3(-2.0f, 2.0f);
        const Vector3 d = Vector3::UnitVector(uni3(-0.5f, 0.5f) - o);
        HitRecord rec;
        const bool hit = ball->IsHit(MakeRaySet(o, d), rec);
        printf("sphere_hit");
        p3(o);
        p3(d);
        printf(" %d %.9g %d", hit ? 1 : 0, hit ? rec.distance : 0.0f, rec.pMaterial == glow ? 1 : 0);
        Vector3 nrm;
        Vector2 uv;
        if (hit) ball->CalcParam(o + d * rec.distance, rec.barycentric, &nrm, &uv);
        p3(nrm);
        printf("\n");
    }
    {
        PCG r(77);
        for (int i = 0; i < 50; ++i) {
            const u64 before = r.GetState();
            Vector3 pos;
            float pdf = 0.0f;
            ball->Sample(r, &pos, &pdf);
            printf("sphere_sample %llu %llu", (unsigned long long)before, (unsigned long long)r.GetState());
            p3(pos);
            printf(" %.9g\n", pdf);
        }
    }

    // ---- Materials: Shade, GetEmissive, GetBaseColor, HasDelta ---------------
    IMaterial *mats[5] = {MaterialFactory::CreateLambert(Color3(0.8f, 0.5f, 0.3f)),
                          MaterialFactory::CreateMirror(Color3(0.9f, 0.9f, 0.8f)),
                          MaterialFactory::CreateGlass(Color3(0.95f, 0.95f, 0.95f), 1.5f),
                          MaterialFactory::CreatePhong(Color3(0.7f, 0.7f, 0.6f), 40.0f),
                          MaterialFactory::CreatePlastic(Color3(0.3f, 0.6f, 0.3f), Color3(0.5f, 0.5f, 0.5f), 60.0f,
                                                         nullptr, Color3(0.5f, 0.25f, 0.0f))};
    for (int m = 0; m < 5; ++m) {
        const Color3 e = mats[m]->GetEmissive();
        const Color3 b = mats[m]->GetBaseColor(Vector2(0.3f, 0.7f));
        printf("material %d %d", m, mats[m]->HasDelta() ? 1 : 0);
        p3(e);
        p3(b);
        printf("\n");
        for (int i = 0; i < 40; ++i) {
            ShadingArg arg;
            arg.normal = unit_dir();
            arg.input = unit_dir();
            if (Vector3::Dot(arg.input, arg.normal) > 0.0f) arg.input = -arg.input;
            arg.texcoord = Vector2(uni(0, 1), uni(0, 1));
            arg.random = PCG(1000u + 97u * (u32)i + (u32)m);
            const u64 before = arg.random.GetState();
            const Color3 w = mats[m]->Shade(arg);
            printf("shade %d %llu", m, (unsigned long long)before);
            p3(arg.input);
            p3(arg.normal);
            printf(" %llu", (unsigned long long)arg.random.GetState());
            p3(arg.output);
            p3(w);
            printf(" %d\n", arg.dice ? 1 : 0);
        }
    }

    // ---- Cameras --------------------------------------------------------------
    PinholeCamera pin;
    pin.Update(Vector3(0.5f, 0.8f, 3.0f), Vector3(0.4f, 0.3f, 0.0f), Vector3(0, 1, 0), 0.8f, 1.5f, 1.0f);
    ThinLensCamera lens;
    lens.Update(Vector3(0.5f, 0.8f, 3.0f), Vector3(0.4f, 0.3f, 0.0f), Vector3(0, 1, 0), 0.8f, 1.5f, 1.0f, 0.2f);
    ICamera *cams[2] = {&pin, &lens};
    for (int c = 0; c < 2; ++c) {
        for (int i = 0; i < 40; ++i) {
            const f32 x = uni(-0.5f, 0.5f), y = uni(-0.5f, 0.5f);
            const u64 before = lens.GetRandom().GetState();
            const Ray r = cams[c]->GetRay(x, y);
            printf("camera %d %.9g %.9g %llu", c, x, y, (unsigned long long)before);
            p3(r.pos);
            p3(r.dir);
            printf("\n");
        }
    }

    // ---- Scene ---------------------------------------------------------------
    std::vector<Vertex> grid;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const f32 x0 = -1.5f + 0.5f * i, z0 = -1.5f + 0.5f * j;
            const Vector3 a(x0, 0, z0), b(x0 + 0.5f, 0.05f * i, z0), c2(x0 + 0.5f, 0, z0 + 0.5f), d(x0, 0.05f * j, z0 + 0.5f);
            for (const Vector3 &p : {a, b, c2, a, c2, d}) grid.push_back(vtx(p, Vector3(0, 1, 0), Vector2()));
        }
    IShape *mesh = Mesh::Create((u32)grid.size(), grid.data(), white);
    IShape *lamp = Sphere::Create(0.3f, Vector3(0.2f, 1.5f, 0.1f), glow);
    PinholeCamera scam;
    scam.Update(Vector3(0.0f, 2.0f, 4.0f), Vector3(0, 0, 0), Vector3(0, 1, 0), 0.9f, 1.0f, 1.0f);
    {
        DemoScene scene(mesh, lamp, &scam);
        printf("scene_grid %d", (int)grid.size());
        for (const Vertex &v : grid) p3(v.Position);
        printf("\n");
        std::vector<RaySet> sets;
        std::vector<HitRecord> singles;
        std::vector<int> single_hit;
        for (int i = 0; i < 200; ++i) {
            const f32 x = uni(-0.5f, 0.5f), y = uni(-0.5f, 0.5f);
            const Ray r = scene.GetRay(x, y);
            HitRecord rec;
            const bool hit = scene.Intersect(MakeRaySet(r.pos, r.dir), rec);
            sets.push_back(MakeRaySet(r.pos, r.dir));
            singles.push_back(rec);
            single_hit.push_back(hit ? 1 : 0);
            int kind = -1;
            if (hit) kind = rec.pShape == lamp ? 1 : 0;
            Vector3 nrm;
            Vector2 uv;
            if (hit) rec.pShape->CalcParam(r.pos + r.dir * rec.distance, rec.barycentric, &nrm, &uv);
            printf("scene_hit %.9g %.9g", x, y);
            p3(r.pos);
            p3(r.dir);
            printf(" %d %.9g %d", kind, hit ? rec.distance : 0.0f, hit && rec.pMaterial == (kind ? glow : white));
            p3(nrm);
            printf("\n");
        }
        // the drop-in's batched overload (an extension, not a reference
        // signature): the same 200 rays in one trace give the same records
        {
            std::vector<HitRecord> batch(sets.size());
            const size_t hits = scene.Intersect(sets.data(), batch.data(), sets.size());
            int equal = 0, nhit = 0;
            for (size_t i = 0; i < sets.size(); ++i) {
                const HitRecord &a = singles[i], &b = batch[i];
                nhit += single_hit[i];
                equal += !single_hit[i] ? (b.pShape == nullptr)
                                        : (a.distance == b.distance && a.barycentric.x == b.barycentric.x &&
                                           a.barycentric.y == b.barycentric.y && a.pShape == b.pShape &&
                                           a.pMaterial == b.pMaterial);
            }
            printf("scene_batch %d %d %d %d\n", equal, (int)sets.size(), (int)hits, nhit);
        }
        const Color3 env = scene.SampleIBL(Vector3(0, 1, 0));
        PCG lr(5);
        IShape *l = scene.GetLight(lr);
        printf("scene_misc");
        p3(env);
        printf(" %d\n", l == lamp ? 1 : 0);
    }

    // ---- ToneMapper::Map and FilterNLM ----------------------------------------
    const int W = 24, H = 16;
    std::vector<Color3> hdr(W * H), out(W * H);
    for (Color3 &c : hdr) c = Color3(uni(0, 4), uni(0, 3), uni(0, 2));
    printf("hdr");
    for (const Color3 &c : hdr) p3(c);
    printf("\n");
    const TONE_MAPPING_TYPE types[3] = {TONE_MAPPING_REINHARD, TONE_MAPPING_UNCHARTED2_FILMIC,
                                        TONE_MAPPING_ACES_FILMIC};
    for (int t = 0; t < 3; ++t) {
        ToneMapper::Map(types[t], W, H, hdr.data(), out.data());
        printf("tonemap %d", (int)types[t]);
        for (const Color3 &c : out) p3(c);
        printf("\n");
    }
    std::vector<Color4> src(W * H), dst(W * H);
    for (size_t i = 0; i < src.size(); ++i) src[i] = Color4{hdr[i].x * 0.25f, hdr[i].y * 0.25f, hdr[i].z * 0.25f, 1.0f};
    FilterNLM(W, H, 0.2f, src.data(), dst.data());
    printf("nlm");
    for (const Color4 &c : dst) printf(" %.9g %.9g %.9g %.9g", c.x, c.y, c.z, c.w);
    printf("\n");

    for (IShape *s : tris) s->Release();
    for (IShape *s : trees) SafeRelease(s);
    leaf->Release();
    tri->Release();
    ball->Release();
    mesh->Release();
    lamp->Release();
    for (IMaterial *m : mats) m->Release();
    white->Release();
    glow->Release();
    printf("done\n");
    return 0;
}
