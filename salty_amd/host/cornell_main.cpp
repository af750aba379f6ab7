// cornell_main.cpp -- renders the 36-triangle Cornell box through the C++
// drop-in layer (s3d_gpu.h), the way a reference TestScene + main() would:
// materials from MaterialFactory, a Mesh from vertex triples, a quad light,
// a pinhole camera, PathTracer::Run with a pass budget, then Capture.
//
// usage: s3d_cornell <width> <height> <passes> <out.ppm> [sum.bin [geometry.bin]]
// geometry.bin: per mesh a u32 vertex count and the vertices, then the camera fov.
//
//        s3d_cornell --smd <file.smd> <width> <height> <passes> <out.ppm> <sum.bin> <setup.bin>
// renders an SMD mesh (Mesh::Create(filename): its materials and BMP textures)
// under a spherical light placed from its bounding box; setup.bin records the
// light centre and radius, the camera position and target and the fov (f32).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "s3d_gpu.h"

using namespace s3d;

namespace {

void quad(std::vector<Vertex> &v, Vector3 a, Vector3 b, Vector3 c, Vector3 d) {
    const Vector3 e1(b.x - a.x, b.y - a.y, b.z - a.z), e2(c.x - a.x, c.y - a.y, c.z - a.z);
    Vector3 n(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    const f32 l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    n = Vector3(n.x / l, n.y / l, n.z / l);
    const Vector3 tri[6] = {a, b, c, a, c, d};
    for (const Vector3 &p : tri) v.push_back(Vertex{p, n, Vector2()});
}

void box(std::vector<Vertex> &v, f32 sx, f32 sy, f32 sz, f32 cx, f32 cz, f32 deg) {
    const f32 a = deg * 3.14159265358979f / 180.0f, c = std::cos(a), s = std::sin(a);
    Vector3 p[8];
    const f32 hx = sx / 2, hz = sz / 2;
    const f32 xs[8] = {-hx, hx, hx, -hx, -hx, hx, hx, -hx}, ys[8] = {0, 0, 0, 0, sy, sy, sy, sy},
              zs[8] = {-hz, -hz, hz, hz, -hz, -hz, hz, hz};
    for (int i = 0; i < 8; ++i) p[i] = Vector3(xs[i] * c + zs[i] * s + cx, ys[i], -xs[i] * s + zs[i] * c + cz);
    const int f[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {3, 0, 4, 7}};
    for (auto &q : f) quad(v, p[q[0]], p[q[1]], p[q[2]], p[q[3]]);
}

}  // namespace

int render_smd(int argc, char **argv) {
    if (argc < 9) {
        fprintf(stderr, "usage: %s --smd file.smd width height passes out.ppm sum.bin setup.bin\n", argv[0]);
        return 2;
    }
    const int W = atoi(argv[3]), H = atoi(argv[4]), passes = atoi(argv[5]);
    IShape *mesh = Mesh::Create(argv[2]);
    if (!mesh) {
        fprintf(stderr, "cannot load %s\n", argv[2]);
        return 1;
    }
    const std::vector<Vertex> &v = static_cast<Mesh *>(mesh)->vertices;
    Vector3 lo = v[0].Position, hi = v[0].Position;
    for (const Vertex &x : v) {
        lo = Vector3(std::fmin(lo.x, x.Position.x), std::fmin(lo.y, x.Position.y), std::fmin(lo.z, x.Position.z));
        hi = Vector3(std::fmax(hi.x, x.Position.x), std::fmax(hi.y, x.Position.y), std::fmax(hi.z, x.Position.z));
    }
    const Vector3 c((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
    const f32 ext = std::fmax(hi.x - lo.x, std::fmax(hi.y - lo.y, hi.z - lo.z));
    const f32 setup[11] = {c.x, hi.y + 0.6f * ext, c.z + 0.3f * ext, 0.1f * ext, c.x, c.y, c.z + 1.6f * ext,
                           c.x, c.y, c.z, 2.0f * std::tan(20.0f * 3.14159265358979f / 180.0f)};
    IMaterial *glow = MaterialFactory::CreateLambert(Color3(0, 0, 0), nullptr, Color3(20, 20, 20));
    IShape *lamp = Sphere::Create(setup[3], Vector3(setup[0], setup[1], setup[2]), glow);
    Scene scene;
    scene.AddShape(mesh);
    scene.AddShape(lamp);
    scene.AddLight(lamp);
    scene.SetEnvironment(Color3(0.05f, 0.05f, 0.05f));
    scene.GetCamera().Update(Vector3(setup[4], setup[5], setup[6]), Vector3(setup[7], setup[8], setup[9]),
                             Vector3(0, 1, 0), setup[10], (f32)W / (f32)H, 1.0f);
    PathTracer pt;
    PathTracer::Config cfg;
    cfg.Width = W;
    cfg.Height = H;
    cfg.MaxBounceCount = 6;
    cfg.MaxPasses = passes;
    if (!pt.Run(cfg, scene) || !pt.Capture(scene, argv[6])) {
        fprintf(stderr, "render failed: %s\n", s3dg_last_error(scene.GetContext()));
        return 1;
    }
    FILE *f = fopen(argv[7], "wb");
    if (!f) return 1;
    fwrite(pt.GetSum().data(), sizeof(Color3), pt.GetSum().size(), f);
    fclose(f);
    f = fopen(argv[8], "wb");
    if (!f) return 1;
    fwrite(setup, sizeof(f32), 11, f);
    fclose(f);
    printf("passes=%d rays=%llu seconds=%.3f\n", pt.GetPassCount(), (unsigned long long)pt.GetRayCount(),
           pt.GetSeconds());
    mesh->Release();
    lamp->Release();
    glow->Release();
    return 0;
}

int main(int argc, char **argv) {
    if (argc > 1 && std::string(argv[1]) == "--smd") return render_smd(argc, argv);
    if (argc < 5) {
        fprintf(stderr, "usage: %s width height passes out.ppm [sum.bin]\n", argv[0]);
        return 2;
    }
    const int W = atoi(argv[1]), H = atoi(argv[2]), passes = atoi(argv[3]);
    // One mesh per material, as Mesh::Create takes a single IMaterial.
    IMaterial *white = MaterialFactory::CreateLambert(Color3(0.75f, 0.75f, 0.75f));
    IMaterial *red = MaterialFactory::CreateLambert(Color3(0.75f, 0.25f, 0.25f));
    IMaterial *green = MaterialFactory::CreateLambert(Color3(0.25f, 0.75f, 0.25f));
    IMaterial *light = MaterialFactory::CreateLambert(Color3(0, 0, 0), nullptr, Color3(15, 15, 15));
    std::vector<Vertex> vw, vr, vg, vl;
    quad(vw, {0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0});
    quad(vw, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1});
    quad(vw, {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0});
    quad(vr, {0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1});
    quad(vg, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0});
    box(vw, 0.3f, 0.6f, 0.3f, 0.33f, 0.37f, 15.0f);
    box(vw, 0.3f, 0.3f, 0.3f, 0.67f, 0.67f, -18.0f);
    const f32 y = 0.999f, lo = 0.375f, hi = 0.625f;
    quad(vl, {lo, y, lo}, {hi, y, lo}, {hi, y, hi}, {lo, y, hi});

    Scene scene;
    IShape *shapes[4] = {Mesh::Create((u32)vw.size(), vw.data(), white), Mesh::Create((u32)vr.size(), vr.data(), red),
                         Mesh::Create((u32)vg.size(), vg.data(), green), Mesh::Create((u32)vl.size(), vl.data(), light)};
    for (IShape *s : shapes) {
        if (!s) return 1;
        scene.AddShape(s);
    }
    scene.AddQuadLight(shapes[3], 0, 2, Vector3(lo, y, lo), Vector3(0.25f, 0, 0), Vector3(0, 0, 0.25f));
    const f32 fov = 2.0f * std::tan(20.0f * 3.14159265358979f / 180.0f);   // 40 degrees vertical
    scene.GetCamera().Update(Vector3(0.5f, 0.5f, 2.4f), Vector3(0.5f, 0.5f, 0.0f), Vector3(0, 1, 0), fov,
                             (f32)W / (f32)H, 1.0f);

    PathTracer pt;
    PathTracer::Config cfg;
    cfg.Width = W;
    cfg.Height = H;
    cfg.MaxBounceCount = 8;
    cfg.MaxRenderingSec = 59.8f;
    cfg.MaxPasses = passes;
    if (!pt.Run(cfg, scene)) {
        fprintf(stderr, "render failed: %s\n", s3dg_last_error(scene.GetContext()));
        return 1;
    }
    if (!pt.Capture(scene, argv[4])) return 1;
    if (argc > 5) {
        FILE *f = fopen(argv[5], "wb");
        if (!f) return 1;
        fwrite(pt.GetSum().data(), sizeof(Color3), pt.GetSum().size(), f);
        fclose(f);
    }
    if (argc > 6) {
        // geometry as built here: per mesh a u32 vertex count then the vertices
        FILE *f = fopen(argv[6], "wb");
        if (!f) return 1;
        const std::vector<Vertex> *meshes[4] = {&vw, &vr, &vg, &vl};
        for (const std::vector<Vertex> *m : meshes) {
            const u32 n = (u32)m->size();
            fwrite(&n, sizeof(n), 1, f);
            fwrite(m->data(), sizeof(Vertex), n, f);
        }
        fwrite(&fov, sizeof(fov), 1, f);
        fclose(f);
    }
    printf("passes=%d rays=%llu seconds=%.3f\n", pt.GetPassCount(), (unsigned long long)pt.GetRayCount(),
           pt.GetSeconds());
    for (IShape *s : shapes) s->Release();
    white->Release();
    red->Release();
    green->Release();
    light->Release();
    return 0;
}
