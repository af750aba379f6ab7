// scene.cpp -- s3dg_scene_create: host-side build of the two-level compressed
// BVH (reference SAH per mesh, SAH over the top-level shapes), shading records
// and camera, then upload to HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "bvh.h"
#include "internal.h"

namespace s3dg {

int set_error(s3dg_context *ctx, int code, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int hip_check(s3dg_context *ctx, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, e == hipErrorOutOfMemory ? S3DG_ERR_NOMEM : S3DG_ERR_DEVICE, m);
}

int ensure_buffer(s3dg_context *ctx, void **ptr, size_t *have, size_t need) {
    if (*have >= need && *ptr) return S3DG_OK;
    if (*ptr) {
        hipError_t e = hipFree(*ptr);
        *ptr = nullptr;
        *have = 0;
        if (e != hipSuccess) return hip_check(ctx, e, "hipFree");
    }
    if (need == 0) return S3DG_OK;
    hipError_t e = hipMalloc(ptr, need);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return hip_check(ctx, e, "hipMalloc");
    }
    *have = need;
    return S3DG_OK;
}

namespace {

struct F3 {
    float x, y, z;
};
inline F3 f3(const float *p) { return F3{p[0], p[1], p[2]}; }
inline F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 cross(F3 a, F3 b) {
    return F3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
inline float dot(F3 a, F3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
inline F3 unit(F3 v) {
    const float m = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return F3{v.x / m, v.y / m, v.z / m};
}
inline F3 scale(F3 v, float s) { return F3{v.x * s, v.y * s, v.z * s}; }
inline F3 add(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }

// Matrix::Det and Matrix::Invert (src/include/s3d_math.h:1776-1842).  Each
// cofactor is a sum of three triple products minus three more, in the
// reference's term order; the reference's operator precedence divides only
// the last term by the determinant, which is reproduced (exact when det = 1).
float det4(const float m[4][4]) {
    // 24 signed quadruple products in the reference's order
    static const int P[24][4] = {
        {1, 2, 3, 4}, {1, 3, 4, 2}, {1, 4, 2, 3}, {2, 1, 4, 3}, {2, 3, 1, 4}, {2, 4, 3, 1},
        {3, 1, 2, 4}, {3, 2, 4, 1}, {3, 4, 1, 2}, {4, 1, 3, 2}, {4, 2, 1, 3}, {4, 3, 2, 1},
        {1, 2, 4, 3}, {1, 3, 2, 4}, {1, 4, 3, 2}, {2, 1, 3, 4}, {2, 3, 4, 1}, {2, 4, 1, 3},
        {3, 1, 4, 2}, {3, 2, 1, 4}, {3, 4, 2, 1}, {4, 1, 2, 3}, {4, 2, 3, 1}, {4, 3, 1, 2}};
    // P[k] lists, for rows 1..4, the column of the factor: _1P0 * _2P1 * _3P2 * _4P3
    float acc = 0.0f;
    for (int k = 0; k < 24; ++k) {
        const float prod = m[0][P[k][0] - 1] * m[1][P[k][1] - 1] * m[2][P[k][2] - 1] * m[3][P[k][3] - 1];
        if (k == 0) acc = prod;
        else if (k < 12) acc = acc + prod;
        else acc = acc - prod;
    }
    return acc;
}
void invert_ref(const float m[4][4], float out[4][4]) {
    // For entry (r,c): three positive and three negative triple products; the
    // factors are given as (row, col) pairs, 1-based, in the reference's order.
    static const unsigned char T[16][6][6] = {
        // _11
        {{2,2,3,3,4,4},{2,3,3,4,4,2},{2,4,3,2,4,3},{2,2,3,4,4,3},{2,3,3,2,4,4},{2,4,3,3,4,2}},
        // _12
        {{1,2,3,4,4,3},{1,3,3,2,4,4},{1,4,3,3,4,2},{1,2,3,3,4,4},{1,3,3,4,4,2},{1,4,3,2,4,3}},
        // _13
        {{1,2,2,3,4,4},{1,3,2,4,4,2},{1,4,2,2,4,3},{1,2,2,4,4,3},{1,3,2,2,4,4},{1,4,2,3,4,2}},
        // _14
        {{1,2,2,4,3,3},{1,3,2,2,3,4},{1,4,2,3,3,2},{1,2,2,3,3,4},{1,3,2,4,3,2},{1,4,2,2,3,3}},
        // _21
        {{2,1,3,4,4,3},{2,3,3,1,4,4},{2,4,3,3,4,1},{2,1,3,3,4,4},{2,3,3,4,4,1},{2,4,3,1,4,3}},
        // _22
        {{1,1,3,3,4,4},{1,3,3,4,4,1},{1,4,3,1,4,3},{1,1,3,4,4,3},{1,3,3,1,4,4},{1,4,3,3,4,1}},
        // _23
        {{1,1,2,4,4,3},{1,3,2,1,4,4},{1,4,2,3,4,1},{1,1,2,3,4,4},{1,3,2,4,4,1},{1,4,2,1,4,3}},
        // _24
        {{1,1,2,3,3,4},{1,3,2,4,3,1},{1,4,2,1,3,3},{1,1,2,4,3,3},{1,3,2,1,3,4},{1,4,2,3,3,1}},
        // _31
        {{2,1,3,2,4,4},{2,2,3,4,4,1},{2,4,3,1,4,2},{2,1,3,4,4,2},{2,2,3,1,4,4},{2,4,3,2,4,1}},
        // _32
        {{1,1,3,4,4,2},{1,2,3,1,4,4},{1,4,3,2,4,1},{1,1,3,2,4,4},{1,2,3,4,4,1},{1,4,3,1,4,2}},
        // _33
        {{1,1,2,2,4,4},{1,2,2,4,4,1},{1,4,2,1,4,2},{1,1,2,4,4,2},{1,2,2,1,4,4},{1,4,2,2,4,1}},
        // _34
        {{1,1,2,4,3,2},{1,2,2,1,3,4},{1,4,2,2,3,1},{1,1,2,2,3,4},{1,2,2,4,3,1},{1,4,2,1,3,2}},
        // _41
        {{2,1,3,3,4,2},{2,2,3,1,4,3},{2,3,3,2,4,1},{2,1,3,2,4,3},{2,2,3,3,4,1},{2,3,3,1,4,2}},
        // _42
        {{1,1,3,2,4,3},{1,2,3,3,4,1},{1,3,3,1,4,2},{1,1,3,3,4,2},{1,2,3,1,4,3},{1,3,3,2,4,1}},
        // _43
        {{1,1,2,3,4,2},{1,2,2,1,4,3},{1,3,2,2,4,1},{1,1,2,2,4,3},{1,2,2,3,4,1},{1,3,2,1,4,2}},
        // _44
        {{1,1,2,2,3,3},{1,2,2,3,3,1},{1,3,2,1,3,2},{1,1,2,3,3,2},{1,2,2,1,3,3},{1,3,2,2,3,1}}};
    const float d = det4(m);
    for (int e = 0; e < 16; ++e) {
        float p[6];
        for (int k = 0; k < 6; ++k) {
            const unsigned char *f = T[e][k];
            p[k] = m[f[0] - 1][f[1] - 1] * m[f[2] - 1][f[3] - 1] * m[f[4] - 1][f[5] - 1];
        }
        out[e / 4][e % 4] = p[0] + p[1] + p[2] - p[3] - p[4] - p[5] / d;
    }
}

// Plain double-precision inverse, used only to bound instance boxes.
bool inverse_double(const float mf[4][4], double inv[4][4]) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) a[r][c] = c < 4 ? (double)mf[r][c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (fabs(a[piv][c]) < 1e-300) return false;
        if (piv != c)
            for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[piv][k]);
        const double inv_p = 1.0 / a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] *= inv_p;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[r][c] = a[r][c + 4];
    return true;
}

}  // namespace

// PinholeCamera::Update / ThinLensCamera::Update (src/include/s3d_camera.h:63-93,
// 171-205): the screen vectors in the reference's operation order.
GCamera make_camera(const s3dg_camera_desc &c) {
    GCamera g;
    memset(&g, 0, sizeof(g));
    const F3 pos = f3(c.position), target = f3(c.target), up = f3(c.upward);
    const F3 dir = unit(sub(target, pos));
    const F3 cx = scale(scale(unit(cross(dir, up)), c.fov), c.aspect);
    const F3 cy = scale(unit(cross(dir, cx)), c.fov);
    const F3 cz = add(pos, scale(dir, c.near_clip));
    g.pos[0] = pos.x; g.pos[1] = pos.y; g.pos[2] = pos.z;
    g.cx[0] = cx.x; g.cx[1] = cx.y; g.cx[2] = cx.z;
    g.cy[0] = cy.x; g.cy[1] = cy.y; g.cy[2] = cy.z;
    g.cz[0] = cz.x; g.cz[1] = cz.y; g.cz[2] = cz.z;
    g.type = c.type;
    g.lens_radius = c.lens_radius;
    const F3 tp = sub(target, pos);
    g.focal_distance = sqrtf((tp.x * tp.x) + (tp.y * tp.y) + (tp.z * tp.z));
    return g;
}

void make_material(const s3dg_material &m, GMat &g) {
    memset(&g, 0, sizeof(g));
    for (int a = 0; a < 3; ++a) {
        g.diffuse[a] = m.diffuse[a];
        g.specular[a] = m.specular[a];
        g.emissive[a] = m.emissive[a];
    }
    g.ior = m.ior;
    g.power = m.power;
    g.type = m.type;
    g.delta = (m.type == S3DG_MAT_MIRROR || m.type == S3DG_MAT_GLASS) ? 1 : 0;
    // Russian-roulette thresholds: max component of the colour the material
    // rolls on, floored at 0.01 (s3d_lambert.cpp:23-30, s3d_phong.cpp:23-31,
    // s3d_plastic.cpp:23-37, where the specular threshold ends up equal to the
    // diffuse one).
    const float *base = (m.type == S3DG_MAT_LAMBERT || m.type == S3DG_MAT_PLASTIC) ? m.diffuse : m.specular;
    float t = base[0];
    t = base[1] > t ? base[1] : t;
    t = base[2] > t ? base[2] : t;
    t = 0.01f > t ? 0.01f : t;
    g.thr0 = t;
    g.thr1 = t;
}

namespace {

float scale_margin(const Aabb &b) {
    float m = 0.0f;
    for (int a = 0; a < 3; ++a) {
        m = std::max(m, fabsf(b.lo[a]));
        m = std::max(m, fabsf(b.hi[a]));
        m = std::max(m, b.hi[a] - b.lo[a]);
    }
    return 4e-6f * m + 1e-30f;
}

// Inverse matrix and world box of an instance of a mesh whose box is mb
// (Instance ctor / UpdateMatrix, src/src/s3d_instance.cpp:22-32, 95-101):
// the inverse is the reference's Matrix::Invert; the box is the region the
// inverse actually used maps into mb, grown to cover rounding.  False when the
// matrix is singular.
bool instance_transform(const float world[16], const Aabb &mb, GInst &g, Aabb &wbox) {
    float w[4][4];
    memcpy(w, world, sizeof(w));
    invert_ref(w, g.inv);
    // a singular matrix divides by a zero determinant: NaN or inf entries
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            if (!std::isfinite(g.inv[r][c])) return false;
    double back[4][4];
    if (!inverse_double(g.inv, back)) return false;
    for (int a = 0; a < 3; ++a) { wbox.lo[a] = 3.4e38f; wbox.hi[a] = -3.4e38f; }
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? mb.hi[0] : mb.lo[0], (c & 2) ? mb.hi[1] : mb.lo[1],
                             (c & 4) ? mb.hi[2] : mb.lo[2]};
        double q[4];
        for (int k = 0; k < 4; ++k) q[k] = p[0] * back[0][k] + p[1] * back[1][k] + p[2] * back[2][k] + back[3][k];
        for (int a = 0; a < 3; ++a) {
            const double v = q[a] / q[3];
            wbox.lo[a] = std::min(wbox.lo[a], (float)(v - 1e-6 * fabs(v)));
            wbox.hi[a] = std::max(wbox.hi[a], (float)(v + 1e-6 * fabs(v)));
        }
    }
    // a small extra margin covers the rounding of the transformed ray
    const float m = scale_margin(wbox) * 4.0f;
    for (int a = 0; a < 3; ++a) { wbox.lo[a] -= m; wbox.hi[a] += m; }
    return true;
}

// Top-level tree over the shapes' world boxes, appended to comp (whose
// node_offset places it in the device node array).  Shape records come out in
// the tree's leaf order: inst_dev[k] is the record of shape inst_dev[k].shape,
// inst_pos[shape] its position.  Ties between shapes go by shape index,
// whatever the tree.
uint32_t build_tlas(const s3dg_context *ctx, const std::vector<GInst> &inst, const std::vector<Aabb> &wbox,
                    Compressed &comp, std::vector<GInst> &inst_dev, std::vector<int32_t> &inst_pos,
                    int32_t &depth) {
    RefTree top;
    if (ctx->leaf_size) build_traversal_tree(wbox, 1, top);
    else build_reference_tree(wbox, 8, top);
    comp.max_depth = 0;
    std::vector<int32_t> top_order;
    // no margin: the traversal widens its slab distances by their rounding
    // bound (the instance boxes in wbox carry their own for the transform)
    const uint32_t root = compress_tree(top, wbox, 0u, true, 0.0f, comp, top_order);
    inst_dev.resize(inst.size());
    inst_pos.assign(inst.size(), 0);
    for (size_t k = 0; k < inst.size(); ++k) {
        const int32_t shape = top.order[top_order[k]];
        inst_dev[k] = inst[shape];
        inst_pos[shape] = (int32_t)k;
    }
    depth = comp.max_depth;
    return root;
}

template <class T>
int upload(s3dg_context *ctx, void **dst, const std::vector<T> &src, size_t &total) {
    const size_t bytes = std::max<size_t>(src.size() * sizeof(T), 256);
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return hip_check(ctx, e, "hipMalloc(scene)");
    total += bytes;
    if (!src.empty()) {
        e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_check(ctx, e, "hipMemcpy(scene)");
    }
    return S3DG_OK;
}

}  // namespace
}  // namespace s3dg

using namespace s3dg;

extern "C" int s3dg_scene_create(s3dg_context *ctx, const s3dg_scene_desc *d, s3dg_scene **out) {
    if (!ctx || !d || !out) return set_error(ctx, S3DG_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->shape_count == 0) return set_error(ctx, S3DG_ERR_INVALID, "scene has no shapes");
    if ((d->mesh_count && !d->meshes) || !d->shapes || (d->material_count && !d->materials) ||
        (d->light_count && !d->lights))
        return set_error(ctx, S3DG_ERR_INVALID, "null array in scene description");
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_check(ctx, he, "hipSetDevice");

    s3dg_scene *sc = new (std::nothrow) s3dg_scene();
    if (!sc) return set_error(ctx, S3DG_ERR_NOMEM, "host allocation");
    sc->ctx = ctx;

    // ---- textures (TexturedMaterial, IBL) -----------------------------------
    if (d->texture_count && !d->textures) {
        delete sc;
        return set_error(ctx, S3DG_ERR_INVALID, "null texture array");
    }
    std::vector<GTexture> texrecs(d->texture_count);
    uint64_t texel_floats = 0;
    for (uint32_t i = 0; i < d->texture_count; ++i) {
        const s3dg_texture_desc &t = d->textures[i];
        if (t.width <= 0 || t.height <= 0 || !t.rgb || (int64_t)t.width * t.height > (1ll << 28)) {
            delete sc;
            return set_error(ctx, S3DG_ERR_INVALID, "bad texture");
        }
        texrecs[i].offset = texel_floats;
        texrecs[i].width = t.width;
        texrecs[i].height = t.height;
        texel_floats += 3ull * (uint64_t)t.width * (uint64_t)t.height;
    }
    if (d->environment_texture < 0 || (uint32_t)d->environment_texture > d->texture_count) {
        delete sc;
        return set_error(ctx, S3DG_ERR_INVALID, "environment texture out of range");
    }

    std::vector<GMat> mats(d->material_count);
    bool textured = false;
    for (uint32_t i = 0; i < d->material_count; ++i) {
        make_material(d->materials[i], mats[i]);
        const int32_t tx = d->materials[i].texture;
        if (tx < 0 || (uint32_t)tx > d->texture_count) {
            delete sc;
            return set_error(ctx, S3DG_ERR_INVALID, "material texture out of range");
        }
        mats[i].texture = tx - 1;
        textured = textured || tx > 0;
    }

    std::vector<GTri> tris;
    std::vector<GTriShade> shade;
    std::vector<uint32_t> rank;   // depth-first (tie) rank of every stored triangle
    std::vector<GTriUV> uvs;      // texture coordinates, kept when a material is textured
    Compressed comp;
    std::vector<Aabb> mesh_box(d->mesh_count);
    int32_t blas_depth = 0;
    size_t total_tris = 0;
    for (uint32_t mi = 0; mi < d->mesh_count; ++mi) total_tris += d->meshes[mi].vertex_count / 3;
    tris.reserve(total_tris);
    shade.reserve(total_tris);
    rank.reserve(total_tris);
    std::vector<int32_t> leaf_order;

    for (uint32_t mi = 0; mi < d->mesh_count; ++mi) {
        const s3dg_mesh_desc &md = d->meshes[mi];
        if (md.vertex_count % 3 != 0 || (md.vertex_count && !md.vertices)) {
            delete sc;
            return set_error(ctx, S3DG_ERR_INVALID, "mesh vertex_count must be a multiple of 3");
        }
        const uint32_t n = md.vertex_count / 3;
        if (n == 0) {
            delete sc;
            return set_error(ctx, S3DG_ERR_INVALID, "mesh without triangles");
        }
        const int kind = (md.bvh_kind == S3DG_BVH2 || md.bvh_kind == S3DG_BVH4) ? md.bvh_kind : S3DG_BVH8;
        std::vector<Aabb> boxes(n);
        for (uint32_t i = 0; i < n; ++i) {
            const s3dg_vertex *v = md.vertices + 3 * i;
            Aabb &b = boxes[i];
            for (int a = 0; a < 3; ++a) {
                float lo = v[0].position[a], hi = v[0].position[a];
                lo = v[1].position[a] < lo ? v[1].position[a] : lo;
                lo = v[2].position[a] < lo ? v[2].position[a] : lo;
                hi = v[1].position[a] > hi ? v[1].position[a] : hi;
                hi = v[2].position[a] > hi ? v[2].position[a] : hi;
                b.lo[a] = lo;
                b.hi[a] = hi;
            }
        }
        // the reference tree fixes the tie rank (its depth-first order); the
        // kernels traverse the traversal tree unless leaf_size is 0
        RefTree tree, trav;
        build_reference_tree(boxes, kind, tree);
        if (ctx->leaf_size) build_traversal_tree(boxes, (int)ctx->leaf_size, trav);
        const RefTree &walk = ctx->leaf_size ? trav : tree;
        SceneMeshInfo info;
        info.tri_offset = (uint32_t)tris.size();
        info.tri_count = n;
        info.order = tree.order;
        std::vector<int32_t> ref_pos(n);
        for (uint32_t j = 0; j < n; ++j) ref_pos[tree.order[j]] = (int32_t)j;
        const Aabb &rb = tree.nodes[tree.root].box;
        mesh_box[mi] = rb;
        comp.max_depth = 0;
        info.root = compress_tree(walk, boxes, info.tri_offset, false, 0.0f, comp, leaf_order);
        blas_depth = std::max(blas_depth, comp.max_depth);
        // triangles in the order of the compressed leaves, ranked by their
        // depth-first position in the reference tree
        info.stored_of_input.assign(n, 0);
        for (uint32_t j = 0; j < n; ++j) {
            const int32_t src = walk.order[leaf_order[j]];
            info.stored_of_input[src] = (int32_t)j;
            rank.push_back(info.tri_offset + (uint32_t)ref_pos[src]);
            const s3dg_vertex *v = md.vertices + 3 * src;
            GTri t;
            GTriShade s;
            for (int a = 0; a < 3; ++a) {
                // Triangle ctor (s3d_triangle.cpp:38-39): edges from v0
                t.v0[a] = v[0].position[a];
                t.e0[a] = v[1].position[a] - v[0].position[a];
                t.e1[a] = v[2].position[a] - v[0].position[a];
                s.n0[a] = v[0].normal[a];
                s.n1[a] = v[1].normal[a];
                s.n2[a] = v[2].normal[a];
            }
            s.material = md.triangle_material ? md.triangle_material[src] : md.material;
            s.input_index = src;
            s.pad = (int32_t)mi;
            if (s.material < 0 || (uint32_t)s.material >= d->material_count) {
                delete sc;
                return set_error(ctx, S3DG_ERR_INVALID, "triangle material index out of range");
            }
            tris.push_back(t);
            shade.push_back(s);
            if (textured) {
                GTriUV w;
                for (int k = 0; k < 3; ++k) {
                    w.uv[2 * k] = v[k].texcoord[0];
                    w.uv[2 * k + 1] = v[k].texcoord[1];
                }
                uvs.push_back(w);
            }
        }
        sc->meshes.push_back(std::move(info));
    }

    // ---- top-level shapes ----------------------------------------------
    std::vector<GInst> inst(d->shape_count);
    std::vector<Aabb> wbox(d->shape_count);
    for (uint32_t i = 0; i < d->shape_count; ++i) {
        const s3dg_shape_desc &s = d->shapes[i];
        GInst &g = inst[i];
        memset(&g, 0, sizeof(g));
        g.shape = (int32_t)i;
        if (s.type == S3DG_SHAPE_SPHERE) {
            if (s.material < 0 || (uint32_t)s.material >= d->material_count || !(s.radius > 0.0f)) {
                delete sc;
                return set_error(ctx, S3DG_ERR_INVALID, "bad sphere");
            }
            g.type = INST_SPHERE;
            for (int a = 0; a < 3; ++a) g.center[a] = s.center[a];
            g.radius = s.radius;
            g.material = s.material;
            for (int a = 0; a < 3; ++a) {
                wbox[i].lo[a] = s.center[a] - s.radius;
                wbox[i].hi[a] = s.center[a] + s.radius;
            }
        } else if (s.type == S3DG_SHAPE_MESH || s.type == S3DG_SHAPE_INSTANCE) {
            if (s.mesh < 0 || (uint32_t)s.mesh >= d->mesh_count) {
                delete sc;
                return set_error(ctx, S3DG_ERR_INVALID, "shape mesh index out of range");
            }
            g.blas_root = sc->meshes[s.mesh].root;
            const Aabb &mb = mesh_box[s.mesh];
            g.material = -1;
            if (s.type == S3DG_SHAPE_MESH) {
                g.type = INST_MESH;
                wbox[i] = mb;
            } else {
                g.type = INST_TRANSFORMED;
                if (s.material >= (int32_t)d->material_count) {
                    delete sc;
                    return set_error(ctx, S3DG_ERR_INVALID, "instance material override out of range");
                }
                g.material = s.material >= 0 ? s.material : -1;
                if (!instance_transform(s.world, mb, g, wbox[i])) {
                    delete sc;
                    return set_error(ctx, S3DG_ERR_INVALID, "singular instance matrix");
                }
            }
        } else {
            delete sc;
            return set_error(ctx, S3DG_ERR_INVALID, "unknown shape type");
        }
    }

    SceneDev &dv = sc->dev;
    memset(&dv, 0, sizeof(dv));
    std::vector<GInst> inst_dev;
    int32_t tlas_depth = 0;
    const bool single = d->shape_count == 1 && d->shapes[0].type == S3DG_SHAPE_MESH;
    std::vector<int32_t> inst_pos(d->shape_count, 0);
    if (single) {
        sc->rec.single_shape = 0;
        sc->rec.root = inst[0].blas_root;
        inst_dev = inst;
        inst_pos[0] = 0;
    } else {
        // the top-level tree follows the meshes' nodes, with room for every
        // rebuild an instance update can make (at most one node per shape)
        const uint32_t base = (uint32_t)comp.nodes.size();
        sc->rec.root = build_tlas(ctx, inst, wbox, comp, inst_dev, inst_pos, tlas_depth);
        sc->host.tlas_base = base;
        sc->host.tlas_cap = std::max<uint32_t>((uint32_t)(comp.nodes.size() - base), d->shape_count + 1);
        comp.nodes.resize(base + sc->host.tlas_cap);
        sc->rec.single_shape = -1;
        sc->stats.tlas_nodes = 0;
    }

    // ---- lights ----------------------------------------------------------
    std::vector<GLight> lights(d->light_count);
    for (uint32_t i = 0; i < d->light_count; ++i) {
        const s3dg_light_desc &l = d->lights[i];
        GLight &g = lights[i];
        memset(&g, 0, sizeof(g));
        if (l.shape < 0 || (uint32_t)l.shape >= d->shape_count) {
            delete sc;
            return set_error(ctx, S3DG_ERR_INVALID, "light shape out of range");
        }
        g.shape = l.shape;
        const s3dg_shape_desc &s = d->shapes[l.shape];
        if (l.type == S3DG_LIGHT_SHAPE) {
            if (s.type != S3DG_SHAPE_SPHERE) {
                delete sc;
                return set_error(ctx, S3DG_ERR_UNSUPPORTED, "shape lights must be spheres (only Sphere::Sample exists)");
            }
            g.type = LIGHT_SPHERE;
            g.tri[0] = inst_pos[l.shape];
            const GInst &sp = inst[l.shape];
            for (int a = 0; a < 3; ++a) {
                g.corner[a] = sp.center[a];
                g.e2[a] = mats[sp.material].emissive[a];
            }
            g.e1[0] = sp.radius;
            // Sphere::Sample pdf (s3d_sphere.cpp:113): 1 / (4 pi r^2), left to right
            g.pdf = 1.0f / (4.0f * 3.1415926535897932384626433832795f * s.radius * s.radius);
        } else if (l.type == S3DG_LIGHT_QUAD) {
            if (s.type != S3DG_SHAPE_MESH || l.triangle_count < 1 || l.triangle_count > 3) {
                delete sc;
                return set_error(ctx, S3DG_ERR_UNSUPPORTED, "quad lights need a MESH shape and 1..3 triangles");
            }
            const SceneMeshInfo &mi = sc->meshes[s.mesh];
            if (l.first_triangle < 0 || (uint32_t)(l.first_triangle + l.triangle_count) > mi.tri_count) {
                delete sc;
                return set_error(ctx, S3DG_ERR_INVALID, "quad light triangles out of range");
            }
            g.type = LIGHT_QUAD;
            g.count = l.triangle_count;
            for (int k = 0; k < l.triangle_count; ++k)
                g.tri[k] = (int32_t)(mi.tri_offset + (uint32_t)mi.stored_of_input[l.first_triangle + k]);
            for (int a = 0; a < 3; ++a) {
                g.corner[a] = l.corner[a];
                g.e1[a] = l.edge1[a];
                g.e2[a] = l.edge2[a];
            }
            const F3 c = cross(f3(l.edge1), f3(l.edge2));
            g.pdf = 1.0f / sqrtf(dot(c, c));
        } else {
            delete sc;
            return set_error(ctx, S3DG_ERR_INVALID, "unknown light type");
        }
    }

    // ---- camera: PinholeCamera::Update / ThinLensCamera::Update ----------
    sc->rec.cam = make_camera(d->camera);
    for (int a = 0; a < 3; ++a) sc->rec.env[a] = d->environment[a];
    sc->rec.env_tex = d->environment_texture - 1;
    sc->rec.light_count = (int32_t)d->light_count;
    sc->rec.light_base = 0;
    // at most two entries per node on the path (a group of inner children and
    // a group of leaves or shapes) plus the return from an instance
    dv.stack_need = 2 * blas_depth + 2 * tlas_depth + 4;
    dv.tri_batch = ctx->tri_batch;

    sc->counts.nodes = (uint32_t)comp.nodes.size();
    sc->counts.tris = (uint32_t)tris.size();
    sc->counts.insts = (uint32_t)inst_dev.size();
    std::vector<GShapeRef> shape_ref(d->shape_count);
    for (uint32_t k = 0; k < d->shape_count; ++k) {
        shape_ref[k].inst = inst_pos[k];
        shape_ref[k].material = inst[k].material;
    }
    sc->counts.mats = (uint32_t)mats.size();
    sc->counts.lights = (uint32_t)lights.size();
    sc->counts.shapes = (uint32_t)inst_pos.size();
    sc->counts.textures = (uint32_t)texrecs.size();
    sc->counts.texel_floats = texel_floats;
    size_t total = 0;
    int rc;
    if ((rc = upload(ctx, &sc->d_nodes, comp.nodes, total)) || (rc = upload(ctx, &sc->d_tris, tris, total)) ||
        (rc = upload(ctx, &sc->d_shade, shade, total)) || (rc = upload(ctx, &sc->d_insts, inst_dev, total)) ||
        (rc = upload(ctx, &sc->d_mats, mats, total)) || (rc = upload(ctx, &sc->d_lights, lights, total)) ||
        (rc = upload(ctx, &sc->d_shape_ref, shape_ref, total)) ||
        (rc = upload(ctx, &sc->d_rec, std::vector<SceneRec>{sc->rec}, total)) ||
        (rc = upload(ctx, &sc->d_rank, rank, total)) ||
        (!texrecs.empty() && (rc = upload(ctx, &sc->d_textures, texrecs, total))) ||
        (!uvs.empty() && (rc = upload(ctx, &sc->d_uv, uvs, total)))) {
        s3dg_scene_destroy(sc);
        return rc;
    }
    if (texel_floats) {
        // texels go up texture by texture, with no host copy of them all
        hipError_t e = hipMalloc(&sc->d_texels, texel_floats * sizeof(float));
        for (uint32_t i = 0; e == hipSuccess && i < d->texture_count; ++i)
            e = hipMemcpy((float *)sc->d_texels + texrecs[i].offset, d->textures[i].rgb,
                          3ull * (uint64_t)texrecs[i].width * (uint64_t)texrecs[i].height * sizeof(float),
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rc = hip_check(ctx, e, "texel upload");
            s3dg_scene_destroy(sc);
            return rc;
        }
        total += texel_floats * sizeof(float);
    }
    if ((rc = workspace_init(ctx, sc->ws))) {
        s3dg_scene_destroy(sc);
        return rc;
    }
    dv.nodes = (const GNode *)sc->d_nodes;
    dv.tris = (const GTri *)sc->d_tris;
    dv.shade = (const GTriShade *)sc->d_shade;
    dv.insts = (const GInst *)sc->d_insts;
    dv.mats = (const GMat *)sc->d_mats;
    dv.lights = (const GLight *)sc->d_lights;
    dv.shape_ref = (const GShapeRef *)sc->d_shape_ref;
    dv.recs = (const SceneRec *)sc->d_rec;
    dv.tri_rank = (const uint32_t *)sc->d_rank;
    dv.textures = (const GTexture *)sc->d_textures;
    dv.texels = (const float *)sc->d_texels;
    dv.tri_uv = (const GTriUV *)sc->d_uv;

    sc->host.inst = inst;
    sc->host.shape_mesh.resize(d->shape_count);
    for (uint32_t i = 0; i < d->shape_count; ++i) sc->host.shape_mesh[i] = d->shapes[i].mesh;
    sc->host.wbox = wbox;
    sc->host.mesh_box = mesh_box;
    sc->host.lights = lights;
    sc->host.blas_depth = blas_depth;

    sc->stats.triangles = tris.size();
    sc->stats.blas_nodes = comp.nodes.size();
    sc->stats.device_bytes = total;
    sc->stats.max_depth = dv.stack_need;
    sc->stats.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = sc;
    return S3DG_OK;
}

extern "C" int s3dg_scene_destroy(s3dg_scene *sc) {
    if (!sc) return S3DG_ERR_INVALID;
    void *ptrs[] = {sc->d_nodes, sc->d_tris, sc->d_shade, sc->d_insts, sc->d_mats, sc->d_lights,
                    sc->d_shape_ref, sc->d_rec, sc->d_rank, sc->d_textures, sc->d_texels, sc->d_uv};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    workspace_free(sc->ws);
    delete sc;
    return S3DG_OK;
}

extern "C" int s3dg_scene_get_stats(const s3dg_scene *sc, s3dg_scene_stats *out) {
    if (!sc || !out) return S3DG_ERR_INVALID;
    *out = sc->stats;
    return S3DG_OK;
}

extern "C" int s3dg_build_order(uint32_t vertex_count, const s3dg_vertex *v, int32_t kind, int32_t *order,
                                uint32_t *node_count) {
    if (!v || !order || vertex_count == 0 || vertex_count % 3 != 0) return S3DG_ERR_INVALID;
    if (kind != S3DG_BVH2 && kind != S3DG_BVH4 && kind != S3DG_BVH8) return S3DG_ERR_INVALID;
    const uint32_t n = vertex_count / 3;
    std::vector<Aabb> boxes(n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float p0 = v[3 * i].position[a], p1 = v[3 * i + 1].position[a], p2 = v[3 * i + 2].position[a];
            float lo = p0, hi = p0;
            lo = p1 < lo ? p1 : lo;
            lo = p2 < lo ? p2 : lo;
            hi = p1 > hi ? p1 : hi;
            hi = p2 > hi ? p2 : hi;
            boxes[i].lo[a] = lo;
            boxes[i].hi[a] = hi;
        }
    RefTree tree;
    build_reference_tree(boxes, kind, tree);
    std::copy(tree.order.begin(), tree.order.end(), order);
    if (node_count) *node_count = (uint32_t)tree.nodes.size();
    return S3DG_OK;
}

extern "C" int s3dg_scene_mesh_order(const s3dg_scene *sc, uint32_t mesh, int32_t *order) {
    if (!sc || !order || mesh >= sc->meshes.size()) return S3DG_ERR_INVALID;
    const auto &o = sc->meshes[mesh].order;
    std::copy(o.begin(), o.end(), order);
    return S3DG_OK;
}

// Instance::UpdateMatrix (src/src/s3d_instance.cpp:95-101) for several
// instances at once: new inverse matrices and world boxes, then the top-level
// tree is rebuilt on the host over the unchanged mesh trees and uploaded into
// the node range reserved for it, with the reordered shape records, shape
// references and light records.
extern "C" int s3dg_scene_update_instances(s3dg_scene *sc, uint32_t count, const uint32_t *shapes,
                                           const float *worlds) {
    if (!sc || (count && (!shapes || !worlds))) return S3DG_ERR_INVALID;
    s3dg_context *ctx = sc->ctx;
    if (sc->rec.single_shape >= 0) return set_error(ctx, S3DG_ERR_INVALID, "scene has no instances");
    std::vector<GInst> inst = sc->host.inst;
    std::vector<Aabb> wbox = sc->host.wbox;
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t sh = shapes[k];
        if (sh >= inst.size() || inst[sh].type != INST_TRANSFORMED)
            return set_error(ctx, S3DG_ERR_INVALID, "shape is not an instance");
        const int32_t mesh = sc->host.shape_mesh[sh];
        if (!instance_transform(worlds + 16 * (size_t)k, sc->host.mesh_box[mesh], inst[sh], wbox[sh]))
            return set_error(ctx, S3DG_ERR_INVALID, "singular instance matrix");
    }
    Compressed comp;
    comp.node_offset = sc->host.tlas_base;
    std::vector<GInst> inst_dev;
    std::vector<int32_t> inst_pos;
    int32_t tlas_depth = 0;
    const uint32_t root = build_tlas(ctx, inst, wbox, comp, inst_dev, inst_pos, tlas_depth);
    if (comp.nodes.size() > sc->host.tlas_cap || root != sc->rec.root)
        return set_error(ctx, S3DG_ERR_INVALID, "top-level tree outgrew its reserved nodes");
    std::vector<GShapeRef> refs(inst.size());
    for (size_t k = 0; k < inst.size(); ++k) {
        refs[k].inst = inst_pos[k];
        refs[k].material = inst[k].material;
    }
    std::vector<GLight> lights = sc->host.lights;
    for (GLight &l : lights)
        if (l.type == LIGHT_SPHERE) l.tri[0] = inst_pos[l.shape];
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    // the context's stream may still be rendering with the old records
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    S3DG_HIP(ctx, hipMemcpy((GNode *)sc->d_nodes + sc->host.tlas_base, comp.nodes.data(),
                            comp.nodes.size() * sizeof(GNode), hipMemcpyHostToDevice));
    S3DG_HIP(ctx, hipMemcpy(sc->d_insts, inst_dev.data(), inst_dev.size() * sizeof(GInst), hipMemcpyHostToDevice));
    S3DG_HIP(ctx, hipMemcpy(sc->d_shape_ref, refs.data(), refs.size() * sizeof(GShapeRef), hipMemcpyHostToDevice));
    if (!lights.empty())
        S3DG_HIP(ctx, hipMemcpy(sc->d_lights, lights.data(), lights.size() * sizeof(GLight), hipMemcpyHostToDevice));
    sc->host.inst = std::move(inst);
    sc->host.wbox = std::move(wbox);
    sc->host.lights = std::move(lights);
    sc->dev.stack_need = std::max(sc->dev.stack_need, 2 * sc->host.blas_depth + 2 * tlas_depth + 4);
    sc->stats.max_depth = sc->dev.stack_need;
    return S3DG_OK;
}
