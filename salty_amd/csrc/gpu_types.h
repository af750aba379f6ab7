// gpu_types.h -- data layout in HBM shared by the host builder and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace s3dg {

// Compressed 8-wide BVH node: one 128-byte line per node.
// Child boxes are quantised to 8 bits per plane relative to `origin` with a
// power-of-two step per axis (2^(exp-127)); quantisation rounds outward so a
// child box always contains the exact child box it stands for.
// Internal children of a node are stored contiguously from child_base in slot
// order; the leaves' triangles (or top-level shape records) are stored
// contiguously from leaf_base in slot order, cnt[s] of them for slot s.
struct alignas(128) GNode {
    float    origin[3];     //   0
    uint8_t  exp[3];        //  12  biased exponent of the quantisation step
    uint8_t  imask;         //  15  bit s set: slot s is an internal child
    uint32_t child_base;    //  16  index of the first internal child; bit 31: top-level node
    uint32_t leaf_base;     //  20  first triangle / shape record of the leaf slots
    uint8_t  cnt[8];        //  24  leaf slot: triangle (or shape record) count
    uint8_t  qlo[3][8];     //  32  quantised lower planes, [axis][slot]
    uint8_t  qhi[3][8];     //  56  quantised upper planes
    uint8_t  pad[48];       //  80
};
constexpr uint32_t kTopLevelBit = 0x80000000u;
static_assert(sizeof(GNode) == 128, "GNode must be one 128-byte line");

// Triangle as Triangle::IsHit uses it: v0 and the two edges computed once on
// the host exactly like the Triangle constructor (36 bytes).
struct GTri {
    float v0[3];
    float e0[3];   // v1 - v0
    float e1[3];   // v2 - v0
};
static_assert(sizeof(GTri) == 36, "GTri is 9 floats");

// Per-triangle shading data, read once per hit.
struct GTriShade {
    float n0[3], n1[3], n2[3];
    int32_t material;
    int32_t input_index;   // index in the mesh's input order
    int32_t pad;
};
static_assert(sizeof(GTriShade) == 48, "GTriShade is 48 bytes");

enum { INST_MESH = 0, INST_TRANSFORMED = 1, INST_SPHERE = 2 };

// Top-level shape record (96 bytes).
struct GInst {
    float    inv[4][4];    // inverse world matrix (row-vector convention)
    float    center[3];    // sphere
    float    radius;
    int32_t  type;         // INST_*
    uint32_t blas_root;    // root node of the mesh tree
    int32_t  shape;        // top-level shape index (tie order)
    int32_t  material;     // sphere material
};
static_assert(sizeof(GInst) == 96, "GInst is 96 bytes");

// Material record (64 bytes).
struct GMat {
    float   diffuse[3];
    float   thr0;
    float   specular[3];
    float   thr1;
    float   emissive[3];
    float   ior;
    float   power;
    int32_t type;
    int32_t delta;
    int32_t texture;        // texture record of a TexturedMaterial, -1: none
};
static_assert(sizeof(GMat) == 64, "GMat is 64 bytes");

// Texture record: RGB texels from `offset` (in floats) of the texel array,
// texel (x, y) at offset + 3 * (width * y + x).
struct GTexture {
    uint64_t offset;
    int32_t  width;
    int32_t  height;
};
static_assert(sizeof(GTexture) == 16, "GTexture is 16 bytes");

// Vertex texture coordinates of a stored triangle (u0 v0 u1 v1 u2 v2), read
// only when its material is textured.
struct GTriUV {
    float uv[6];
};
static_assert(sizeof(GTriUV) == 24, "GTriUV is 24 bytes");

// Per top-level shape: its shape record and the material its hits shade with
// in place of the triangles' own (-1: none; a sphere's own material).
struct GShapeRef {
    int32_t inst;
    int32_t material;
};
static_assert(sizeof(GShapeRef) == 8, "GShapeRef is 8 bytes");

enum { LIGHT_SPHERE = 0, LIGHT_QUAD = 1 };

// Light record (64 bytes), everything its sampling and its NEE term read.
struct GLight {
    float   corner[3];      // QUAD: corner; SPHERE: centre
    int32_t type;           // LIGHT_*
    float   e1[3];          // QUAD: first edge; SPHERE: e1[0] = radius
    int32_t shape;          // top-level shape index
    float   e2[3];          // QUAD: second edge; SPHERE: emission of the sphere's material
    float   pdf;            // QUAD: 1/area; SPHERE: 1/(4 pi r^2)
    int32_t tri[3];         // QUAD: global depth-first indices of its triangles;
                            // SPHERE: tri[0] = instance record of the sphere
    int32_t count;          // QUAD: triangle count (<= 3)
};
static_assert(sizeof(GLight) == 64, "GLight is 64 bytes");

// Camera in the reference's precomputed form (PinholeCamera::Update).
struct GCamera {
    float pos[3];
    int32_t type;
    float cx[3];
    float lens_radius;
    float cy[3];
    float focal_distance;
    float cz[3];
    float pad;
};

}  // namespace s3dg
