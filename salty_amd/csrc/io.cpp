// io.cpp -- host-side readers of the reference's file formats: 24-bit BMP and
// Radiance HDR images (src/src/s3d_texture.cpp) and SMD meshes
// (src/src/s3d_mesh.cpp).  Each follows the reference reader's behaviour,
// quirks included, and turns what would be undefined behaviour there (reads
// past a buffer, a run before the first texel) into an error.
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/s3dg.h"

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(const char *path) : f(path ? std::fopen(path, "rb") : nullptr) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

uint32_t le32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint16_t le16(const uint8_t *p) { return (uint16_t)(p[0] | p[1] << 8); }

float *alloc_rgb(int64_t texels) {
    if (texels <= 0 || texels > (int64_t)1 << 28) return nullptr;
    return (float *)std::calloc((size_t)texels * 3, sizeof(float));
}

// ---------------------------------------------------------------------------
// BMP: LoadFromBMP (src/src/s3d_texture.cpp:1029-1077).  The 14-byte file
// header and the 40-byte info header (MSVC `long` is 32 bits) are read; the
// texels follow at byte 54 whatever OffBits says, width*height*3 bytes with no
// row padding, B, G, R per texel.
// ---------------------------------------------------------------------------
int load_bmp(const char *path, int32_t *w_out, int32_t *h_out, float **rgb_out) {
    File file(path);
    if (!file.f) return S3DG_ERR_INVALID;
    uint8_t hdr[54];
    if (std::fread(hdr, 1, 54, file.f) != 54) return S3DG_ERR_INVALID;
    if (le16(hdr) != 0x4D42u) return S3DG_ERR_INVALID;   // 'BM'
    const int32_t width = (int32_t)le32(hdr + 18), height = (int32_t)le32(hdr + 22);
    const uint16_t bits = le16(hdr + 28);
    if (bits != 24 || width <= 0 || height <= 0) return S3DG_ERR_UNSUPPORTED;
    const int64_t n = (int64_t)width * height;
    float *rgb = alloc_rgb(n);
    if (!rgb) return S3DG_ERR_NOMEM;
    std::vector<uint8_t> bytes((size_t)n * 3, 0);
    // a short file leaves the rest of the texels at zero (the reference
    // leaves them unread)
    (void)std::fread(bytes.data(), 1, bytes.size(), file.f);
    // the reference's loop stops at i < size - 3, so the last texel is never
    // converted; it stays 0 here
    for (int64_t i = 0, j = 0; i < n * 3 - 3; i += 3, ++j) {
        rgb[3 * j + 0] = (float)bytes[(size_t)i + 2] / 255.0f;
        rgb[3 * j + 1] = (float)bytes[(size_t)i + 1] / 255.0f;
        rgb[3 * j + 2] = (float)bytes[(size_t)i + 0] / 255.0f;
    }
    *w_out = width;
    *h_out = height;
    *rgb_out = rgb;
    return S3DG_OK;
}

// ---------------------------------------------------------------------------
// Radiance HDR: ReadHdrHeader / ReadColor / ReadOldColors / ReadHdrData
// (src/src/s3d_texture.cpp:587-850).
// ---------------------------------------------------------------------------
struct Rgbe {
    uint8_t v[4];
};

bool hdr_header(FILE *f, int32_t &width, int32_t &height) {
    char buf[256];
    if (std::fread(buf, 1, 2, f) != 2 || buf[0] != '#' || buf[1] != '?') return false;
    bool valid = false;
    for (;;) {
        if (!std::fgets(buf, sizeof(buf), f)) break;
        if (buf[0] == '\n') break;
        if (buf[0] == '#') continue;
        float g, e;
        if (std::sscanf(buf, "GAMMA=%f\n", &g) != 0) continue;           // parsed, unused
        if (std::sscanf(buf, "EXPOSURE=%f\n", &e) != 0) continue;        // parsed, unused
        if (std::strcmp(buf, "FORMAT=32-bit_rle_rgbe\n") == 0) valid = true;
    }
    if (!valid) return false;
    if (std::fgets(buf, sizeof(buf), f)) {
        int w = 0, h = 0;
        if (std::sscanf(buf, "-Y %d +X %d\n", &h, &w) != 0 || std::sscanf(buf, "+X %d -Y %d\n", &w, &h) != 0) {
            width = w;
            height = h;
        } else {
            return false;
        }
    }
    return true;
}

// Old run-length form: (1, 1, 1, n) repeats the previous texel n << shift times.
bool hdr_old(FILE *f, Rgbe *line, int32_t pos, int32_t count) {
    int shift = 0;
    while (pos < count) {
        Rgbe px;
        for (int k = 0; k < 4; ++k) px.v[k] = (uint8_t)std::getc(f);
        if (std::feof(f) || std::ferror(f)) return false;
        if (px.v[0] == 1 && px.v[1] == 1 && px.v[2] == 1) {
            if (pos == 0) return false;   // nothing to repeat
            // a fifth consecutive repeat code would count in units of 2^32
            // texels, more than any line holds (the reference shifts past
            // the width of its int here)
            if (shift >= 32) return false;
            const int64_t rep = (int64_t)px.v[3] << shift;
            if (pos + rep > count) return false;
            for (int64_t i = 0; i < rep; ++i, ++pos) line[pos] = line[pos - 1];
            shift += 8;
        } else {
            line[pos++] = px;
            shift = 0;
        }
    }
    return true;
}

bool hdr_line(FILE *f, Rgbe *line, int32_t count) {
    if (count < 8 || 0x7fff < count) return hdr_old(f, line, 0, count);
    int i = std::getc(f);
    if (i == EOF) return false;
    if (i != 2) {
        std::ungetc(i, f);
        return hdr_old(f, line, 0, count);
    }
    const int g = std::getc(f), b = std::getc(f);
    if ((i = std::getc(f)) == EOF) return false;
    if (g != 2 || (b & 128)) {
        line[0].v[0] = 2;
        line[0].v[1] = (uint8_t)g;
        line[0].v[2] = (uint8_t)b;
        line[0].v[3] = (uint8_t)i;
        return hdr_old(f, line, 1, count);
    }
    if ((b << 8 | i) != count) return false;
    for (int c = 0; c < 4; ++c) {
        for (int32_t j = 0; j < count;) {
            int code = std::getc(f);
            if (code == EOF) return false;
            if (128 < code) {
                code &= 127;
                const int val = std::getc(f);
                if (j + code > count) return false;
                while (code--) line[j++].v[c] = (uint8_t)val;
            } else {
                if (j + code > count) return false;
                while (code--) line[j++].v[c] = (uint8_t)std::getc(f);
            }
        }
    }
    return !std::feof(f);
}

int load_hdr(const char *path, int32_t *w_out, int32_t *h_out, float **rgb_out) {
    File file(path);
    if (!file.f) return S3DG_ERR_INVALID;
    int32_t width = 0, height = 0;
    if (!hdr_header(file.f, width, height) || width <= 0 || height <= 0) return S3DG_ERR_INVALID;
    float *rgb = alloc_rgb((int64_t)width * height);
    if (!rgb) return S3DG_ERR_NOMEM;
    std::vector<Rgbe> line((size_t)width);
    for (int32_t y = 0; y < height; ++y) {
        if (!hdr_line(file.f, line.data(), width)) {
            std::free(rgb);
            return S3DG_ERR_INVALID;
        }
        for (int32_t x = 0; x < width; ++x) {
            const Rgbe &p = line[(size_t)x];
            float *o = rgb + 3 * ((size_t)y * width + x);
            if (p.v[3]) {
                // RGBEToVec3: channel * 2^(e - 136) in double, then to float
                const double f = std::ldexp(1.0, (int)p.v[3] - (128 + 8));
                o[0] = (float)(p.v[0] * f);
                o[1] = (float)(p.v[1] * f);
                o[2] = (float)(p.v[2] * f);
            }
        }
    }
    *w_out = width;
    *h_out = height;
    *rgb_out = rgb;
    return S3DG_OK;
}

// ---------------------------------------------------------------------------
// SMD: Mesh::LoadFromFile (src/src/s3d_mesh.cpp:250-465) and its structs
// (s3d_mesh.cpp:27-177).
// ---------------------------------------------------------------------------
enum { SMD_MATTE = 0, SMD_MIRROR = 1, SMD_DIELECTRIC = 2, SMD_GLOSSY = 3, SMD_PLASTIC = 4 };

bool read_floats(FILE *f, float *dst, int n) { return std::fread(dst, sizeof(float), (size_t)n, f) == (size_t)n; }
bool read_i32(FILE *f, int32_t &v) { return std::fread(&v, sizeof(v), 1, f) == 1; }

void free_smd(s3dg_smd *m) {
    if (!m) return;
    std::free(m->vertices);
    std::free(m->triangle_material);
    std::free(m->materials);
    std::free(m->texture_paths);
    delete m;
}

int load_smd(const char *path, s3dg_smd **out) {
    if (!path || !out) return S3DG_ERR_INVALID;
    *out = nullptr;
    File file(path);
    if (!file.f) return S3DG_ERR_INVALID;
    FILE *f = file.f;
    // folder of the file: what precedes the last '/' (empty when there is
    // none, which makes the texture paths "/name" as in the reference)
    std::string dir;
    {
        const std::string p(path);
        const size_t k = p.rfind('/');
        if (k != std::string::npos) dir = p.substr(0, k);
    }
    uint8_t hdr[36];
    if (std::fread(hdr, 1, 36, f) != 36) return S3DG_ERR_INVALID;
    if (std::memcmp(hdr, "SMD\0", 4) != 0) return S3DG_ERR_INVALID;
    if (le32(hdr + 4) != 2u) return S3DG_ERR_INVALID;        // SMD_CURRENT_VERSION
    if (le32(hdr + 8) != 24u) return S3DG_ERR_INVALID;       // sizeof(SMD_DATA_HEADER)
    const uint32_t ntri = le32(hdr + 12), nmat = le32(hdr + 16), ntex = le32(hdr + 20);
    if (ntri == 0 || nmat == 0) return S3DG_ERR_INVALID;
    if (ntri > (1u << 26) || nmat > (1u << 20) || ntex > (1u << 16)) return S3DG_ERR_UNSUPPORTED;

    std::unique_ptr<s3dg_smd, void (*)(s3dg_smd *)> m(new s3dg_smd(), free_smd);
    std::memset(m.get(), 0, sizeof(s3dg_smd));
    m->vertex_count = 3 * ntri;
    m->vertices = (s3dg_vertex *)std::calloc((size_t)3 * ntri, sizeof(s3dg_vertex));
    m->triangle_material = (int32_t *)std::calloc(ntri, sizeof(int32_t));
    m->material_count = nmat;
    m->materials = (s3dg_material *)std::calloc(nmat, sizeof(s3dg_material));
    m->texture_count = ntex;
    m->texture_paths = ntex ? (char(*)[512])std::calloc(ntex, 512) : nullptr;
    if (!m->vertices || !m->triangle_material || !m->materials || (ntex && !m->texture_paths))
        return S3DG_ERR_NOMEM;

    for (uint32_t i = 0; i < ntex; ++i) {
        char name[257];
        if (std::fread(name, 1, 256, f) != 256) return S3DG_ERR_INVALID;
        name[256] = 0;
        // the reference runs on Windows, where '\\' in a stored name is a
        // folder separator too
        for (char *c = name; *c; ++c)
            if (*c == '\\') *c = '/';
        const std::string p = dir + "/" + name;
        std::snprintf(m->texture_paths[i], 512, "%s", p.c_str());
    }
    for (uint32_t i = 0; i < nmat; ++i) {
        int32_t type = -1;
        if (!read_i32(f, type)) return S3DG_ERR_INVALID;
        s3dg_material &o = m->materials[i];
        float c[3], e[3], x;
        int32_t map = -1, map2 = -1;
        switch (type) {
        case SMD_MATTE:        // Color, Emissive, ColorMap -> CreateLambert
        case SMD_MIRROR:       // Color, Emissive, ColorMap -> CreateMirror
            if (!read_floats(f, c, 3) || !read_floats(f, e, 3) || !read_i32(f, map)) return S3DG_ERR_INVALID;
            o.type = type == SMD_MATTE ? S3DG_MAT_LAMBERT : S3DG_MAT_MIRROR;
            std::memcpy(type == SMD_MATTE ? o.diffuse : o.specular, c, sizeof(c));
            break;
        case SMD_DIELECTRIC:   // Color, Ior, Emissive, ColorMap -> CreateGlass
        case SMD_GLOSSY:       // Color, Power, Emissive, ColorMap -> CreatePhong
            if (!read_floats(f, c, 3) || !read_floats(f, &x, 1) || !read_floats(f, e, 3) || !read_i32(f, map))
                return S3DG_ERR_INVALID;
            o.type = type == SMD_DIELECTRIC ? S3DG_MAT_GLASS : S3DG_MAT_PHONG;
            std::memcpy(o.specular, c, sizeof(c));
            if (type == SMD_DIELECTRIC) o.ior = x;
            else o.power = x;
            break;
        case SMD_PLASTIC: {    // Diffuse, Specular, Power, Emissive, DiffuseMap, SpecularMap -> CreatePlastic
            float sp[3];
            if (!read_floats(f, c, 3) || !read_floats(f, sp, 3) || !read_floats(f, &x, 1) || !read_floats(f, e, 3) ||
                !read_i32(f, map) || !read_i32(f, map2))
                return S3DG_ERR_INVALID;
            o.type = S3DG_MAT_PLASTIC;
            std::memcpy(o.diffuse, c, sizeof(c));
            std::memcpy(o.specular, sp, sizeof(sp));
            o.power = x;
            break;   // the SpecularMap is read and not used, as there
        }
        default:
            return S3DG_ERR_UNSUPPORTED;   // the reference asserts
        }
        std::memcpy(o.emissive, e, sizeof(e));
        if (map >= 0 && (uint32_t)map >= ntex) return S3DG_ERR_INVALID;
        o.texture = map >= 0 ? map + 1 : 0;
    }
    for (uint32_t i = 0; i < ntri; ++i) {
        float v[24];
        int32_t mat;
        if (!read_floats(f, v, 24) || !read_i32(f, mat)) return S3DG_ERR_INVALID;
        for (int k = 0; k < 3; ++k) {
            s3dg_vertex &d = m->vertices[3 * i + k];
            std::memcpy(d.position, v + 8 * k, 12);
            std::memcpy(d.normal, v + 8 * k + 3, 12);
            std::memcpy(d.texcoord, v + 8 * k + 6, 8);
        }
        // the reference indexes its material table with no bound check
        if (mat >= (int32_t)nmat) return S3DG_ERR_INVALID;
        m->triangle_material[i] = mat < 0 ? -1 : mat;
    }
    *out = m.release();
    return S3DG_OK;
}

}  // namespace

extern "C" int s3dg_image_load_bmp(const char *path, int32_t *width, int32_t *height, float **rgb) {
    if (!path || !width || !height || !rgb) return S3DG_ERR_INVALID;
    return load_bmp(path, width, height, rgb);
}

extern "C" int s3dg_image_load_hdr(const char *path, int32_t *width, int32_t *height, float **rgb) {
    if (!path || !width || !height || !rgb) return S3DG_ERR_INVALID;
    return load_hdr(path, width, height, rgb);
}

extern "C" void s3dg_image_free(float *rgb) { std::free(rgb); }

extern "C" int s3dg_smd_load(const char *path, s3dg_smd **out) { return load_smd(path, out); }

extern "C" void s3dg_smd_free(s3dg_smd *smd) { free_smd(smd); }
