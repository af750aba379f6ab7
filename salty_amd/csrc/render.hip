// render.hip -- wavefront path tracer: PathTracer::TracePath / Radiance /
// NextEventEstimation (src/src/s3d_pt.cpp:182-353) as a pipeline of gfx950
// kernels over device-resident ray queues.
//
// One pass adds one pixel-centre sample per pixel (like one iteration of the
// reference's progressive loop).  Several passes are traced together: path
// slot p holds the (p % npix)-th pixel, in the band order of slot_pixel, of
// pass p / npix, so the deep bounces, where
// Russian roulette has thinned the queues, still fill the GPU.  Per bounce:
// extend (closest hit of the queued rays), shade (emission, light sample,
// material sample, Russian roulette, compaction into the next queue and the
// shadow queue) and shadow (visibility of the light sample, then the NEE term).
// The order in which a path adds emission, NEE and miss radiance to L is the
// reference's, and the passes are added to the accumulator in pass order, so
// the sum is the same whatever the number of passes in flight.  Renders of
// few paths run the same per-path code in one launch instead (k_paths: one
// thread carries a path through all its bounces), with the same sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "internal.h"
#include "shading.h"
#include "traverse.h"

using namespace s3dg;


namespace {

// Path state over N path slots.  What lives from bounce to bounce is one
// 64-byte record per path slot (a gather after compaction then touches one
// line per path, not one per field):
//   path[4p]     ray origin x y z, direction x
//   path[4p + 1] direction y z, throughput W r g
//   path[4p + 2] W b, radiance L r g b
//   path[4p + 3] PCG state (low, high word), scene, unused
// The first bounce's records are not stored before that bounce: its extend
// and shade launches compute them from the slot (camera_record), and the
// shade launch stores them, whole for a path that goes on, L (and W b) only
// for one that ends there.
// What one bounce hands to the next kernel is kept in the order that kernel
// reads it: the hit of queue entry i of the extension launch at hit[i], the
// shadow ray of the path the shade launch visited i-th at shadow[3 i ...].
struct PathBuf {
    float4 *path;                           // path records
    float4 *hit;                            // extension hit by queue position: t, u, v, prim (bits)
    int32_t *hit_shape;                     // top-level shape of that hit
    float4 *shadow;                         // shadow rays by shade position, three float4 each:
                                            // origin + distance to the light (< 0: light missed),
                                            // direction + path slot (bits), W * NEE (added to L
                                            // when the light is visible) + scene (bits)
    int32_t single_scene;                   // >= 0: every path is in this scene (record scene unused)
    uint32_t *q[2];                         // extension queues (ping-pong)
    uint32_t *qs;                           // shadow queue (shade positions)
    uint32_t *cnt;                          // [0],[1] extension counts, [2] shadow count,
                                            // [4],[5] extend / shadow work-fetch cursors,
                                            // [8] a tail launch has finished the paths,
                                            // [9] the launches after it have seen that
    unsigned long long *tot;                // [0] extension rays, [1] shadow rays, [2..5] node visits
                                            // and triangle tests (extend, shadow), [6], [7]
                                            // stack pushes to global memory (extend, shadow)
};

// bytes per path slot: the 64-byte record, a 16-byte hit and its shape, a
// 48-byte shadow record and three queues
constexpr size_t kPathBytes = 64 + 16 + 4 + 48 + 3 * 4;

size_t path_bytes(uint32_t n) { return (size_t)n * kPathBytes + 4096; }

PathBuf carve(void *base, uint32_t n) {
    char *p = (char *)base;
    auto take = [&](size_t bytes) {
        char *r = p;
        p += (bytes + 255) & ~(size_t)255;
        return r;
    };
    PathBuf b;
    b.cnt = (uint32_t *)take(64);
    b.tot = (unsigned long long *)take(64);
    b.path = (float4 *)take((size_t)n * 64);
    b.hit = (float4 *)take((size_t)n * 16);
    b.hit_shape = (int32_t *)take((size_t)n * 4);
    b.shadow = (float4 *)take((size_t)n * 48);
    b.q[0] = (uint32_t *)take((size_t)n * 4);
    b.q[1] = (uint32_t *)take((size_t)n * 4);
    b.qs = (uint32_t *)take((size_t)n * 4);
    b.single_scene = -1;
    return b;
}

// HIP events owned for the duration of one call.
struct EventSet {
    std::vector<hipEvent_t> ev;
    hipError_t create(size_t n) {
        ev.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            hipEvent_t e = nullptr;
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            ev.push_back(e);
        }
        return hipSuccess;
    }
    ~EventSet() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// A shadow ray as k_shade hands it to k_shadow.
struct ShadowRay {
    V3 o, d;
    float tl;    // distance to the light along d (< 0: the light was missed)
    V3 c;        // W * NEE term, added to L when nothing is hit before the light
    uint32_t scene;
};

__device__ __forceinline__ uint32_t scene_of(const PathBuf &P, uint32_t p) {
    return P.single_scene >= 0 ? (uint32_t)P.single_scene : __float_as_uint(P.path[4 * (size_t)p + 3].z);
}

// The ray of path p.
__device__ __forceinline__ DRay path_ray(const PathBuf &P, uint32_t p) {
    const float4 a = P.path[4 * (size_t)p], b = P.path[4 * (size_t)p + 1];
    DRay r;
    r.ox = a.x; r.oy = a.y; r.oz = a.z;
    r.dx = a.w; r.dy = b.x; r.dz = b.y;
    return r;
}

// L += c for path p (the W b word of the same float4 is kept).
__device__ __forceinline__ void path_add_l(const PathBuf &P, uint32_t p, V3 c) {
    float4 *w = P.path + 4 * (size_t)p + 2;
    const float4 v = *w;
    *w = make_float4(v.x, v.y + c.x, v.z + c.y, v.w + c.z);
}
size_t carve_bytes(uint32_t n) { return path_bytes(n) + 40 * 256; }

// A batch of paths is a list of segments: consecutive passes of one frame
// (a scene, its resolution, seed and accumulation buffer).  Path slot p of
// segment g is pixel slot_pixel((p - path0) % npix) of pass
// pass0 + (p - path0) / npix.
struct Seg {
    uint32_t path0;     // first path slot
    uint32_t npix;      // width * height
    uint32_t pix0;      // first accumulation thread (prefix of npix over the batch)
    uint32_t scene;     // index into SceneDev::recs
    int32_t width, height;
    uint32_t pass0, passes;
    uint64_t seed;
    float *accum;       // RGB sum buffer of the frame
};

struct BatchDesc {
    uint32_t nseg, paths, pixels, pad;
};

// Pixel (row-major index) of the j-th slot of a pass: the image is cut into
// bands of kBand rows and each band is walked column by column, so the 64
// paths of a wave start in a kBand x 8 block of pixels and their primary and
// first shadow rays stay coherent.  A bijection on [0, W*H) for any W, H.
constexpr uint32_t kBand = 8;
__device__ __forceinline__ uint32_t slot_pixel(uint32_t j, uint32_t W, uint32_t H) {
    const uint32_t band = j / (kBand * W);
    const uint32_t y0 = band * kBand;
    const uint32_t bh = min(kBand, H - y0);
    const uint32_t r = j - y0 * W;
    return (y0 + r % bh) * W + r / bh;
}

// Segment holding path slot (or accumulation thread) i: binary search on
// path0 (or pix0); segments are sorted by both.
template <bool BY_PIX>
__device__ __forceinline__ const Seg &find_seg(const Seg *__restrict__ segs, uint32_t nseg, uint32_t i) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t start = BY_PIX ? segs[mid].pix0 : segs[mid].path0;
        if (start <= i) lo = mid;
        else hi = mid;
    }
    return segs[lo];
}

// Paths traced together when the caller leaves passes_in_flight at 0, and the
// most ever held at once: 2^30 paths are 155 GB of path state, so that a
// batch of many frames runs in few, large batches (the 64-scene
// configuration's 1.93 G paths in two; each batch ends with a thin tail of
// deep bounces).  The automatic choice also stays below half of the free
// device memory.
constexpr uint32_t kAutoPathsInFlight = 1u << 30;
constexpr uint32_t kMaxPathsInFlight = 1u << 30;

// Appends of a workgroup to two queues, staged in LDS over several loop
// iterations: a wave reserves its entries in the stage with one LDS atomic,
// and the workgroup reserves the global space of a full stage with one global
// atomic per queue and writes it out contiguously.  The few queue counters
// every workgroup of the grid appends to are thus hit once per ~kStage
// entries (the hardware serialises atomics on one address), not once per
// wave or per iteration.  Every thread of the workgroup must make the same
// calls (they synchronise the workgroup).
struct StagedAppend2 {
    static constexpr uint32_t kStage = 2048;   // entries per queue
    uint32_t buf[2][kStage];
    uint32_t count[2];
    uint32_t base[2];

    __device__ __forceinline__ void init() {
        if (threadIdx.x < 2) count[threadIdx.x] = 0;
        __syncthreads();
    }
    // Empties the stage when it could not take another iteration (all = true:
    // whatever it holds).
    __device__ __forceinline__ void flush_if_full(uint32_t *c0, uint32_t *q0, uint32_t *c1, uint32_t *q1, bool all) {
        __syncthreads();
        const uint32_t n0 = count[0], n1 = count[1];
        if (all ? (n0 | n1) != 0u : (n0 > kStage - kBlock || n1 > kStage - kBlock)) {
            if (threadIdx.x == 0) {
                base[0] = n0 ? atomicAdd(c0, n0) : 0u;
                base[1] = n1 ? atomicAdd(c1, n1) : 0u;
            }
            __syncthreads();
            const uint32_t b0 = base[0], b1 = base[1];
            for (uint32_t k = threadIdx.x; k < n0; k += kBlock) q0[b0 + k] = buf[0][k];
            for (uint32_t k = threadIdx.x; k < n1; k += kBlock) q1[b1 + k] = buf[1][k];
            __syncthreads();
            if (threadIdx.x < 2) count[threadIdx.x] = 0;
        }
        __syncthreads();
    }
    // Stages v0 in queue 0 if w0 and v1 in queue 1 if w1 (at most kBlock
    // entries per queue between two flush_if_full calls).
    __device__ __forceinline__ void add(bool w0, uint32_t v0, bool w1, uint32_t v1) {
        const uint32_t lane = __lane_id();
        const uint64_t m0 = __ballot(w0), m1 = __ballot(w1);
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t s0 = 0, s1 = 0;
        if (lane == 0) {
            if (m0) s0 = atomicAdd(&count[0], (uint32_t)__popcll(m0));
            if (m1) s1 = atomicAdd(&count[1], (uint32_t)__popcll(m1));
        }
        s0 = __shfl(s0, 0) + (uint32_t)__popcll(m0 & below);
        s1 = __shfl(s1, 0) + (uint32_t)__popcll(m1 & below);
        if (w0) buf[0][s0] = v0;
        if (w1) buf[1][s1] = v1;
    }
};

// Persistent-kernel work fetch: lane 0 of the wave takes the next 64 queue
// entries; the wave leaves its loop once the cursor has passed the queue end.
__device__ __forceinline__ uint32_t wave_fetch(uint32_t *cursor) {
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(cursor, 64u);
    return __shfl(base, 0);
}

// Radiance of a ray leaving the scene: Scene::SampleIBL (s3d_scene.h:80-82),
// the lat-long texture times 10, or the scene's constant environment.
__device__ __forceinline__ V3 environment(const SceneDev &S, const SceneRec &rec, V3 dir) {
    if (rec.env_tex < 0) return ld3(rec.env);
    return tex_sample_dir(S.texels, S.textures[rec.env_tex], dir) * 10.0f;
}

// The four float4 words of a path record (see PathBuf::path).
struct PathRec {
    float4 a, b, c, d;
};

__device__ __forceinline__ PathRec load_rec(const PathBuf &P, uint32_t p) {
    const float4 *r = P.path + 4 * (size_t)p;
    return PathRec{r[0], r[1], r[2], r[3]};
}

// Path slot p as ray generation leaves it (TracePath, s3d_pt.cpp:326-328):
// the camera ray, W = 1, L = 0, the RNG after the camera's draws and the
// scene.  The pipeline's first bounce computes it where it is needed (the
// extension and the shading launch each, bit for bit alike) instead of
// reading it back from HBM.
__device__ __forceinline__ PathRec camera_record(const SceneDev &S, uint32_t p, uint32_t nseg,
                                                 const Seg *__restrict__ segs) {
    const Seg &g = find_seg<false>(segs, nseg, p);
    const uint32_t local = p - g.path0;
    const uint32_t pix = slot_pixel(local % g.npix, (uint32_t)g.width, (uint32_t)g.height);
    const uint32_t pass = g.pass0 + local / g.npix;
    const uint32_t x = pix % (uint32_t)g.width, y = pix / (uint32_t)g.width;
    Pcg rng;
    rng.seed(path_seed(g.seed, pix, pass, 0u));
    // TracePath pixel-centre coordinates (s3d_pt.cpp:326-328)
    const float fx = (0.5f + (float)x) / (float)g.width - 0.5f;
    const float fy = (0.5f + (float)y) / (float)g.height - 0.5f;
    V3 o, d;
    camera_ray(S.recs[g.scene].cam, fx, fy, rng, o, d);
    return PathRec{make_float4(o.x, o.y, o.z, d.x), make_float4(d.y, d.z, 1.0f, 1.0f),
                   make_float4(1.0f, 0.0f, 0.0f, 0.0f),
                   make_float4(__uint_as_float((uint32_t)rng.state), __uint_as_float((uint32_t)(rng.state >> 32)),
                               __uint_as_float(g.scene), 0.0f)};
}

__device__ __forceinline__ DRay rec_ray(const PathRec &r) {
    DRay d;
    d.ox = r.a.x; d.oy = r.a.y; d.oz = r.a.z;
    d.dx = r.a.w; d.dy = r.b.x; d.dz = r.b.y;
    return d;
}

__device__ __forceinline__ uint32_t rec_scene(const PathBuf &P, const PathRec &r) {
    return P.single_scene >= 0 ? (uint32_t)P.single_scene : __float_as_uint(r.d.z);
}

// Renders of no bounce: every path's radiance word as ray generation leaves
// it (W b = 1, L = 0), for the accumulation.
__global__ __launch_bounds__(kBlock) void k_no_bounce(PathBuf P, const BatchDesc *__restrict__ bd) {
    const uint32_t npaths = bd->paths;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npaths; p += gridDim.x * blockDim.x)
        P.path[4 * (size_t)p + 2] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
}

// Counters of a batch's first bounce: the extension queue is the identity
// over the batch's path slots (implicit: the first bounce reads no queue).
__global__ void k_start_batch(PathBuf P, const BatchDesc *__restrict__ bd) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        P.cnt[0] = bd->paths;
        P.cnt[1] = 0;
        P.cnt[2] = 0;
        P.cnt[4] = 0;
        P.cnt[5] = 0;
        P.cnt[8] = 0;
        P.cnt[9] = 0;
    }
}

template <bool COUNT, int LDS, int WAVES>
__global__ __attribute__((amdgpu_flat_work_group_size(kBlock, kBlock), amdgpu_waves_per_eu(WAVES, WAVES))) void
k_extend(SceneDev S, PathBuf P, int32_t cur, uint2 *gstack, int32_t gcap, int32_t *err, FetchParams fetch,
         int32_t first, const BatchDesc *__restrict__ bd, const Seg *__restrict__ segs) {
    __shared__ uint2 lds[LDS * kBlock];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    if (P.cnt[8]) {
        // a tail launch has finished every path: tell the later tails
        if (gtid == 0) P.cnt[9] = 1;
        return;
    }
    const uint32_t n = P.cnt[cur];
    if (gtid == 0) {
        P.cnt[cur ^ 1] = 0;
        P.cnt[2] = 0;
        P.cnt[5] = 0;   // the shadow launch of this bounce fetches from zero
        P.tot[0] += n;
    }
    Stack<LDS> st;
    st.init(lds, threadIdx.x, gstack + gtid, stride, gcap, err);
    const uint32_t *q = P.q[cur];
    const uint32_t nseg = first ? bd->nseg : 0u;
    Tracer<false, COUNT> tr;
    tr.reset_counts();
    uint32_t qi = 0;
    refill_loop(
        &P.cnt[4], n, fetch,
        [&](uint32_t i) {
            qi = i;
            st.sp = 0;
            if (first) {
                // the first bounce: slot i's camera ray, computed here
                const PathRec pr = camera_record(S, i, nseg, segs);
                tr.start(S.recs[rec_scene(P, pr)], rec_ray(pr), kRefHitMin, kRefHitMax);
            } else {
                const uint32_t p = q[i];
                tr.start(S.recs[scene_of(P, p)], path_ray(P, p), kRefHitMin, kRefHitMax);
            }
            return true;
        },
        [&]() { return tr.step(S, st); },
        [&]() {
            // by queue position: the lanes of a wave hold nearby entries
            P.hit[qi] = make_float4(tr.best.t, tr.best.u, tr.best.v, __int_as_float(tr.best.prim));
            P.hit_shape[qi] = tr.best.shape;
        });
    tr.flush(P.tot + 2);
    if (COUNT) st.flush_spills(P.tot + 6);
}

// One path vertex (Radiance, s3d_pt.cpp:182-246, with the sampling half of
// NextEventEstimation) at the closest hit h of the path's ray: emission or
// environment, the light sample and its NEE term (returned as the shadow ray
// sr), the material sample and Russian roulette.
// shadow: sr is to be traced; cont: the path goes on.
// pr: the path's record as it stands (fresh: as ray generation leaves it and
// not yet stored, so that L is stored whatever the vertex adds).
__device__ __forceinline__ void shade_path(const SceneDev &S, const PathBuf &P, uint32_t p, const PathRec &pr,
                                           bool fresh, const DHit &h, int32_t depth, int32_t rr_start, bool &shadow,
                                           ShadowRay &sr, bool &cont) {
    float4 *rp = P.path + 4 * (size_t)p;
    const float4 ra = pr.a, rb = pr.b, rc = pr.c, rd = pr.d;
    V3 W = v3(rb.z, rb.w, rc.x);
    const V3 o = v3(ra.x, ra.y, ra.z);
    const V3 d = v3(ra.w, rb.x, rb.y);
    const uint32_t scene = P.single_scene >= 0 ? (uint32_t)P.single_scene : __float_as_uint(rd.z);
    V3 Ladd;   // what this vertex adds to the path's radiance L
    const int32_t prim = h.prim;
    if (prim == -1) {
        // miss: L += W * SampleIBL(dir) (s3d_pt.cpp:199-203), or the
        // constant environment when the scene has no IBL texture
        Ladd = W * environment(S, S.recs[scene], d);
    } else {
        const float t = h.t;
        const V3 pos = o + d * t;
        V3 normal;
        const GShapeRef sref = S.shape_ref[h.shape];
        int32_t mat = sref.material;
        if (prim == -2) {
            normal = unit3(pos - ld3(S.insts[sref.inst].center));
        } else {
            const GTriShade &sh = S.shade[prim];
            normal = tri_normal(sh, h.u, h.v);
            if (mat < 0) mat = sh.material;
        }
        const GMat &m = S.mats[mat];
        Ladd = W * ld3(m.emissive);
        // TexturedMaterial: Shade and GetBaseColor times the texture
        // at the hit's texcoord (1 when untextured, an exact no-op)
        V3 tcol = v3(1.0f, 1.0f, 1.0f);
        if (m.texture >= 0) {
            float tu, tv;
            if (prim == -2) sphere_texcoord(normal, tu, tv);
            else tri_texcoord(S.tri_uv[prim], h.u, h.v, tu, tv);
            tcol = tex_sample(S.texels, S.textures[m.texture], tu, tv);
        }
        Pcg rng;
        rng.state = (uint64_t)__float_as_uint(rd.x) | ((uint64_t)__float_as_uint(rd.y) << 32);
        const SceneRec &rec = S.recs[scene];
        if (!m.delta && rec.light_count > 0) {
            // NextEventEstimation sampling half (s3d_pt.cpp:258-268)
            const uint32_t li = (uint32_t)rec.light_base + rng.u32() % (uint32_t)rec.light_count;
            const GLight &lt = S.lights[li];
            V3 lpos;
            if (lt.type == LIGHT_SPHERE) {
                lpos = sphere_sample_point(ld3(lt.corner), lt.e1[0], rng);
            } else {
                const float u = rng.f32();
                const float v = rng.f32();
                lpos = ld3(lt.corner) + ld3(lt.e1) * u + ld3(lt.e2) * v;
            }
            V3 ldir = lpos - pos;
            const float d2 = dot3(ldir, ldir);
            ldir = safe_unit3(ldir);
            // Where the shadow ray meets the sampled light, by that
            // shape's own test, and the NextEventEstimation term
            // (s3d_pt.cpp:274-289) it adds if nothing is hit before.
            DRay r;
            r.ox = pos.x; r.oy = pos.y; r.oz = pos.z;
            r.dx = ldir.x; r.dy = ldir.y; r.dz = ldir.z;
            float tl = -1.0f, lu = 0.0f, lv = 0.0f;
            int32_t ltri = -1;
            if (lt.type == LIGHT_SPHERE) {
                tl = sphere_test(lt.corner[0], lt.corner[1], lt.corner[2], lt.e1[0], r, kRefHitMin);
            } else {
                float best = kFMaxF;
                for (int k = 0; k < lt.count; ++k) {
                    float t, u, v;
                    if (tri_test(S.tris, (uint32_t)lt.tri[k], r, kRefHitMin, kRefHitMax, t, u, v) &&
                        t < best) {
                        best = t; lu = u; lv = v; ltri = lt.tri[k];
                    }
                }
                if (ltri >= 0) tl = best;
            }
            V3 c = v3(0.0f, 0.0f, 0.0f);
            if (tl >= 0.0f) {
                const V3 ph = pos + ldir * tl;
                V3 ln, le;
                if (lt.type == LIGHT_SPHERE) {
                    ln = unit3(ph - ld3(lt.corner));
                    le = ld3(lt.e2);
                } else {
                    ln = tri_normal(S.shade[ltri], lu, lv);
                    le = ld3(S.mats[S.shade[ltri].material].emissive);
                }
                const V3 fs = (base_color(m) * tcol) / kPi;
                const float cos_shadow = fabsf(dot3(normal, ldir));
                const float cos_light = fabsf(dot3(ln, -ldir));
                const float G = (cos_shadow * cos_light) / d2;
                const float brdf_pdf = fabsf(ldir.z / kPi) * cos_light / d2;
                const float mis = lt.pdf / (brdf_pdf + lt.pdf);
                const V3 nee = le * fs * mis * (G / lt.pdf);
                c = W * nee;
            }
            sr.o = pos;
            sr.d = ldir;
            sr.tl = tl;
            sr.c = c;
            sr.scene = scene;
            shadow = true;
        }
        V3 out;
        bool dice = false;
        W = W * (shade(m, d, normal, rng, out, dice) * tcol);
        const bool stop = (dice && depth >= rr_start) ||
                          (W.x < FLT_EPSILON && W.y < FLT_EPSILON && W.z < FLT_EPSILON);
        if (!stop) {
            // the next ray, W and the RNG (the state of a finished path is
            // never read again); L follows below
            rp[0] = make_float4(pos.x, pos.y, pos.z, out.x);
            rp[1] = make_float4(out.y, out.z, W.x, W.y);
            rp[3] = make_float4(__uint_as_float((uint32_t)rng.state), __uint_as_float((uint32_t)(rng.state >> 32)),
                                rd.z, rd.w);
            cont = true;
        }
    }
    // L += Ladd, skipped when Ladd is zero in every channel: L starts
    // at +0 and a sum is -0 only when both terms are, so adding a zero
    // leaves L's bits as they are (a NaN compares unequal and is added).
    // The float4 holding L also holds W b, written when the path goes on.
    const bool add = Ladd.x != 0.0f || Ladd.y != 0.0f || Ladd.z != 0.0f;
    if (add || cont || fresh) {
        const V3 L = add ? v3(rc.y + Ladd.x, rc.z + Ladd.y, rc.w + Ladd.z) : v3(rc.y, rc.z, rc.w);
        rp[2] = make_float4(cont ? W.z : rc.x, L.x, L.y, L.z);
    }
}

__global__ __launch_bounds__(kBlock) void k_shade(SceneDev S, PathBuf P, int32_t cur, int32_t depth,
                                                   int32_t rr_start, const BatchDesc *__restrict__ bd,
                                                   const Seg *__restrict__ segs) {
    if (P.cnt[8]) return;   // a tail launch has finished every path
    const uint32_t n = P.cnt[cur];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t *q = P.q[cur];
    __shared__ StagedAppend2 stage;
    stage.init();
    if (gtid == 0) P.cnt[4] = 0;   // the next extend launch fetches from zero
    // the first bounce's queue is the identity and its records are computed here
    const bool first = depth == 0;
    const uint32_t nseg = first ? bd->nseg : 0u;
    const uint32_t trips = (n + stride - 1) / stride;
    for (uint32_t it = 0; it < trips; ++it) {
        const uint32_t i = gtid + it * stride;
        const bool active = i < n;
        bool cont = false, shadow = false;
        uint32_t p = 0;
        if (active) {
            p = first ? i : q[i];
            const PathRec pr = first ? camera_record(S, p, nseg, segs) : load_rec(P, p);
            const float4 hv = P.hit[i];
            DHit h;
            h.t = hv.x;
            h.u = hv.y;
            h.v = hv.z;
            h.prim = __float_as_int(hv.w);
            h.shape = P.hit_shape[i];
            ShadowRay sr;
            shade_path(S, P, p, pr, first, h, depth, rr_start, shadow, sr, cont);
            if (shadow) {
                float4 *rec = P.shadow + 3 * (size_t)i;
                rec[0] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tl);
                rec[1] = make_float4(sr.d.x, sr.d.y, sr.d.z, __uint_as_float(p));
                rec[2] = make_float4(sr.c.x, sr.c.y, sr.c.z, __uint_as_float(sr.scene));
            }
        }
        stage.flush_if_full(&P.cnt[2], P.qs, &P.cnt[cur ^ 1], P.q[cur ^ 1], false);
        // the shadow queue takes the record's position, the extension queue the path
        stage.add(shadow, i, cont, p);
    }
    stage.flush_if_full(&P.cnt[2], P.qs, &P.cnt[cur ^ 1], P.q[cur ^ 1], true);
}

template <bool COUNT, int LDS, int WAVES>
__global__ __attribute__((amdgpu_flat_work_group_size(kBlock, kBlock), amdgpu_waves_per_eu(WAVES, WAVES))) void
k_shadow(SceneDev S, PathBuf P, uint2 *gstack, int32_t gcap, int32_t *err, FetchParams fetch) {
    __shared__ uint2 lds[LDS * kBlock];
    if (P.cnt[8]) return;   // a tail launch has finished every path
    const uint32_t n = P.cnt[2];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gtid == 0) P.tot[1] += n;
    Stack<LDS> st;
    st.init(lds, threadIdx.x, gstack + gtid, stride, gcap, err);
    Tracer<true, COUNT> tr;
    tr.reset_counts();
    uint32_t p = 0, j = 0;
    refill_loop(
        &P.cnt[5], n, fetch,
        [&](uint32_t i) {
            j = P.qs[i];
            const float4 a = P.shadow[3 * (size_t)j];
            // the light was missed (no NEE term): nothing to trace
            const float tl = a.w;
            if (!(tl >= 0.0f)) return false;
            const float4 b = P.shadow[3 * (size_t)j + 1];
            p = __float_as_uint(b.w);
            DRay r;
            r.ox = a.x; r.oy = a.y; r.oz = a.z;
            r.dx = b.x; r.dy = b.y; r.dz = b.z;
            // the scene from the shadow record: a path that ended at its first
            // vertex has no record of its own in memory
            const uint32_t scene =
                P.single_scene >= 0 ? (uint32_t)P.single_scene : __float_as_uint(P.shadow[3 * (size_t)j + 2].w);
            // "closest hit is the light" == no other hit strictly before it
            st.sp = 0;
            tr.start(S.recs[scene], r, kRefHitMin, nextafterf(tl, 0.0f));
            return true;
        },
        [&]() { return tr.step(S, st); },
        [&]() {
            if (tr.best.prim != -1) return;
            const float4 c = P.shadow[3 * (size_t)j + 2];
            path_add_l(P, p, v3(c.x, c.y, c.z));
        });
    tr.flush(P.tot + 4);
    if (COUNT) st.flush_spills(P.tot + 7);
}

// Renders of few paths: each thread carries its paths (grid-stride) through
// all of their bounces in a single launch (the camera ray, then per bounce the
// closest hit, shade_path and the shadow ray), the per-path arithmetic of the
// wavefront kernels in the same order, so the sums are the same bit for bit.
// A small render then costs one launch instead of three per bounce; with many
// paths the wavefront pipeline's coherent, compacted launches win.
template <int LDS>
__global__ __launch_bounds__(kBlock) void k_paths(SceneDev S, PathBuf P, const BatchDesc *__restrict__ bd,
                                                  const Seg *__restrict__ segs, int32_t max_bounce,
                                                  int32_t rr_start, uint2 *gstack, int32_t gcap, int32_t *err) {
    __shared__ uint2 lds[LDS * kBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack<LDS> st;
    st.init(lds, threadIdx.x, gstack + gtid, stride, gcap, err);
    unsigned long long ext = 0, shadows = 0;
    const uint32_t npaths = bd->paths;
    for (uint32_t p = gtid; p < npaths; p += stride) {
        PathRec pr = camera_record(S, p, bd->nseg, segs);
        if (max_bounce <= 0) P.path[4 * (size_t)p + 2] = pr.c;   // no vertex: L stays 0
        const SceneRec &rec = S.recs[rec_scene(P, pr)];
        for (int32_t depth = 0; depth < max_bounce; ++depth) {
            if (depth > 0) pr = load_rec(P, p);
            const DHit h = traverse<false, false>(S, rec, rec_ray(pr), kRefHitMin, kRefHitMax, st, nullptr);
            ++ext;
            bool shadow = false, cont = false;
            ShadowRay sh;
            shade_path(S, P, p, pr, depth == 0, h, depth, rr_start, shadow, sh, cont);
            if (shadow) {
                ++shadows;
                if (sh.tl >= 0.0f) {
                    DRay sr;
                    sr.ox = sh.o.x; sr.oy = sh.o.y; sr.oz = sh.o.z;
                    sr.dx = sh.d.x; sr.dy = sh.d.y; sr.dz = sh.d.z;
                    // "closest hit is the light" == no other hit strictly before it
                    const DHit hs = traverse<true, false>(S, rec, sr, kRefHitMin, nextafterf(sh.tl, 0.0f), st, nullptr);
                    if (hs.prim == -1) path_add_l(P, p, sh.c);
                }
            }
            if (!cont) break;
        }
    }
    // ray counts of the render statistics: one atomic per wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ext += __shfl_xor(ext, off);
        shadows += __shfl_xor(shadows, off);
    }
    if (__lane_id() == 0) {
        if (ext) atomicAdd(&P.tot[0], ext);
        if (shadows) atomicAdd(&P.tot[1], shadows);
    }
}

// The last bounces of the wavefront: once the extension queue holds at most
// `limit` paths (at most one per thread of the grid), one thread carries each
// of them through every remaining bounce, depth0 .. max_bounce - 1, with the
// per-path code of k_paths, instead of three launches per bounce that each
// wait on the slowest of a few rays.  Same arithmetic per path in the same
// order, so the sums and ray counts are those of the pipeline.  Launched after
// every bounce's shadow launch; the first one whose queue is short enough
// runs, sets cnt[8] so that the later bounce launches return at once, and the
// later tails see cnt[9], which the first skipped extend launch sets.
template <int LDS>
__global__ __launch_bounds__(kBlock) void k_tail(SceneDev S, PathBuf P, int32_t cur, int32_t depth0,
                                                 int32_t max_bounce, int32_t rr_start, uint32_t limit,
                                                 uint2 *gstack, int32_t gcap, int32_t *err) {
    __shared__ uint2 lds[LDS * kBlock];
    if (P.cnt[9]) return;
    const uint32_t n = P.cnt[cur];
    if (n > limit) return;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (gtid == 0) P.cnt[8] = 1;
    Stack<LDS> st;
    st.init(lds, threadIdx.x, gstack + gtid, stride, gcap, err);
    unsigned long long ext = 0, shadows = 0;
    for (uint32_t i = gtid; i < n; i += stride) {
        const uint32_t p = P.q[cur][i];
        const SceneRec &rec = S.recs[scene_of(P, p)];
        for (int32_t depth = depth0; depth < max_bounce; ++depth) {
            const PathRec pr = load_rec(P, p);
            const DHit h = traverse<false, false>(S, rec, rec_ray(pr), kRefHitMin, kRefHitMax, st, nullptr);
            ++ext;
            bool shadow = false, cont = false;
            ShadowRay sh;
            shade_path(S, P, p, pr, false, h, depth, rr_start, shadow, sh, cont);
            if (shadow) {
                ++shadows;
                if (sh.tl >= 0.0f) {
                    DRay sr;
                    sr.ox = sh.o.x; sr.oy = sh.o.y; sr.oz = sh.o.z;
                    sr.dx = sh.d.x; sr.dy = sh.d.y; sr.dz = sh.d.z;
                    // "closest hit is the light" == no other hit strictly before it
                    const DHit hs = traverse<true, false>(S, rec, sr, kRefHitMin, nextafterf(sh.tl, 0.0f), st, nullptr);
                    if (hs.prim == -1) path_add_l(P, p, sh.c);
                }
            }
            if (!cont) break;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ext += __shfl_xor(ext, off);
        shadows += __shfl_xor(shadows, off);
    }
    if (__lane_id() == 0) {
        if (ext) atomicAdd(&P.tot[0], ext);
        if (shadows) atomicAdd(&P.tot[1], shadows);
    }
}

// m_RenderTarget[idx] += L for each pass in turn (s3d_pt.cpp:330-332).
__global__ __launch_bounds__(kBlock) void k_accumulate(PathBuf P, const BatchDesc *__restrict__ bd,
                                                        const Seg *__restrict__ segs) {
    const uint32_t npixels = bd->pixels, nseg = bd->nseg;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < npixels; t += gridDim.x * blockDim.x) {
        const Seg &g = find_seg<true>(segs, nseg, t);
        const uint32_t j = t - g.pix0;
        float *acc = g.accum + 3 * (size_t)slot_pixel(j, (uint32_t)g.width, (uint32_t)g.height);
        float r = acc[0], gg = acc[1], b = acc[2];
        for (uint32_t s = 0; s < g.passes; ++s) {
            const size_t q = (size_t)g.path0 + (size_t)s * g.npix + j;
            const float4 l = P.path[4 * q + 2];
            r = r + l.y;
            gg = gg + l.z;
            b = b + l.w;
        }
        acc[0] = r;
        acc[1] = gg;
        acc[2] = b;
    }
}

__global__ void k_zero_tot(PathBuf P) {
    if (threadIdx.x < 8 && blockIdx.x == 0) P.tot[threadIdx.x] = 0;
}

// First-hit guide buffers for the denoisers.
__global__ __launch_bounds__(kBlock) void k_guides(SceneDev S, int32_t W, int32_t H, uint64_t seed, uint32_t pass,
                                                    float *albedo, float *normal, float *depth, uint2 *gstack,
                                                    int32_t gcap, int32_t *err) {
    __shared__ uint2 lds[kLdsStack * kBlock];
    const uint32_t n = (uint32_t)W * (uint32_t)H;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack<kLdsStack> st;
    st.init(lds, threadIdx.x, gstack + gtid, stride, gcap, err);
    for (uint32_t p = gtid; p < n; p += stride) {
        const uint32_t x = p % (uint32_t)W, y = p / (uint32_t)W;
        Pcg rng;
        rng.seed(path_seed(seed, p, pass, 0u));
        V3 o, d;
        camera_ray(S.recs[0].cam, (0.5f + (float)x) / (float)W - 0.5f, (0.5f + (float)y) / (float)H - 0.5f, rng, o,
                   d);
        DRay r;
        r.ox = o.x; r.oy = o.y; r.oz = o.z;
        r.dx = d.x; r.dy = d.y; r.dz = d.z;
        const DHit h = traverse<false, false>(S, S.recs[0], r, kRefHitMin, kRefHitMax, st, nullptr);
        V3 a = v3(0, 0, 0), nn = v3(0, 0, 0);
        float t = 0.0f;
        if (h.prim != -1) {
            const V3 pos = o + d * h.t;
            int32_t mat;
            const GShapeRef sr = S.shape_ref[h.shape];
            mat = sr.material;
            if (h.prim == -2) {
                nn = unit3(pos - ld3(S.insts[sr.inst].center));
            } else {
                nn = tri_normal(S.shade[h.prim], h.u, h.v);
                if (mat < 0) mat = S.shade[h.prim].material;
            }
            const GMat &m = S.mats[mat];
            a = base_color(m);
            if (m.texture >= 0) {
                float tu, tv;
                if (h.prim == -2) sphere_texcoord(nn, tu, tv);
                else tri_texcoord(S.tri_uv[h.prim], h.u, h.v, tu, tv);
                a = a * tex_sample(S.texels, S.textures[m.texture], tu, tv);
            }
            t = h.t;
        }
        albedo[3 * p] = a.x; albedo[3 * p + 1] = a.y; albedo[3 * p + 2] = a.z;
        normal[3 * p] = nn.x; normal[3 * p + 1] = nn.y; normal[3 * p + 2] = nn.z;
        depth[p] = t;
    }
}

// Extension and shadow kernels of one traversal variant ([0] plain, [1] counting).
struct TraceKernels {
    const void *ext[2];
    const void *sh[2];
    const void *paths;
    const void *tail;
    int lds;
};

template <int V>
TraceKernels trace_kernels_of() {
    constexpr int L = kVariants[V].lds, W = kVariants[V].waves;
    return TraceKernels{{reinterpret_cast<const void *>(&k_extend<false, L, W>),
                         reinterpret_cast<const void *>(&k_extend<true, L, W>)},
                        {reinterpret_cast<const void *>(&k_shadow<false, L, W>),
                         reinterpret_cast<const void *>(&k_shadow<true, L, W>)},
                        reinterpret_cast<const void *>(&k_paths<L>), reinterpret_cast<const void *>(&k_tail<L>),
                        L};
}

TraceKernels trace_kernels(uint32_t v) {
    static_assert(kVariantCount == 3, "one case per variant");
    switch (v) {
        case 0: return trace_kernels_of<0>();
        case 2: return trace_kernels_of<2>();
        default: return trace_kernels_of<1>();
    }
}

}  // namespace

namespace s3dg {

// The render of a list of frames (RenderJob: scene, resolution, samples, seed,
// accumulation buffer) as batches of path slots: each batch holds consecutive
// passes of one or more frames, up to the path capacity; every batch runs
// its counters' reset, max_bounce x (extend, shade, shadow; the first
// bounce's extend and shade launches generate the camera rays themselves)
// and the accumulation.  Passes
// of a frame are accumulated in pass order whatever the batching, so a frame's
// sum does not depend on what else shares its batches.
int render_jobs(s3dg_context *ctx, const SceneDev &S, Workspace &ws, const RenderJob *jobs, size_t njobs,
                int32_t max_bounce, int32_t rr_start, uint32_t flags, uint32_t passes_in_flight,
                s3dg_render_stats *stats) {
    if (njobs == 0 || !jobs) return set_error(ctx, S3DG_ERR_INVALID, "render of no frames");
    uint64_t total_paths = 0;
    uint32_t max_npix = 1;
    for (size_t j = 0; j < njobs; ++j) {
        const RenderJob &J = jobs[j];
        if (J.width <= 0 || J.height <= 0 || J.spp < 0 || !J.d_accum || (int64_t)J.width * J.height > (1ll << 27))
            return set_error(ctx, S3DG_ERR_INVALID, "bad render size");
        const uint32_t npix = (uint32_t)J.width * (uint32_t)J.height;
        max_npix = std::max(max_npix, npix);
        total_paths += (uint64_t)npix * (uint32_t)J.spp;
    }
    if (max_bounce < 0) return set_error(ctx, S3DG_ERR_INVALID, "bad bounce count");
    if (!passes_in_flight) passes_in_flight = ctx->passes_in_flight;
    uint64_t cap = passes_in_flight ? (uint64_t)passes_in_flight * max_npix : (uint64_t)kAutoPathsInFlight;
    if (!passes_in_flight) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            cap = std::min<uint64_t>(cap, (free_b / 2 + ws.render_bytes) / kPathBytes);
    }
    cap = std::min<uint64_t>(cap, kMaxPathsInFlight);
    cap = std::min<uint64_t>(cap, std::max<uint64_t>(total_paths, 1));
    cap = std::max<uint64_t>(cap, max_npix);
    const uint32_t slots = (uint32_t)cap;

    // ---- batches of segments ----------------------------------------------
    std::vector<std::vector<Seg>> batches;
    {
        std::vector<Seg> cur;
        uint32_t used = 0, pixels = 0;
        auto flush = [&]() {
            if (!cur.empty()) batches.push_back(cur);
            cur.clear();
            used = pixels = 0;
        };
        for (size_t j = 0; j < njobs; ++j) {
            const RenderJob &J = jobs[j];
            const uint32_t npix = (uint32_t)J.width * (uint32_t)J.height;
            uint32_t done = 0;
            while (done < (uint32_t)J.spp) {
                const uint32_t fit = (slots - used) / npix;
                if (fit == 0) {
                    flush();
                    continue;
                }
                const uint32_t k = std::min(fit, (uint32_t)J.spp - done);
                Seg g{};
                g.path0 = used;
                g.npix = npix;
                g.pix0 = pixels;
                g.scene = J.scene;
                g.width = J.width;
                g.height = J.height;
                g.pass0 = (uint32_t)J.first_pass + done;
                g.passes = k;
                g.seed = J.seed;
                g.accum = J.d_accum;
                cur.push_back(g);
                used += k * npix;
                pixels += npix;
                done += k;
            }
        }
        flush();
    }
    size_t max_seg = 1;
    for (auto &b : batches) max_seg = std::max(max_seg, b.size());

    // ---- buffers -------------------------------------------------------------
    int rc = ensure_buffer(ctx, &ws.d_render, &ws.render_bytes, carve_bytes(slots));
    if (rc) return rc;
    PathBuf P = carve(ws.d_render, slots);
    // paths skip the per-path scene index when every job renders one scene
    P.single_scene = (int32_t)jobs[0].scene;
    for (size_t j = 1; j < njobs; ++j)
        if (jobs[j].scene != jobs[0].scene) P.single_scene = -1;
    // device tables: the slot the kernels read (descriptor + segments), then
    // every batch's descriptor and segments, copied into the slot in turn
    std::vector<BatchDesc> descs(batches.size());
    std::vector<Seg> all_segs;
    std::vector<size_t> seg_off(batches.size());
    for (size_t b = 0; b < batches.size(); ++b) {
        const std::vector<Seg> &bs = batches[b];
        descs[b].nseg = (uint32_t)bs.size();
        descs[b].paths = bs.back().path0 + bs.back().passes * bs.back().npix;
        descs[b].pixels = bs.back().pix0 + bs.back().npix;
        seg_off[b] = all_segs.size();
        all_segs.insert(all_segs.end(), bs.begin(), bs.end());
    }
    const size_t slot_bytes = 256 + ((max_seg * sizeof(Seg) + 255) & ~(size_t)255);
    const size_t desc_bytes = (descs.size() * sizeof(BatchDesc) + 255) & ~(size_t)255;
    rc = ensure_buffer(ctx, &ws.d_batch, &ws.batch_bytes,
                       slot_bytes + desc_bytes + std::max<size_t>(all_segs.size(), 1) * sizeof(Seg));
    if (rc) return rc;
    BatchDesc *d_desc = (BatchDesc *)ws.d_batch;
    Seg *d_segs = (Seg *)((char *)ws.d_batch + 256);
    BatchDesc *d_all_desc = (BatchDesc *)((char *)ws.d_batch + slot_bytes);
    Seg *d_all_segs = (Seg *)((char *)ws.d_batch + slot_bytes + desc_bytes);
    S3DG_HIP(ctx, hipMemcpyAsync(d_all_desc, descs.data(), descs.size() * sizeof(BatchDesc), hipMemcpyHostToDevice,
                                 ctx->stream));
    if (!all_segs.empty())
        S3DG_HIP(ctx, hipMemcpyAsync(d_all_segs, all_segs.data(), all_segs.size() * sizeof(Seg),
                                     hipMemcpyHostToDevice, ctx->stream));
    const uint32_t grid = batch_grid(slots);
    const TraceKernels K = trace_kernels(ctx->variant);
    uint32_t tgrid = 0;
    rc = resident_grid(ctx, K.ext[0], &tgrid);
    if (rc) return rc;
    rc = ensure_gstack(ctx, ws, tgrid * kBlock, S.stack_need, K.lds);
    if (rc) return rc;
    int32_t gcap = stack_capacity(S.stack_need, K.lds);
    uint2 *gs = (uint2 *)ws.d_gstack;
    int32_t *d_err = ws.d_error;
    FetchParams fetch{ctx->refill, ctx->fetch_chunk};
    SceneDev Sx = S;
    Sx.tri_batch = ctx->tri_batch;

    const bool timed = (flags & S3DG_RENDER_TIME_KERNELS) != 0;
    const bool count = (flags & S3DG_RENDER_COUNT_WORK) != 0;
    const size_t launches = batches.size() * (size_t)max_bounce;
    // events: [0], [1] around the render, then four per bounce launch when timed
    // (destroyed on every return path)
    EventSet events;
    S3DG_HIP(ctx, events.create(2 + (timed ? 4 * launches : 0)));
    const hipEvent_t e0 = events.ev[0], e1 = events.ev[1];
    hipEvent_t *ev = events.ev.data() + 2;
    S3DG_HIP(ctx, hipEventRecord(e0, ctx->stream));
    hipLaunchKernelGGL(k_zero_tot, dim3(1), dim3(64), 0, ctx->stream, P);

    // one batch's kernels; li = index of its first bounce launch (timing)
    // One batch of up to kPerPathRounds x the resident traversal threads is
    // traced thread per path: each thread carries its paths through every
    // bounce in one launch (same sums as the pipeline).  Measured on the
    // Cornell box at 256x256 the two break even near 16 spp (1 M paths); on
    // the mesh scene at 480x270 thread per path is still 1.8x faster at 16 spp.
    constexpr uint64_t kPerPathRounds = 4;
    const bool per_path = !timed && !count && !(flags & S3DG_RENDER_WAVEFRONT) && batches.size() == 1 &&
                          ((flags & S3DG_RENDER_PER_PATH) || (uint64_t)slots <= kPerPathRounds * tgrid * kBlock);
    // its grid: one thread per path slot, at most as many threads as the
    // traversal stack area was sized for
    const uint32_t pgrid = std::min<uint32_t>((slots + kBlock - 1) / kBlock, tgrid);
    // the pipeline's tail launches take over queues of at most this many
    // paths (automatic: two per thread of the traversal grid, which on the
    // 1080p mesh scene beat one per thread by 2 % and four by 3 %); not with
    // instrumentation, whose per-launch timings and work counts cover the
    // pipeline kernels
    uint32_t tail_limit = (timed || count)                  ? 0u
                          : ctx->tail_paths == 0xffffffffu ? 2u * tgrid * kBlock
                                                            : ctx->tail_paths;
    auto launch_batch = [&](size_t li) {
        if (per_path) {
            void *args[] = {&Sx, &P, &d_desc, &d_segs, &max_bounce, &rr_start, &gs, &gcap, &d_err};
            (void)hipLaunchKernel(K.paths, dim3(pgrid), dim3(kBlock), args, 0, ctx->stream);
            hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(kBlock), 0, ctx->stream, P, d_desc, d_segs);
            return;
        }
        hipLaunchKernelGGL(k_start_batch, dim3(1), dim3(64), 0, ctx->stream, P, d_desc);
        if (max_bounce <= 0) hipLaunchKernelGGL(k_no_bounce, dim3(grid), dim3(kBlock), 0, ctx->stream, P, d_desc);
        int32_t cur = 0;
        for (int32_t depth = 0; depth < max_bounce; ++depth, ++li) {
            if (timed) (void)hipEventRecord(ev[4 * li], ctx->stream);
            int32_t first = depth == 0;
            void *ext_args[] = {&Sx, &P, &cur, &gs, &gcap, &d_err, &fetch, &first, &d_desc, &d_segs};
            (void)hipLaunchKernel(K.ext[count ? 1 : 0], dim3(tgrid), dim3(kBlock), ext_args, 0, ctx->stream);
            if (timed) (void)hipEventRecord(ev[4 * li + 1], ctx->stream);
            hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kBlock), 0, ctx->stream, Sx, P, cur, depth, rr_start, d_desc,
                               d_segs);
            if (timed) (void)hipEventRecord(ev[4 * li + 2], ctx->stream);
            void *sh_args[] = {&Sx, &P, &gs, &gcap, &d_err, &fetch};
            (void)hipLaunchKernel(K.sh[count ? 1 : 0], dim3(tgrid), dim3(kBlock), sh_args, 0, ctx->stream);
            if (timed) (void)hipEventRecord(ev[4 * li + 3], ctx->stream);
            cur ^= 1;
            if (tail_limit && depth + 1 < max_bounce) {
                int32_t depth0 = depth + 1;
                void *tail_args[] = {&Sx, &P, &cur, &depth0, &max_bounce, &rr_start, &tail_limit, &gs, &gcap, &d_err};
                (void)hipLaunchKernel(K.tail, dim3(tgrid), dim3(kBlock), tail_args, 0, ctx->stream);
            }
        }
        hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(kBlock), 0, ctx->stream, P, d_desc, d_segs);
    };

    // Without instrumentation the batch's kernels are replayed from a HIP
    // graph captured once for these launch parameters (the batch tables they
    // read are rewritten before each replay).
    const bool use_graph = !timed && !count;
    if (use_graph) {
        // every device address the captured launches bake in: the stack,
        // path state, batch tables (d_desc and d_segs sit at fixed offsets in
        // ws.d_batch, which a render needing more batches or segments
        // reallocates), the error word and the scene arrays
        const uint64_t key[kGraphKeyWords] = {(uint64_t)(uintptr_t)ws.d_gstack, (uint64_t)(uint32_t)gcap,
                                              (uint64_t)(uintptr_t)ws.d_render,
                                              slots | ((uint64_t)(uint32_t)(P.single_scene + 1) << 32),
                                              (uint64_t)(uint32_t)max_bounce | ((uint64_t)per_path << 32),
                                              ((uint64_t)ctx->variant << 32) | (uint32_t)rr_start, tgrid,
                                              (uint64_t)(uintptr_t)S.nodes, (uint64_t)(uintptr_t)S.recs,
                                              ((uint64_t)fetch.refill << 48) | ((uint64_t)fetch.chunk << 16) |
                                                  Sx.tri_batch,
                                              (uint64_t)(uintptr_t)ws.d_batch,
                                              (uint64_t)(uintptr_t)ws.d_error, tail_limit};
        if (!ws.graph || memcmp(key, ws.graph_key, sizeof(key)) != 0) {
            if (ws.graph) {
                (void)hipGraphExecDestroy(ws.graph);
                ws.graph = nullptr;
            }
            hipGraph_t g = nullptr;
            S3DG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            launch_batch(0);
            S3DG_HIP(ctx, hipStreamEndCapture(ctx->stream, &g));
            const hipError_t ie = hipGraphInstantiate(&ws.graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            S3DG_HIP(ctx, ie);
            memcpy(ws.graph_key, key, sizeof(key));
        }
    }
    size_t li = 0;
    for (size_t b = 0; b < batches.size(); ++b) {
        S3DG_HIP(ctx, hipMemcpyAsync(d_desc, d_all_desc + b, sizeof(BatchDesc), hipMemcpyDeviceToDevice, ctx->stream));
        S3DG_HIP(ctx, hipMemcpyAsync(d_segs, d_all_segs + seg_off[b], batches[b].size() * sizeof(Seg),
                                     hipMemcpyDeviceToDevice, ctx->stream));
        if (use_graph) S3DG_HIP(ctx, hipGraphLaunch(ws.graph, ctx->stream));
        else launch_batch(li);
        li += (size_t)max_bounce;
    }
    S3DG_HIP(ctx, hipGetLastError());
    S3DG_HIP(ctx, hipEventRecord(e1, ctx->stream));
    unsigned long long tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    S3DG_HIP(ctx, hipMemcpyAsync(tot, P.tot, sizeof(tot), hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    S3DG_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    double ext_ms = 0.0, sh_ms = 0.0;
    if (timed) {
        for (size_t i = 0; i < launches; ++i) {
            float a = 0.0f, b = 0.0f;
            S3DG_HIP(ctx, hipEventElapsedTime(&a, ev[4 * i], ev[4 * i + 1]));
            S3DG_HIP(ctx, hipEventElapsedTime(&b, ev[4 * i + 2], ev[4 * i + 3]));
            ext_ms += a;
            sh_ms += b;
        }
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->extension_rays = tot[0];
        stats->shadow_rays = tot[1];
        stats->paths = total_paths;
        stats->milliseconds = ms;
        stats->extend_ms = ext_ms;
        stats->shadow_ms = sh_ms;
        stats->extend_launches = timed ? launches : 0;
        stats->shadow_launches = timed ? launches : 0;
        stats->extend_nodes = tot[2];
        stats->extend_tris = tot[3];
        stats->shadow_nodes = tot[4];
        stats->shadow_tris = tot[5];
        stats->extend_stack_spills = tot[6];
        stats->shadow_stack_spills = tot[7];
    }
    return check_error_flag(ctx, ws);
}

}  // namespace s3dg

extern "C" int s3dg_render_device(s3dg_scene *sc, const s3dg_render_config *cfg, float *d_accum,
                                  s3dg_render_stats *stats) {
    if (!sc || !cfg || !d_accum) return S3DG_ERR_INVALID;
    s3dg_context *ctx = sc->ctx;
    if (cfg->width <= 0 || cfg->height <= 0 || cfg->spp < 0 || cfg->max_bounce_count < 0 ||
        (int64_t)cfg->width * cfg->height > (1ll << 27))
        return set_error(ctx, S3DG_ERR_INVALID, "bad render size");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    RenderJob job{};
    job.scene = 0;
    job.width = cfg->width;
    job.height = cfg->height;
    job.spp = cfg->spp;
    job.seed = cfg->seed;
    job.first_pass = cfg->first_pass;
    job.d_accum = d_accum;
    return render_jobs(ctx, sc->dev, sc->ws, &job, 1, cfg->max_bounce_count, cfg->rr_start_depth, cfg->flags,
                       cfg->passes_in_flight, stats);
}

extern "C" int s3dg_get_tuning(s3dg_context *ctx, s3dg_tuning *out) {
    if (!ctx || !out) return S3DG_ERR_INVALID;
    out->traversal_variant = ctx->variant;
    out->refill = ctx->refill;
    out->tri_batch = ctx->tri_batch;
    out->passes_in_flight = ctx->passes_in_flight;
    out->fetch_chunk = ctx->fetch_chunk;
    out->leaf_size = ctx->leaf_size;
    out->tail_paths = ctx->tail_paths;
    return S3DG_OK;
}

extern "C" int s3dg_set_tuning(s3dg_context *ctx, const s3dg_tuning *t) {
    if (!ctx || !t) return S3DG_ERR_INVALID;
    if (t->traversal_variant >= (uint32_t)kVariantCount || t->refill < 1 || t->refill > 64 || t->tri_batch < 1 ||
        t->tri_batch > 65 || t->fetch_chunk < 64 || t->fetch_chunk > 65536 || t->leaf_size > 8)
        return set_error(ctx, S3DG_ERR_INVALID, "tuning out of range");
    ctx->variant = t->traversal_variant;
    ctx->refill = t->refill;
    ctx->tri_batch = t->tri_batch;
    ctx->passes_in_flight = t->passes_in_flight;
    ctx->fetch_chunk = t->fetch_chunk;
    ctx->leaf_size = t->leaf_size;
    ctx->tail_paths = t->tail_paths;
    return S3DG_OK;
}

// Coordinate descent over the traversal variant, the refill threshold, the
// triangle batching threshold, the fetch chunk and the tail queue length;
// each candidate is timed (device time) as the faster of two trial renders,
// after a first one that captures its graph.  The trials render the caller's
// configuration (samples reduced only beyond 2^27 paths): how deep the
// queues run, and so which settings win, depends on the number of paths in
// flight.
extern "C" int s3dg_autotune(s3dg_scene *sc, const s3dg_render_config *cfg, s3dg_tuning *best, double *best_ms) {
    if (!sc || !cfg || cfg->width <= 0 || cfg->height <= 0) return S3DG_ERR_INVALID;
    s3dg_context *ctx = sc->ctx;
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = (size_t)cfg->width * cfg->height * 3 * sizeof(float);
    int rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, bytes);
    if (rc) return rc;
    s3dg_render_config trial = *cfg;
    const uint64_t npix = (uint64_t)cfg->width * (uint64_t)cfg->height;
    const uint64_t max_spp = std::max<uint64_t>(1, (1ull << 27) / npix);
    trial.spp = (int32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)std::max(cfg->spp, 1), max_spp));
    trial.flags = 0;
    auto time_one = [&](const s3dg_tuning &t, double &ms) -> int {
        int r = s3dg_set_tuning(ctx, &t);
        if (r) return r;
        s3dg_render_stats st{};
        for (int k = 0; k < 3; ++k) {
            r = s3dg_render_device(sc, &trial, (float *)ctx->scratch, &st);
            if (r) return r;
            if (k == 1) ms = st.milliseconds;
            else if (k == 2) ms = std::min(ms, st.milliseconds);
        }
        return S3DG_OK;
    };
    s3dg_tuning cur{};
    s3dg_get_tuning(ctx, &cur);
    const s3dg_tuning start = cur;
    double cur_ms = 0.0;
    rc = time_one(cur, cur_ms);
    if (rc) return rc;
    const uint32_t refills[] = {16, 24, 32, 48};
    const uint32_t batches[] = {8, 16, 24};
    const uint32_t chunks[] = {64, 128, 256};
    const uint32_t tails[] = {0u, 1u << 16, 1u << 19, 0xffffffffu, 1u << 21};
    for (int axis = 0; axis < 5; ++axis) {
        const int n = axis == 0 ? kVariantCount : axis == 1 ? 4 : axis == 4 ? 5 : 3;
        s3dg_tuning axis_best = cur;
        for (int i = 0; i < n; ++i) {
            s3dg_tuning t = cur;
            if (axis == 0) t.traversal_variant = (uint32_t)i;
            else if (axis == 1) t.refill = refills[i];
            else if (axis == 2) t.tri_batch = batches[i];
            else if (axis == 3) t.fetch_chunk = chunks[i];
            else t.tail_paths = tails[i];
            if (memcmp(&t, &cur, sizeof(t)) == 0) continue;
            double ms = 0.0;
            rc = time_one(t, ms);
            if (rc) {
                s3dg_set_tuning(ctx, &start);
                return rc;
            }
            if (ms < cur_ms) {
                cur_ms = ms;
                axis_best = t;
            }
        }
        cur = axis_best;
    }
    s3dg_set_tuning(ctx, &cur);
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return S3DG_OK;
}

extern "C" int s3dg_render(s3dg_scene *sc, const s3dg_render_config *cfg, float *accum, s3dg_render_stats *stats) {
    if (!sc || !cfg || !accum) return S3DG_ERR_INVALID;
    s3dg_context *ctx = sc->ctx;
    if (cfg->width <= 0 || cfg->height <= 0) return set_error(ctx, S3DG_ERR_INVALID, "bad render size");
    const size_t bytes = (size_t)cfg->width * cfg->height * 3 * sizeof(float);
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    int rc = ensure_buffer(ctx, &ctx->scratch, &ctx->scratch_bytes, bytes);
    if (rc) return rc;
    S3DG_HIP(ctx, hipMemcpyAsync(ctx->scratch, accum, bytes, hipMemcpyHostToDevice, ctx->stream));
    rc = s3dg_render_device(sc, cfg, (float *)ctx->scratch, stats);
    if (rc) return rc;
    S3DG_HIP(ctx, hipMemcpyAsync(accum, ctx->scratch, bytes, hipMemcpyDeviceToHost, ctx->stream));
    S3DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return S3DG_OK;
}

extern "C" int s3dg_render_guides_device(s3dg_scene *sc, const s3dg_render_config *cfg, float *d_albedo,
                                         float *d_normal, float *d_depth) {
    if (!sc || !cfg || !d_albedo || !d_normal || !d_depth) return S3DG_ERR_INVALID;
    s3dg_context *ctx = sc->ctx;
    if (cfg->width <= 0 || cfg->height <= 0) return set_error(ctx, S3DG_ERR_INVALID, "bad render size");
    S3DG_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t n = (uint32_t)cfg->width * (uint32_t)cfg->height;
    const uint32_t grid = batch_grid(n);
    int rc = ensure_gstack(ctx, sc->ws, grid * kBlock, sc->dev.stack_need, kLdsStack);
    if (rc) return rc;
    const int32_t gcap = stack_capacity(sc->dev.stack_need);
    hipLaunchKernelGGL(k_guides, dim3(grid), dim3(kBlock), 0, ctx->stream, sc->dev, cfg->width, cfg->height,
                       cfg->seed, (uint32_t)cfg->first_pass, d_albedo, d_normal, d_depth, (uint2 *)sc->ws.d_gstack,
                       gcap, sc->ws.d_error);
    S3DG_HIP(ctx, hipGetLastError());
    return check_error_flag(ctx, sc->ws);
}
