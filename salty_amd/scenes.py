"""Seeded procedural scenes for the rendering hot path.

* ``cornell_box``      -- 36-triangle Cornell box, quad light (radiance 15).
* ``textured_cornell`` -- the same box with textured materials, a textured
                          sphere and a lat-long IBL.
* ``smd_scene``        -- an SMD mesh file with its materials and textures,
                          lit by a spherical light.
* ``mesh_scene``       -- ~1.05 M triangles: displaced level-7 icosphere, 64
                          instanced level-3 icospheres, 640 k-triangle ground
                          grid, one spherical light.
* ``triangle_soup`` / ``soup_rays`` -- random triangle soup and random rays for
                          the intersection microbenchmark.
* ``batch_scene``      -- one member of the uneven multi-scene batch.

All generators are deterministic functions of their seed (numpy PCG64).
"""
from __future__ import annotations

import math

import numpy as np

from .abi import (BVH8, CAMERA_PINHOLE, CAMERA_THINLENS, F_HIT_MAX, LIGHT_QUAD, LIGHT_SHAPE, MAT_GLASS,
                  MAT_LAMBERT, MAT_MIRROR, MAT_PHONG, MAT_PLASTIC,
                  RAY_DTYPE, SHAPE_INSTANCE, SHAPE_MESH, SHAPE_SPHERE, VERTEX_DTYPE, Camera, Light, Mesh,
                  SceneDesc, Shape)

# The reference uses the camera's fov linearly (vertical FOV = 2*atan(fov/2)),
# so a 40 degree vertical field of view is fov = 2*tan(20 deg).
FOV_40 = float(np.float32(2.0 * math.tan(math.radians(20.0))))


def _tri_vertices(p: np.ndarray, n: np.ndarray | None = None) -> np.ndarray:
    """(T,3,3) positions (+ optional (T,3,3) normals) -> VERTEX_DTYPE array of 3T."""
    t = p.shape[0]
    v = np.zeros(t * 3, dtype=VERTEX_DTYPE)
    v["position"] = p.reshape(-1, 3)
    if n is None:
        fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        ln = np.linalg.norm(fn, axis=1, keepdims=True)
        fn = fn / np.maximum(ln, 1e-30)
        n = np.repeat(fn[:, None, :], 3, axis=1)
    v["normal"] = n.reshape(-1, 3)
    return v


def _quad(a, b, c, d):
    """Two triangles (a,b,c), (a,c,d)."""
    return [[a, b, c], [a, c, d]]


def _box_tris(size, center_xz, angle_deg):
    sx, sy, sz = size
    hx, hz = sx / 2, sz / 2
    corners = np.array([[-hx, 0, -hz], [hx, 0, -hz], [hx, 0, hz], [-hx, 0, hz],
                        [-hx, sy, -hz], [hx, sy, -hz], [hx, sy, hz], [-hx, sy, hz]], dtype=np.float64)
    a = math.radians(angle_deg)
    rot = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    c = corners @ rot.T + np.array([center_xz[0], 0.0, center_xz[1]])
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (3, 0, 4, 7)]
    tris = []
    for f in faces:
        tris += _quad(c[f[0]], c[f[1]], c[f[2]], c[f[3]])
    return tris


def cornell_box() -> SceneDesc:
    """Cornell box in [0,1]^3 open towards +z: five walls (10 triangles), a
    tall box rotated 15 degrees and a short box rotated -18 degrees (12
    triangles each), and a 0.25 x 0.25 quad light of radiance 15 just under the
    ceiling (2 triangles): 36 triangles, Lambert materials."""
    tris, mats = [], []
    white, red, green, light = 0, 1, 2, 3
    P = lambda *x: np.array(x, dtype=np.float64)  # noqa: E731
    tris += _quad(P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0)); mats += [white] * 2      # floor
    tris += _quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1)); mats += [white] * 2      # ceiling
    tris += _quad(P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0)); mats += [white] * 2      # back
    tris += _quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1)); mats += [red] * 2        # left
    tris += _quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0)); mats += [green] * 2      # right
    tris += _box_tris((0.3, 0.6, 0.3), (0.33, 0.37), 15.0); mats += [white] * 12
    tris += _box_tris((0.3, 0.3, 0.3), (0.67, 0.67), -18.0); mats += [white] * 12
    y = 0.999
    lo, hi = 0.5 - 0.125, 0.5 + 0.125
    tris += _quad(P(lo, y, lo), P(hi, y, lo), P(hi, y, hi), P(lo, y, hi)); mats += [light] * 2
    p = np.array(tris, dtype=np.float32)
    assert p.shape == (36, 3, 3)
    verts = _tri_vertices(p)
    materials = [
        dict(type=MAT_LAMBERT, diffuse=(0.75, 0.75, 0.75)),
        dict(type=MAT_LAMBERT, diffuse=(0.75, 0.25, 0.25)),
        dict(type=MAT_LAMBERT, diffuse=(0.25, 0.75, 0.25)),
        dict(type=MAT_LAMBERT, diffuse=(0.0, 0.0, 0.0), emissive=(15.0, 15.0, 15.0)),
    ]
    desc = SceneDesc()
    desc.meshes = [Mesh(verts, triangle_material=np.array(mats, dtype=np.int32))]
    desc.shapes = [Shape(SHAPE_MESH, mesh=0)]
    desc.materials = materials
    desc.lights = [Light(LIGHT_QUAD, shape=0, first_triangle=34, triangle_count=2, corner=(lo, y, lo),
                         edge1=(0.25, 0.0, 0.0), edge2=(0.0, 0.0, 0.25))]
    desc.camera = Camera(position=(0.5, 0.5, 2.4), target=(0.5, 0.5, 0.0), upward=(0.0, 1.0, 0.0), fov=FOV_40,
                         aspect=1.0, near_clip=1.0)
    desc.environment = (0.0, 0.0, 0.0)
    return desc


def textured_cornell(seed: int = 0) -> SceneDesc:
    """The Cornell box with textures and an IBL (TexturedMaterial and
    Scene::SampleIBL): a seeded 16 x 16 checker on the floor with texture
    coordinates running over [-1.25, 2.75] (so both wrap directions and the
    negative-coordinate quirk are hit), a 37 x 23 noise texture on the back wall
    and on the tall box (a textured mirror), a textured Lambert sphere, and a
    64 x 32 lat-long environment seen through the open front."""
    desc = cornell_box()
    rng = np.random.default_rng(seed + 1000)
    v = desc.meshes[0].vertices
    mats = desc.meshes[0].triangle_material.copy()

    def quad_uv(t0, u0, v0, u1, v1):
        # triangles (a, b, c), (a, c, d) of _quad with uv corners a..d
        uv = np.array([[u0, v0], [u1, v0], [u1, v1], [u0, v1]], dtype=np.float32)
        for k, idx in enumerate(((0, 1, 2), (0, 2, 3))):
            v["texcoord"][3 * (t0 + k):3 * (t0 + k) + 3] = uv[list(idx)]

    quad_uv(0, -1.25, -1.25, 2.75, 2.75)     # floor
    quad_uv(4, 0.0, 0.0, 1.0, 1.0)           # back wall
    for t in range(10, 22):                  # tall box faces
        v["texcoord"][3 * t:3 * t + 3] = [[0.1, 0.2], [0.9, 0.3], [0.5, 0.8]]
    mats[0:2] = 4                            # floor: textured white Lambert
    mats[4:6] = 5                            # back wall: textured Lambert
    mats[10:22] = 6                          # tall box: textured mirror
    desc.meshes[0].triangle_material = mats
    checker = np.zeros((16, 16, 3), dtype=np.float32)
    yy, xx = np.mgrid[0:16, 0:16]
    on = ((xx // 2 + yy // 2) % 2 == 0)
    checker[on] = rng.uniform(0.7, 1.0, 3)
    checker[~on] = rng.uniform(0.05, 0.3, 3)
    noise = rng.uniform(0.1, 1.0, (23, 37, 3)).astype(np.float32)
    env = np.zeros((32, 64, 3), dtype=np.float32)
    env[..., 0] = np.linspace(0.0, 0.6, 64)[None, :]
    env[..., 1] = np.linspace(0.5, 0.0, 32)[:, None]
    env[..., 2] = 0.2
    desc.textures = [checker, noise, env]
    desc.environment_texture = 2
    desc.materials += [dict(type=MAT_LAMBERT, diffuse=(0.8, 0.8, 0.8), texture=0),
                       dict(type=MAT_LAMBERT, diffuse=(0.9, 0.85, 0.8), texture=1),
                       dict(type=MAT_MIRROR, specular=(0.95, 0.95, 0.95), texture=1),
                       dict(type=MAT_LAMBERT, diffuse=(0.9, 0.9, 0.9), texture=0)]
    desc.shapes.append(Shape(SHAPE_SPHERE, material=7, center=(0.72, 0.45, 0.55), radius=0.12))
    return desc


def smd_scene(path: str, light_radiance: float = 20.0, environment=(0.05, 0.05, 0.05)) -> SceneDesc:
    """A scene around an SMD mesh file (Mesh::Create(filename), its materials
    and BMP textures): the mesh, a spherical light above it and a camera
    framing its bounding box.  A texture that does not load leaves its material
    untextured (the reference logs a warning and keeps the empty texture)."""
    from . import S3dgError, load_bmp, load_smd
    m = load_smd(path)
    textures, slot = [], {}
    for i, tp in enumerate(m["texture_paths"]):
        try:
            textures.append(load_bmp(tp))
            slot[i] = len(textures) - 1
        except S3dgError:
            pass
    mats = []
    for mt in m["materials"]:
        mt = dict(mt)
        mt["texture"] = slot.get(mt.get("texture", -1), -1)
        mats.append(mt)
    light = len(mats)
    mats.append(dict(type=MAT_LAMBERT, diffuse=(0.0, 0.0, 0.0), emissive=(light_radiance,) * 3))
    v = m["vertices"]
    lo, hi = v["position"].min(axis=0), v["position"].max(axis=0)
    c, ext = (lo + hi) / 2, float(np.max(hi - lo))
    desc = SceneDesc()
    tm = m["triangle_material"].copy()
    tm[tm < 0] = 0
    desc.meshes = [Mesh(v, triangle_material=tm)]
    desc.materials = mats
    desc.textures = textures
    r = 0.1 * ext
    centre = (float(c[0]), float(hi[1] + 0.6 * ext), float(c[2] + 0.3 * ext))
    desc.shapes = [Shape(SHAPE_MESH, mesh=0), Shape(SHAPE_SPHERE, material=light, center=centre, radius=r)]
    desc.lights = [Light(LIGHT_SHAPE, shape=1)]
    desc.camera = Camera(position=(float(c[0]), float(c[1]), float(c[2] + 1.6 * ext)),
                         target=(float(c[0]), float(c[1]), float(c[2])), upward=(0.0, 1.0, 0.0), fov=FOV_40,
                         aspect=1.0, near_clip=1.0)
    desc.environment = tuple(environment)
    return desc


def cornell_variant(seed: int) -> SceneDesc:
    """A seeded Cornell-box-style scene (the denoiser's training and held-out
    scenes): random wall colours, box sizes, positions and rotations, light
    size and position; Lambert materials, 36 triangles, radiance-15 quad light.
    Seed 0 is the canonical box of cornell_box()."""
    if seed == 0:
        return cornell_box()
    rng = np.random.default_rng(seed)
    P = lambda *x: np.array(x, dtype=np.float64)  # noqa: E731
    tris, mats = [], []
    tris += _quad(P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0)); mats += [0] * 2
    tris += _quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1)); mats += [0] * 2
    tris += _quad(P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0)); mats += [0] * 2
    tris += _quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1)); mats += [1] * 2
    tris += _quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0)); mats += [2] * 2
    for k in range(2):
        sx, sz = rng.uniform(0.15, 0.35, 2)
        sy = rng.uniform(0.15, 0.7)
        cx = rng.uniform(0.2, 0.8)
        cz = rng.uniform(0.2, 0.8)
        tris += _box_tris((sx, sy, sz), (cx, cz), rng.uniform(-45, 45))
        mats += [3 + k] * 12
    y = 0.999
    half = rng.uniform(0.06, 0.2)
    lx, lz = rng.uniform(0.25, 0.75, 2)
    lo_x, hi_x, lo_z, hi_z = lx - half, lx + half, lz - half, lz + half
    tris += _quad(P(lo_x, y, lo_z), P(hi_x, y, lo_z), P(hi_x, y, hi_z), P(lo_x, y, hi_z)); mats += [5] * 2
    p = np.array(tris, dtype=np.float32)
    col = lambda: tuple(float(c) for c in rng.uniform(0.1, 0.85, 3))  # noqa: E731
    materials = [dict(type=MAT_LAMBERT, diffuse=tuple(float(c) for c in rng.uniform(0.6, 0.85, 1).repeat(3))),
                 dict(type=MAT_LAMBERT, diffuse=col()), dict(type=MAT_LAMBERT, diffuse=col()),
                 dict(type=MAT_LAMBERT, diffuse=col()), dict(type=MAT_LAMBERT, diffuse=col()),
                 dict(type=MAT_LAMBERT, diffuse=(0.0, 0.0, 0.0), emissive=(15.0, 15.0, 15.0))]
    desc = SceneDesc()
    desc.meshes = [Mesh(_tri_vertices(p), triangle_material=np.array(mats, dtype=np.int32))]
    desc.shapes = [Shape(SHAPE_MESH, mesh=0)]
    desc.materials = materials
    desc.lights = [Light(LIGHT_QUAD, shape=0, first_triangle=34, triangle_count=2,
                         corner=(float(np.float32(lo_x)), float(np.float32(y)), float(np.float32(lo_z))),
                         edge1=(float(np.float32(hi_x) - np.float32(lo_x)), 0.0, 0.0),
                         edge2=(0.0, 0.0, float(np.float32(hi_z) - np.float32(lo_z))))]
    j = rng.normal(scale=0.05, size=2)
    desc.camera = Camera(position=(0.5 + j[0], 0.5 + j[1], 2.4), target=(0.5, 0.5, 0.0), upward=(0.0, 1.0, 0.0),
                         fov=FOV_40, aspect=1.0, near_clip=1.0)
    return desc


# ---------------------------------------------------------------------------
# Icospheres and value noise
# ---------------------------------------------------------------------------
def icosphere(level: int):
    """Unit icosphere: (vertices (V,3) float64, faces (F,3) int64), F = 20*4^level."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        a = base + inv[:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:]
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v, f


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(73856093)) ^ (iy.astype(np.uint64) * np.uint64(19349663)) ^ \
        (iz.astype(np.uint64) * np.uint64(83492791)) ^ np.uint64(seed * 2654435761 % (1 << 32))
    h ^= h >> np.uint64(13)
    h *= np.uint64(0x5bd1e995)
    h ^= h >> np.uint64(15)
    return (h & np.uint64(0xffffff)).astype(np.float64) / float(0xffffff)


def value_noise(p: np.ndarray, seed: int, octaves: int = 3, base_freq: float = 2.0) -> np.ndarray:
    """Sum of `octaves` octaves of trilinear value noise in [-1, 1], halving the
    amplitude and doubling the frequency per octave, normalised to [-1, 1]."""
    total = np.zeros(len(p))
    amp_sum = 0.0
    for o in range(octaves):
        f = base_freq * (2 ** o)
        a = 0.5 ** o
        q = p * f + 1000.0
        i = np.floor(q).astype(np.int64)
        w = q - i
        w = w * w * (3 - 2 * w)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    r = _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed + 7919 * o)
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += wx * wy * wz * (2 * r - 1)
        total += a * acc
        amp_sum += a
    return total / amp_sum


def _smooth_normals(v, f):
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    vn = np.zeros_like(v)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    return vn


def _mesh_vertices(v, f, normals=None):
    p = v[f].astype(np.float32)
    n = None if normals is None else normals[f].astype(np.float32)
    return _tri_vertices(p, n)


def _grid(x0, x1, z0, z1, nx, nz, y=0.0):
    xs = np.linspace(x0, x1, nx + 1)
    zs = np.linspace(z0, z1, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    v = np.stack([X.reshape(-1), np.full(X.size, y), Z.reshape(-1)], 1)
    idx = np.arange((nx + 1) * (nz + 1)).reshape(nx + 1, nz + 1)
    a = idx[:-1, :-1].reshape(-1)
    b = idx[1:, :-1].reshape(-1)
    c = idx[1:, 1:].reshape(-1)
    d = idx[:-1, 1:].reshape(-1)
    f = np.concatenate([np.stack([a, d, c], 1), np.stack([a, c, b], 1)])
    n = np.zeros_like(v)
    n[:, 1] = 1.0
    return v, f, n


def _random_material(rng, mix: str = "c2") -> dict:
    u = rng.random()
    if mix == "all":
        # every material of the reference: Lambert, Mirror, Glass, Phong, Plastic
        if u < 0.40:
            return dict(type=MAT_LAMBERT, diffuse=tuple(float(x) for x in rng.uniform(0.2, 0.9, 3)))
        if u < 0.50:
            return dict(type=MAT_MIRROR, specular=(0.95, 0.95, 0.95))
        if u < 0.60:
            return dict(type=MAT_GLASS, specular=(0.99, 0.99, 0.99), ior=1.5)
        if u < 0.80:
            return dict(type=MAT_PHONG, specular=tuple(float(x) for x in rng.uniform(0.3, 0.9, 3)),
                        power=float(rng.uniform(5.0, 200.0)))
        return dict(type=MAT_PLASTIC, diffuse=tuple(float(x) for x in rng.uniform(0.2, 0.9, 3)),
                    specular=(0.8, 0.8, 0.8), power=float(rng.uniform(10.0, 500.0)))
    if u < 0.70:
        return dict(type=MAT_LAMBERT, diffuse=tuple(float(x) for x in rng.uniform(0.2, 0.9, 3)))
    if u < 0.85:
        return dict(type=MAT_MIRROR, specular=(0.95, 0.95, 0.95))
    return dict(type=MAT_GLASS, specular=(0.99, 0.99, 0.99), ior=1.5)


def translate(x, y, z) -> np.ndarray:
    """Matrix::Translate (row-vector convention, translation in the last row)."""
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (x, y, z)
    return m


def mesh_scene(seed: int = 0, big_level: int = 7, small_level: int = 3, ground=(800, 400),
               instances: int = 64, materials: str = "c2", lens_radius: float = 0.0) -> SceneDesc:
    """The ~1.05 M-triangle scene: a level-7 icosphere (327,680 triangles) of
    radius 1 displaced by 3-octave value noise of amplitude 0.05; `instances`
    translated copies of a level-3 icosphere of radius 0.4 (1,280 triangles
    each) at seeded uniform positions on a 20 x 20 ground grid of 800 x 400
    quads (640,000 triangles); one spherical light.  Materials per object:
    70 % Lambert (albedo U[0.2,0.9]), 15 % mirror, 15 % glass (IOR 1.5); the
    ground is Lambert 0.7.  materials="all" draws from all five of the
    reference's materials instead (Phong and Plastic included, the ground
    Plastic), and lens_radius > 0 makes the camera a ThinLensCamera focused
    on the target."""
    rng = np.random.default_rng(seed)
    desc = SceneDesc()
    mats = [dict(type=MAT_LAMBERT, diffuse=(0.7, 0.7, 0.7))]
    if materials == "all":
        mats = [dict(type=MAT_PLASTIC, diffuse=(0.6, 0.6, 0.6), specular=(0.5, 0.5, 0.5), power=80.0)]
    # ground
    gv, gf, gn = _grid(-10.0, 10.0, -10.0, 10.0, ground[0], ground[1])
    desc.meshes.append(Mesh(_mesh_vertices(gv, gf, gn), material=0))
    # displaced big icosphere
    v, f = icosphere(big_level)
    disp = 0.05 * value_noise(v, seed, octaves=3, base_freq=2.0)
    vd = v * (1.0 + disp)[:, None] + np.array([0.0, 1.05, 0.0])
    mats.append(_random_material(rng, materials))
    desc.meshes.append(Mesh(_mesh_vertices(vd, f, _smooth_normals(vd, f)), material=1))
    # small icosphere mesh shared by the instances
    sv, sf = icosphere(small_level)
    r_small = 0.4
    desc.meshes.append(Mesh(_mesh_vertices(sv * r_small, sf, sv), material=0))
    desc.shapes.append(Shape(SHAPE_MESH, mesh=0))
    desc.shapes.append(Shape(SHAPE_MESH, mesh=1))
    placed = 0
    while placed < instances:
        x, z = rng.uniform(-9.5, 9.5, 2)
        if x * x + z * z < 1.6 ** 2:
            continue
        mats.append(_random_material(rng, materials))
        desc.shapes.append(Shape(SHAPE_INSTANCE, mesh=2, material=len(mats) - 1,
                                 world=translate(float(x), r_small, float(z))))
        placed += 1
    mats.append(dict(type=MAT_LAMBERT, diffuse=(0.0, 0.0, 0.0), emissive=(60.0, 60.0, 60.0)))
    desc.shapes.append(Shape(SHAPE_SPHERE, material=len(mats) - 1, center=(4.0, 7.0, 3.0), radius=1.0))
    desc.lights = [Light(LIGHT_SHAPE, shape=len(desc.shapes) - 1)]
    desc.materials = mats
    desc.camera = Camera(position=(0.0, 3.5, 11.0), target=(0.0, 0.8, 0.0), upward=(0.0, 1.0, 0.0), fov=FOV_40,
                         aspect=16.0 / 9.0, near_clip=1.0,
                         type=CAMERA_THINLENS if lens_radius > 0 else CAMERA_PINHOLE, lens_radius=lens_radius)
    desc.environment = (0.15, 0.18, 0.22)
    return desc


def triangle_soup(n: int, seed: int = 0) -> SceneDesc:
    """n triangles: v0 ~ U[0,1]^3, v1 and v2 = v0 + (uniform unit direction) *
    N(0.01, 0.003) (lengths clipped below at 1e-3)."""
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3))
    d = rng.normal(size=(n, 2, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    ln = np.maximum(rng.normal(0.01, 0.003, size=(n, 2)), 1e-3)
    p = np.empty((n, 3, 3), dtype=np.float32)
    p[:, 0] = v0
    p[:, 1] = v0 + d[:, 0] * ln[:, 0:1]
    p[:, 2] = v0 + d[:, 1] * ln[:, 1:2]
    desc = SceneDesc()
    desc.meshes = [Mesh(_tri_vertices(p), material=0, bvh_kind=BVH8)]
    desc.shapes = [Shape(SHAPE_MESH, mesh=0)]
    desc.materials = [dict(type=MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))]
    desc.camera = Camera(position=(0.5, 0.5, 3.0), target=(0.5, 0.5, 0.0))
    return desc


def soup_rays(n: int, seed: int = 1, tmin: float = 1e-4, tmax: float = float("inf")) -> np.ndarray:
    """n rays: origins U[-0.1, 1.1]^3, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=RAY_DTYPE)
    r["org"] = rng.uniform(-0.1, 1.1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["dir"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def camera_rays(desc: SceneDesc, width: int, height: int) -> np.ndarray:
    """Pixel-centre pinhole rays computed in float32 exactly as
    PinholeCamera::Update / GetRay do (for trace-level tests)."""
    f = np.float32
    c = desc.camera
    pos = np.array(c.position, f)
    tgt = np.array(c.target, f)
    up = np.array(c.upward, f)

    def unit(v):
        m = np.sqrt(f(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + f(v[..., 2] * v[..., 2]))
        return (v / m[..., None]).astype(f)

    def cross(a, b):
        return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], f)

    d = unit(tgt - pos)
    cx = (unit(cross(d, up)) * f(c.fov) * f(c.aspect)).astype(f)
    cy = (unit(cross(d, cx)) * f(c.fov)).astype(f)
    cz = (pos + d * f(c.near_clip)).astype(f)
    xs = ((f(0.5) + np.arange(width, dtype=f)) / f(width) - f(0.5)).astype(f)
    ys = ((f(0.5) + np.arange(height, dtype=f)) / f(height) - f(0.5)).astype(f)
    X, Y = np.meshgrid(xs, ys)
    p = (cx[None, None] * X[..., None]).astype(f) + (cy[None, None] * Y[..., None]).astype(f)
    p = (p + cz).astype(f)
    dirs = unit((p - pos).astype(f)).reshape(-1, 3)
    r = np.zeros(len(dirs), dtype=RAY_DTYPE)
    r["org"] = pos
    r["dir"] = dirs
    r["tmin"] = f(1e-3)
    r["tmax"] = F_HIT_MAX
    return r


def batch_scene(seed: int, tris: int) -> SceneDesc:
    """One scene of the uneven batch: a displaced icosphere whose level is
    chosen to give about `tris` triangles (levels 3..7, plus a plain ground
    grid making up the remainder), a spherical light, seeded materials."""
    rng = np.random.default_rng(seed)
    level = int(np.clip(round(math.log(max(tris * 0.5, 1280) / 20.0, 4)), 3, 7))
    v, f = icosphere(level)
    disp = 0.05 * value_noise(v, seed, octaves=3)
    vd = v * (1.0 + disp)[:, None] + np.array([0.0, 1.05, 0.0])
    rest = max(tris - len(f), 2)
    q = max(1, int(math.sqrt(rest / 2)))
    gv, gf, gn = _grid(-6, 6, -6, 6, q, q)
    desc = SceneDesc()
    desc.materials = [dict(type=MAT_LAMBERT, diffuse=(0.7, 0.7, 0.7)), _random_material(rng),
                      dict(type=MAT_LAMBERT, diffuse=(0.0, 0.0, 0.0), emissive=(50.0, 50.0, 50.0))]
    desc.meshes = [Mesh(_mesh_vertices(gv, gf, gn), material=0),
                   Mesh(_mesh_vertices(vd, f, _smooth_normals(vd, f)), material=1)]
    desc.shapes = [Shape(SHAPE_MESH, mesh=0), Shape(SHAPE_MESH, mesh=1),
                   Shape(SHAPE_SPHERE, material=2, center=(3.0, 6.0, 2.0), radius=0.8)]
    desc.lights = [Light(LIGHT_SHAPE, shape=2)]
    desc.camera = Camera(position=(0.0, 2.5, 6.5), target=(0.0, 0.9, 0.0), fov=FOV_40, aspect=4.0 / 3.0,
                         near_clip=1.0, type=CAMERA_PINHOLE)
    desc.environment = (0.15, 0.18, 0.22)
    return desc


BATCH_RESOLUTIONS = ((320, 240), (640, 480), (1280, 720), (1920, 1080))
BATCH_SPP = (4, 16, 64)


def batch_config(seed: int = 0, count: int = 64, tri_range=(1_000, 500_000)):
    """The uneven batch: `count` scenes with triangle counts log-uniform in
    tri_range, resolutions drawn from BATCH_RESOLUTIONS and samples per pixel
    from BATCH_SPP.  Returns [(scene description, width, height, spp)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        tris = int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
        w, h = BATCH_RESOLUTIONS[rng.integers(len(BATCH_RESOLUTIONS))]
        spp = int(BATCH_SPP[rng.integers(len(BATCH_SPP))])
        desc = batch_scene(seed * 1000 + k, tris)
        desc.camera.aspect = w / h
        out.append((desc, w, h, spp))
    return out
