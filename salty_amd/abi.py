"""ctypes mirror of include/s3dg.h (structures and constants) and a plain-data
scene description that can be handed to the C ABI."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

OK = 0
ERR_INVALID = -1
ERR_NOMEM = -2
ERR_DEVICE = -3
ERR_NODEVICE = -4
ERR_UNSUPPORTED = -5

MAT_LAMBERT, MAT_MIRROR, MAT_GLASS, MAT_PHONG, MAT_PLASTIC = range(5)
BVH2, BVH4, BVH8 = 2, 4, 8
SHAPE_MESH, SHAPE_INSTANCE, SHAPE_SPHERE = 0, 1, 2
LIGHT_SHAPE, LIGHT_QUAD = 0, 1
CAMERA_PINHOLE, CAMERA_THINLENS = 0, 1
TRACE_CLOSEST, TRACE_ANY = 0, 1
TONEMAP_REINHARD, TONEMAP_UNCHARTED2, TONEMAP_ACES = 0, 1, 2

F_HIT_MIN = np.float32(1e-3)
F_HIT_MAX = np.float32(1e12)
F_MAX = np.float32(3.402823466e38)

VERTEX_DTYPE = np.dtype([("position", np.float32, 3), ("normal", np.float32, 3), ("texcoord", np.float32, 2)])
RAY_DTYPE = np.dtype([("org", np.float32, 3), ("tmin", np.float32), ("dir", np.float32, 3), ("tmax", np.float32)])
HIT_DTYPE = np.dtype([("prim", np.int32), ("t", np.float32), ("u", np.float32), ("v", np.float32)])


class Vertex(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("normal", C.c_float * 3), ("texcoord", C.c_float * 2)]


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("diffuse", C.c_float * 3), ("specular", C.c_float * 3),
                ("emissive", C.c_float * 3), ("ior", C.c_float), ("power", C.c_float), ("texture", C.c_int32)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb", C.c_void_p)]


class Smd(C.Structure):
    _fields_ = [("vertex_count", C.c_uint32), ("vertices", C.c_void_p), ("triangle_material", C.POINTER(C.c_int32)),
                ("material_count", C.c_uint32), ("materials", C.POINTER(Material)),
                ("texture_count", C.c_uint32), ("texture_paths", C.POINTER(C.c_char * 512))]


class MeshDesc(C.Structure):
    _fields_ = [("vertex_count", C.c_uint32), ("vertices", C.c_void_p), ("material", C.c_int32),
                ("triangle_material", C.c_void_p), ("bvh_kind", C.c_int32)]


class ShapeDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("mesh", C.c_int32), ("material", C.c_int32), ("center", C.c_float * 3),
                ("radius", C.c_float), ("world", C.c_float * 16)]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("shape", C.c_int32), ("first_triangle", C.c_int32),
                ("triangle_count", C.c_int32), ("corner", C.c_float * 3), ("edge1", C.c_float * 3),
                ("edge2", C.c_float * 3)]


class CameraDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", C.c_float * 3), ("target", C.c_float * 3),
                ("upward", C.c_float * 3), ("fov", C.c_float), ("aspect", C.c_float), ("near_clip", C.c_float),
                ("lens_radius", C.c_float), ("reserved", C.c_float)]


class SceneDescC(C.Structure):
    _fields_ = [("mesh_count", C.c_uint32), ("meshes", C.POINTER(MeshDesc)),
                ("shape_count", C.c_uint32), ("shapes", C.POINTER(ShapeDesc)),
                ("material_count", C.c_uint32), ("materials", C.POINTER(Material)),
                ("light_count", C.c_uint32), ("lights", C.POINTER(LightDesc)),
                ("camera", CameraDesc), ("environment", C.c_float * 3),
                ("texture_count", C.c_uint32), ("textures", C.POINTER(TextureDesc)),
                ("environment_texture", C.c_int32)]


class RenderConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("max_bounce_count", C.c_int32), ("spp", C.c_int32),
                ("seed", C.c_uint64), ("rr_start_depth", C.c_int32), ("first_pass", C.c_int32),
                ("flags", C.c_uint32), ("passes_in_flight", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("traversal_variant", C.c_uint32), ("refill", C.c_uint32), ("tri_batch", C.c_uint32),
                ("passes_in_flight", C.c_uint32), ("fetch_chunk", C.c_uint32), ("leaf_size", C.c_uint32),
                ("tail_paths", C.c_uint32)]


class FrameJob(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("seed", C.c_uint64), ("first_pass", C.c_int32), ("reserved", C.c_int32),
                ("d_accum_rgb", C.c_void_p)]


RENDER_TIME_KERNELS = 1
RENDER_COUNT_WORK = 2
RENDER_WAVEFRONT = 4
RENDER_PER_PATH = 8


class RenderStats(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
                ("milliseconds", C.c_double), ("extend_ms", C.c_double), ("shadow_ms", C.c_double),
                ("extend_launches", C.c_uint64), ("shadow_launches", C.c_uint64),
                ("extend_nodes", C.c_uint64), ("extend_tris", C.c_uint64),
                ("shadow_nodes", C.c_uint64), ("shadow_tris", C.c_uint64),
                ("extend_stack_spills", C.c_uint64), ("shadow_stack_spills", C.c_uint64)]


class SceneStats(C.Structure):
    _fields_ = [("triangles", C.c_uint64), ("blas_nodes", C.c_uint64), ("tlas_nodes", C.c_uint64),
                ("device_bytes", C.c_uint64), ("max_depth", C.c_int32), ("build_ms", C.c_double)]


# ---------------------------------------------------------------------------
# Plain-data scene description
# ---------------------------------------------------------------------------
@dataclass
class Mesh:
    vertices: np.ndarray                      # VERTEX_DTYPE, 3 per triangle
    material: int = 0
    triangle_material: np.ndarray | None = None
    bvh_kind: int = BVH8


@dataclass
class Shape:
    type: int
    mesh: int = -1
    material: int = -1
    center: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.0
    world: np.ndarray | None = None           # 4x4, row-vector convention


@dataclass
class Light:
    type: int
    shape: int
    first_triangle: int = 0
    triangle_count: int = 0
    corner: tuple = (0.0, 0.0, 0.0)
    edge1: tuple = (0.0, 0.0, 0.0)
    edge2: tuple = (0.0, 0.0, 0.0)


@dataclass
class Camera:
    position: tuple
    target: tuple
    upward: tuple = (0.0, 1.0, 0.0)
    fov: float = 0.7279404
    aspect: float = 1.0
    near_clip: float = 1.0
    type: int = CAMERA_PINHOLE
    lens_radius: float = 0.0


@dataclass
class SceneDesc:
    meshes: list = field(default_factory=list)
    shapes: list = field(default_factory=list)
    materials: list = field(default_factory=list)   # dicts: type, diffuse, specular, emissive, ior, power,
                                                    # texture (index into textures, -1 / absent: none)
    lights: list = field(default_factory=list)
    camera: Camera | None = None
    environment: tuple = (0.0, 0.0, 0.0)
    textures: list = field(default_factory=list)    # float32 arrays (height, width, 3), row y = GetColor row y
    environment_texture: int = -1                   # IBL: index into textures, -1: constant environment

    def triangle_count(self) -> int:
        return sum(len(m.vertices) // 3 for m in self.meshes)

    def to_c(self):
        """Returns (SceneDescC, keepalive) for the C ABI."""
        keep = []
        meshes = (MeshDesc * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            v = np.ascontiguousarray(m.vertices, dtype=VERTEX_DTYPE)
            keep.append(v)
            meshes[i].vertex_count = len(v)
            meshes[i].vertices = v.ctypes.data
            meshes[i].material = m.material
            if m.triangle_material is not None:
                tm = np.ascontiguousarray(m.triangle_material, dtype=np.int32)
                keep.append(tm)
                meshes[i].triangle_material = tm.ctypes.data
            else:
                meshes[i].triangle_material = None
            meshes[i].bvh_kind = m.bvh_kind
        shapes = (ShapeDesc * max(1, len(self.shapes)))()
        for i, s in enumerate(self.shapes):
            shapes[i].type = s.type
            shapes[i].mesh = s.mesh
            shapes[i].material = s.material
            shapes[i].center[:] = [float(x) for x in s.center]
            shapes[i].radius = s.radius
            w = np.eye(4, dtype=np.float32) if s.world is None else np.asarray(s.world, dtype=np.float32)
            shapes[i].world[:] = [float(x) for x in w.reshape(-1)]
        mats = (Material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i].type = m.get("type", MAT_LAMBERT)
            mats[i].diffuse[:] = list(m.get("diffuse", (0.0, 0.0, 0.0)))
            mats[i].specular[:] = list(m.get("specular", (0.0, 0.0, 0.0)))
            mats[i].emissive[:] = list(m.get("emissive", (0.0, 0.0, 0.0)))
            mats[i].ior = m.get("ior", 1.5)
            mats[i].power = m.get("power", 1.0)
            mats[i].texture = int(m.get("texture", -1)) + 1
        lights = (LightDesc * max(1, len(self.lights)))()
        for i, l in enumerate(self.lights):
            lights[i].type = l.type
            lights[i].shape = l.shape
            lights[i].first_triangle = l.first_triangle
            lights[i].triangle_count = l.triangle_count
            lights[i].corner[:] = list(l.corner)
            lights[i].edge1[:] = list(l.edge1)
            lights[i].edge2[:] = list(l.edge2)
        texs = (TextureDesc * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            a = np.ascontiguousarray(t, dtype=np.float32)
            assert a.ndim == 3 and a.shape[2] == 3
            keep.append(a)
            texs[i].width = a.shape[1]
            texs[i].height = a.shape[0]
            texs[i].rgb = a.ctypes.data
        keep += [meshes, shapes, mats, lights, texs]
        d = SceneDescC()
        d.mesh_count = len(self.meshes)
        d.meshes = meshes
        d.shape_count = len(self.shapes)
        d.shapes = shapes
        d.material_count = len(self.materials)
        d.materials = mats
        d.light_count = len(self.lights)
        d.lights = lights
        cam = self.camera
        d.camera.type = cam.type
        d.camera.position[:] = list(cam.position)
        d.camera.target[:] = list(cam.target)
        d.camera.upward[:] = list(cam.upward)
        d.camera.fov = cam.fov
        d.camera.aspect = cam.aspect
        d.camera.near_clip = cam.near_clip
        d.camera.lens_radius = cam.lens_radius
        d.environment[:] = list(self.environment)
        d.texture_count = len(self.textures)
        d.textures = texs
        d.environment_texture = self.environment_texture + 1
        return d, keep


def make_rays(org: np.ndarray, dirs: np.ndarray, tmin=F_HIT_MIN, tmax=F_HIT_MAX) -> np.ndarray:
    r = np.zeros(len(org), dtype=RAY_DTYPE)
    r["org"] = org
    r["dir"] = dirs
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r
