"""salty_amd: MI355X (gfx950) rendering hot path with a C ABI (include/s3dg.h).

This module loads the in-tree ``libs3dg.so`` through ctypes and exposes thin
Python handles over it.  There is no CPU fallback: if the library or a gfx950
device is missing, the constructors raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (HIT_DTYPE, RAY_DTYPE, VERTEX_DTYPE, RenderConfig, RenderStats, SceneDesc,  # noqa: F401
                  SceneStats)

_HERE = os.path.dirname(os.path.abspath(__file__))
# S3DG_LIBRARY names another build of the library (for example one compiled
# with different kernel parameters); the default is the in-tree build.
LIB_PATH = os.environ.get("S3DG_LIBRARY") or os.path.join(_HERE, "libs3dg.so")

_lib = None

# Exported C entry points and their argument types.
_SIGNATURES = {
    "s3dg_create": [C.POINTER(C.c_void_p), C.c_int],
    "s3dg_destroy": [C.c_void_p],
    "s3dg_last_error": [C.c_void_p],
    "s3dg_synchronize": [C.c_void_p],
    "s3dg_malloc": [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)],
    "s3dg_free": [C.c_void_p, C.c_void_p],
    "s3dg_memcpy_htod": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t],
    "s3dg_memcpy_dtoh": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t],
    "s3dg_stream": [C.c_void_p],
    "s3dg_scene_create": [C.c_void_p, C.POINTER(abi.SceneDescC), C.POINTER(C.c_void_p)],
    "s3dg_scene_destroy": [C.c_void_p],
    "s3dg_scene_get_stats": [C.c_void_p, C.POINTER(abi.SceneStats)],
    "s3dg_scene_mesh_order": [C.c_void_p, C.c_uint32, C.c_void_p],
    "s3dg_scene_update_instances": [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p],
    "s3dg_build_order": [C.c_uint32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p],
    "s3dg_image_load_bmp": [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_void_p)],
    "s3dg_image_load_hdr": [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_void_p)],
    "s3dg_image_free": [C.c_void_p],
    "s3dg_smd_load": [C.c_char_p, C.POINTER(C.POINTER(abi.Smd))],
    "s3dg_smd_free": [C.POINTER(abi.Smd)],
    "s3dg_trace": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p],
    "s3dg_trace_device": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p],
    "s3dg_trace_counted_device": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p],
    "s3dg_render": [C.c_void_p, C.POINTER(RenderConfig), C.c_void_p, C.POINTER(RenderStats)],
    "s3dg_render_device": [C.c_void_p, C.POINTER(RenderConfig), C.c_void_p, C.POINTER(RenderStats)],
    "s3dg_render_guides_device": [C.c_void_p, C.POINTER(RenderConfig), C.c_void_p, C.c_void_p, C.c_void_p],
    "s3dg_get_tuning": [C.c_void_p, C.POINTER(abi.Tuning)],
    "s3dg_set_tuning": [C.c_void_p, C.POINTER(abi.Tuning)],
    "s3dg_autotune": [C.c_void_p, C.POINTER(RenderConfig), C.POINTER(abi.Tuning), C.POINTER(C.c_double)],
    "s3dg_scene_set_create": [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
    "s3dg_scene_set_destroy": [C.c_void_p],
    "s3dg_render_batch_device": [C.c_void_p, C.c_uint32, C.POINTER(abi.FrameJob), C.POINTER(RenderConfig),
                                 C.POINTER(RenderStats)],
    "s3dg_tonemap": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p],
    "s3dg_tonemap_device": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p],
    "s3dg_tonemap_device_noalloc": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                    C.c_void_p],
    "s3dg_nlm": [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p],
    "s3dg_nlm_device": [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_int32, C.c_void_p,
                        C.c_void_p],
    "s3dg_shade": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "s3dg_base_color": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "s3dg_camera_rays": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p],
    "s3dg_sphere_sample": [C.c_void_p, C.c_void_p, C.c_float, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p],
    "s3dg_scene_calc_param": [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p],
    "s3dg_scene_environment": [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p],
    "s3dg_cnn_create": [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)],
    "s3dg_cnn_destroy": [C.c_void_p],
    "s3dg_cnn_param_count": [C.c_void_p, C.POINTER(C.c_uint64)],
    "s3dg_cnn_get_params": [C.c_void_p, C.c_void_p, C.c_uint64],
    "s3dg_cnn_set_params": [C.c_void_p, C.c_void_p, C.c_uint64],
    "s3dg_cnn_infer_device": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p],
    "s3dg_denoiser_features_device": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "s3dg_cnn_gradient_device": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_double), C.c_void_p, C.c_uint64],
    "s3dg_cnn_train_step_device": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_float,
                                   C.POINTER(C.c_double)],
}


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libs3dg.so and declares every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(path)
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.s3dg_last_error.restype = C.c_char_p
    lib.s3dg_stream.restype = C.c_void_p
    lib.s3dg_image_free.restype = None
    lib.s3dg_smd_free.restype = None
    _lib = lib
    return lib


class S3dgError(RuntimeError):
    pass


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def build_order(vertices: np.ndarray, kind: int = abi.BVH8):
    """Depth-first permutation of the reference SAH build (host only, no device)."""
    lib = load_library()
    v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
    order = np.empty(len(v) // 3, dtype=np.int32)
    nodes = C.c_uint32()
    rc = lib.s3dg_build_order(len(v), _ptr(v), kind, _ptr(order), C.byref(nodes))
    if rc != abi.OK:
        raise S3dgError(f"s3dg_build_order failed ({rc})")
    return order, nodes.value


def _load_image(fn, path: str) -> np.ndarray:
    w, h, p = C.c_int32(), C.c_int32(), C.c_void_p()
    rc = fn(os.fsencode(path), C.byref(w), C.byref(h), C.byref(p))
    if rc != abi.OK:
        raise S3dgError(f"cannot load {path} ({rc})")
    lib = load_library()
    try:
        buf = (C.c_float * (w.value * h.value * 3)).from_address(p.value)
        return np.frombuffer(buf, dtype=np.float32).reshape(h.value, w.value, 3).copy()
    finally:
        lib.s3dg_image_free(p)


def load_bmp(path: str) -> np.ndarray:
    """24-bit BMP as (height, width, 3) float32 in file row order (host only)."""
    return _load_image(load_library().s3dg_image_load_bmp, path)


def load_hdr(path: str) -> np.ndarray:
    """Radiance RGBE image as (height, width, 3) float32 (host only)."""
    return _load_image(load_library().s3dg_image_load_hdr, path)


def load_smd(path: str) -> dict:
    """SMD mesh file (host only): vertices (VERTEX_DTYPE), triangle_material
    (int32, -1: none), materials (dicts as SceneDesc takes them, with a texture
    index into texture_paths or -1) and texture_paths."""
    lib = load_library()
    p = C.POINTER(abi.Smd)()
    rc = lib.s3dg_smd_load(os.fsencode(path), C.byref(p))
    if rc != abi.OK:
        raise S3dgError(f"cannot load {path} ({rc})")
    try:
        m = p.contents
        nv = m.vertex_count
        vbuf = (C.c_byte * (nv * VERTEX_DTYPE.itemsize)).from_address(m.vertices)
        vertices = np.frombuffer(vbuf, dtype=VERTEX_DTYPE).copy()
        tri_mat = np.ctypeslib.as_array(m.triangle_material, shape=(nv // 3,)).copy()
        mats = []
        for i in range(m.material_count):
            c = m.materials[i]
            mats.append(dict(type=c.type, diffuse=tuple(c.diffuse), specular=tuple(c.specular),
                             emissive=tuple(c.emissive), ior=c.ior, power=c.power, texture=c.texture - 1))
        paths = [m.texture_paths[i].value.decode() for i in range(m.texture_count)]
        return dict(vertices=vertices, triangle_material=tri_mat, materials=mats, texture_paths=paths)
    finally:
        lib.s3dg_smd_free(p)


class Context:
    """Owns a HIP stream and the device memory of everything created from it."""

    def __init__(self, device: int = -1):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.s3dg_create(C.byref(h), device)
        if rc != abi.OK:
            raise S3dgError(f"s3dg_create failed ({rc}): no usable gfx950 device")
        self.handle = h

    def check(self, rc: int, what: str):
        if rc != abi.OK:
            msg = self.lib.s3dg_last_error(self.handle)
            raise S3dgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "handle", None):
            self.lib.s3dg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self.check(self.lib.s3dg_synchronize(self.handle), "s3dg_synchronize")

    @property
    def stream(self) -> int:
        return self.lib.s3dg_stream(self.handle)

    # -- device buffers -------------------------------------------------
    def get_tuning(self) -> abi.Tuning:
        t = abi.Tuning()
        self.check(self.lib.s3dg_get_tuning(self.handle, C.byref(t)), "s3dg_get_tuning")
        return t

    def set_tuning(self, t: abi.Tuning):
        self.check(self.lib.s3dg_set_tuning(self.handle, C.byref(t)), "s3dg_set_tuning")

    def malloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self.check(self.lib.s3dg_malloc(self.handle, nbytes, C.byref(p)), "s3dg_malloc")
        return p.value

    def free(self, ptr: int):
        self.check(self.lib.s3dg_free(self.handle, C.c_void_p(ptr)), "s3dg_free")

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.malloc(arr.nbytes)
        self.check(self.lib.s3dg_memcpy_htod(self.handle, C.c_void_p(p), _ptr(arr), arr.nbytes), "htod")
        return p

    def download(self, ptr: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        self.check(self.lib.s3dg_memcpy_dtoh(self.handle, _ptr(out), C.c_void_p(ptr), out.nbytes), "dtoh")
        return out

    def copy_to_device(self, ptr: int, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        self.check(self.lib.s3dg_memcpy_htod(self.handle, C.c_void_p(ptr), _ptr(arr), arr.nbytes), "htod")

    # -- post -------------------------------------------------------------
    def tonemap(self, rgb: np.ndarray, kind: int = abi.TONEMAP_ACES):
        """Returns (tonemapped float RGB, 8-bit sRGB) of an HxWx3 float image."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        h, w = rgb.shape[:2]
        out = np.empty_like(rgb)
        out8 = np.empty(rgb.shape, dtype=np.uint8)
        self.check(self.lib.s3dg_tonemap(self.handle, kind, w, h, _ptr(rgb), _ptr(out), _ptr(out8)),
                   "s3dg_tonemap")
        return out, out8

    def nlm(self, rgba: np.ndarray, coeff: float, search_radius: int = 6, patch_radius: int = 2) -> np.ndarray:
        rgba = np.ascontiguousarray(rgba, dtype=np.float32)
        h, w = rgba.shape[:2]
        out = np.empty_like(rgba)
        self.check(self.lib.s3dg_nlm(self.handle, w, h, coeff, search_radius, patch_radius, _ptr(rgba),
                                     _ptr(out)), "s3dg_nlm")
        return out

    def nlm_device(self, w: int, h: int, coeff: float, search_radius: int, patch_radius: int, d_in: int,
                   d_out: int):
        self.check(self.lib.s3dg_nlm_device(self.handle, w, h, coeff, search_radius, patch_radius,
                                            C.c_void_p(d_in), C.c_void_p(d_out)), "s3dg_nlm_device")

    def features_device(self, n: int, h: int, w: int, d_accum: int, spp: int, d_albedo: int, d_normal: int,
                        d_depth: int, d_features: int):
        """The denoiser's 10 input features from a render's device buffers."""
        self.check(self.lib.s3dg_denoiser_features_device(self.handle, n, h, w, C.c_void_p(d_accum), spp,
                                                          C.c_void_p(d_albedo), C.c_void_p(d_normal),
                                                          C.c_void_p(d_depth), C.c_void_p(d_features)),
                   "s3dg_denoiser_features_device")


class Scene:
    """A scene uploaded to the device (two-level compressed BVH + shading data)."""

    def __init__(self, ctx: Context, desc: SceneDesc):
        self.ctx = ctx
        self.desc = desc
        cdesc, keep = desc.to_c()
        h = C.c_void_p()
        ctx.check(ctx.lib.s3dg_scene_create(ctx.handle, C.byref(cdesc), C.byref(h)), "s3dg_scene_create")
        del keep
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.s3dg_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> SceneStats:
        s = SceneStats()
        self.ctx.check(self.ctx.lib.s3dg_scene_get_stats(self.handle, C.byref(s)), "stats")
        return s

    def mesh_order(self, mesh: int) -> np.ndarray:
        n = len(self.desc.meshes[mesh].vertices) // 3
        out = np.empty(n, dtype=np.int32)
        self.ctx.check(self.ctx.lib.s3dg_scene_mesh_order(self.handle, mesh, _ptr(out)), "mesh_order")
        return out

    def update_instances(self, shapes, worlds):
        """New world matrices (4x4 each, row-vector convention) for INSTANCE shapes."""
        sh = np.ascontiguousarray(shapes, dtype=np.uint32)
        w = np.ascontiguousarray(worlds, dtype=np.float32).reshape(len(sh), 16)
        self.ctx.check(self.ctx.lib.s3dg_scene_update_instances(self.handle, len(sh), _ptr(sh), _ptr(w)),
                       "s3dg_scene_update_instances")

    def trace(self, rays: np.ndarray, mode: int = abi.TRACE_CLOSEST, with_shapes: bool = False):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.empty(len(rays), dtype=HIT_DTYPE)
        shapes = np.empty(len(rays), dtype=np.int32) if with_shapes else None
        self.ctx.check(self.ctx.lib.s3dg_trace(self.handle, _ptr(rays), len(rays), mode, _ptr(hits),
                                               _ptr(shapes)), "s3dg_trace")
        return (hits, shapes) if with_shapes else hits

    def trace_device(self, d_rays: int, n: int, mode: int, d_hits: int, d_shapes: int | None = None):
        self.ctx.check(self.ctx.lib.s3dg_trace_device(self.handle, C.c_void_p(d_rays), n, mode,
                                                      C.c_void_p(d_hits),
                                                      None if d_shapes is None else C.c_void_p(d_shapes)),
                       "s3dg_trace_device")

    def trace_counted_device(self, d_rays: int, n: int, mode: int, d_hits: int, d_counters: int):
        self.ctx.check(self.ctx.lib.s3dg_trace_counted_device(self.handle, C.c_void_p(d_rays), n, mode,
                                                              C.c_void_p(d_hits), C.c_void_p(d_counters)),
                       "s3dg_trace_counted_device")

    def render(self, width: int, height: int, spp: int, max_depth: int, seed: int = 0, rr_start_depth: int = 0,
               first_pass: int = 0, accum: np.ndarray | None = None, flags: int = 0, passes_in_flight: int = 0):
        """Adds spp passes to accum (HxWx3 sum buffer) and returns (accum, stats)."""
        if accum is None:
            accum = np.zeros((height, width, 3), dtype=np.float32)
        cfg = RenderConfig(width, height, max_depth, spp, seed, rr_start_depth, first_pass, flags, passes_in_flight)
        st = RenderStats()
        self.ctx.check(self.ctx.lib.s3dg_render(self.handle, C.byref(cfg), _ptr(accum), C.byref(st)),
                       "s3dg_render")
        return accum, st

    def render_device(self, cfg: RenderConfig, d_accum: int) -> RenderStats:
        st = RenderStats()
        self.ctx.check(self.ctx.lib.s3dg_render_device(self.handle, C.byref(cfg), C.c_void_p(d_accum),
                                                       C.byref(st)), "s3dg_render_device")
        return st

    def autotune(self, cfg: RenderConfig):
        """Times trial renders of cfg for the candidate kernel settings and keeps
        the fastest in the context; returns (tuning, device ms of the trial)."""
        t = abi.Tuning()
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.s3dg_autotune(self.handle, C.byref(cfg), C.byref(t), C.byref(ms)),
                       "s3dg_autotune")
        return t, ms.value

    def guides_device(self, cfg: RenderConfig, d_albedo: int, d_normal: int, d_depth: int):
        self.ctx.check(self.ctx.lib.s3dg_render_guides_device(self.handle, C.byref(cfg), C.c_void_p(d_albedo),
                                                              C.c_void_p(d_normal), C.c_void_p(d_depth)),
                       "s3dg_render_guides_device")


class SceneSet:
    """Scenes of one context merged for batched rendering (the scenes stay usable)."""

    def __init__(self, ctx: Context, scenes):
        self.ctx = ctx
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[sc.handle.value for sc in self.scenes])
        h = C.c_void_p()
        ctx.check(ctx.lib.s3dg_scene_set_create(ctx.handle, len(self.scenes), arr, C.byref(h)),
                  "s3dg_scene_set_create")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.s3dg_scene_set_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render_batch_device(self, frames, max_depth: int, rr_start_depth: int = 0, flags: int = 0,
                            passes_in_flight: int = 0) -> RenderStats:
        """frames: sequence of (scene, width, height, spp, seed, first_pass, d_accum)."""
        jobs = (abi.FrameJob * len(frames))()
        for j, (scene, w, h, spp, seed, first_pass, d_acc) in enumerate(frames):
            jobs[j] = abi.FrameJob(scene, w, h, spp, seed, first_pass, 0, d_acc)
        cfg = RenderConfig(0, 0, max_depth, 0, 0, rr_start_depth, 0, flags, passes_in_flight)
        st = RenderStats()
        self.ctx.check(self.ctx.lib.s3dg_render_batch_device(self.handle, len(frames), jobs, C.byref(cfg),
                                                             C.byref(st)), "s3dg_render_batch_device")
        return st


class DeviceTimer:
    """Device milliseconds of the work a context's stream runs between
    start() and stop(), from a pair of HIP events recorded on that stream
    (the HIP runtime the library already loaded, through ctypes)."""

    _hip = None

    def __init__(self, ctx: "Context"):
        if DeviceTimer._hip is None:
            try:
                hip = C.CDLL("libamdhip64.so")
            except OSError:
                hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
            hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
            hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
            hip.hipEventSynchronize.argtypes = [C.c_void_p]
            hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
            hip.hipEventDestroy.argtypes = [C.c_void_p]
            DeviceTimer._hip = hip
        self.ctx = ctx
        self.ev = [C.c_void_p(), C.c_void_p()]
        for e in self.ev:
            if DeviceTimer._hip.hipEventCreate(C.byref(e)) != 0:
                raise S3dgError("hipEventCreate failed")

    def start(self):
        if DeviceTimer._hip.hipEventRecord(self.ev[0], C.c_void_p(self.ctx.stream)) != 0:
            raise S3dgError("hipEventRecord failed")

    def stop(self) -> float:
        hip = DeviceTimer._hip
        if hip.hipEventRecord(self.ev[1], C.c_void_p(self.ctx.stream)) != 0 or hip.hipEventSynchronize(self.ev[1]) != 0:
            raise S3dgError("hipEventRecord / hipEventSynchronize failed")
        ms = C.c_float()
        if hip.hipEventElapsedTime(C.byref(ms), self.ev[0], self.ev[1]) != 0:
            raise S3dgError("hipEventElapsedTime failed")
        return float(ms.value)

    def close(self):
        for e in self.ev:
            if e:
                DeviceTimer._hip.hipEventDestroy(e)
        self.ev = []


class Denoiser:
    """The U-Net denoiser (BF16 MFMA kernels) on a context."""

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(ctx.lib.s3dg_cnn_create(ctx.handle, seed, C.byref(h)), "s3dg_cnn_create")
        self.handle = h
        n = C.c_uint64()
        ctx.check(ctx.lib.s3dg_cnn_param_count(h, C.byref(n)), "s3dg_cnn_param_count")
        self.nparam = n.value

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.s3dg_cnn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparam, np.float32)
        self.ctx.check(self.ctx.lib.s3dg_cnn_get_params(self.handle, _ptr(out), self.nparam), "get_params")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self.ctx.check(self.ctx.lib.s3dg_cnn_set_params(self.handle, _ptr(p), self.nparam), "set_params")

    def infer_device(self, n, h, w, d_feat: int, d_out: int):
        self.ctx.check(self.ctx.lib.s3dg_cnn_infer_device(self.handle, n, h, w, C.c_void_p(d_feat),
                                                          C.c_void_p(d_out)), "s3dg_cnn_infer_device")

    def infer(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, dtype=np.float32)
        n, h, w, _ = f.shape
        d_f = self.ctx.upload(f)
        d_o = self.ctx.malloc(n * h * w * 3 * 4)
        try:
            self.infer_device(n, h, w, d_f, d_o)
            return self.ctx.download(d_o, (n, h, w, 3), np.float32)
        finally:
            self.ctx.free(d_f)
            self.ctx.free(d_o)

    def gradient(self, features: np.ndarray, target: np.ndarray):
        f = np.ascontiguousarray(features, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        n, h, w, _ = f.shape
        d_f, d_t = self.ctx.upload(f), self.ctx.upload(t)
        g = np.empty(self.nparam, np.float32)
        loss = C.c_double()
        try:
            self.ctx.check(self.ctx.lib.s3dg_cnn_gradient_device(self.handle, n, h, w, C.c_void_p(d_f),
                                                                 C.c_void_p(d_t), C.byref(loss), _ptr(g),
                                                                 self.nparam), "s3dg_cnn_gradient_device")
        finally:
            self.ctx.free(d_f)
            self.ctx.free(d_t)
        return loss.value, g

    def train_step_device(self, n, h, w, d_feat: int, d_target: int, lr: float) -> float:
        loss = C.c_double()
        self.ctx.check(self.ctx.lib.s3dg_cnn_train_step_device(self.handle, n, h, w, C.c_void_p(d_feat),
                                                               C.c_void_p(d_target), lr, C.byref(loss)),
                       "s3dg_cnn_train_step_device")
        return loss.value
