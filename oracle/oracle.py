"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from salty_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    """Compiles liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_philox4x32_10": (None, [vp, vp, vp]),
            "orc_pcg_stream": (None, [C.c_uint64, C.c_int32, vp]),
            "orc_pcg_srandom_stream": (None, [C.c_uint64, C.c_uint64, C.c_int32, vp]),
            "orc_pcg_f32": (C.c_float, [C.c_uint32]),
            "orc_path_seed": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_triangle_hit": (C.c_int, [vp, vp, vp, vp, vp]),
            "orc_tangent_space": (None, [vp, vp, vp]),
            "orc_sample_lambert": (None, [C.c_float, C.c_float, vp]),
            "orc_sample_phong": (None, [C.c_float, C.c_float, C.c_float, vp]),
            "orc_shade": (C.c_int, [vp, vp, vp, vp, vp, vp]),
            "orc_invert": (None, [vp, vp]),
            "orc_mesh_create": (vp, [C.c_uint32, vp, C.c_int32]),
            "orc_mesh_destroy": (None, [vp]),
            "orc_mesh_order": (C.c_int, [vp, vp]),
            "orc_mesh_tree": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_mesh_trace": (C.c_int, [vp, vp, C.c_uint32, C.c_int32, vp, C.c_int32]),
            "orc_scene_create": (vp, [C.POINTER(abi.SceneDescC)]),
            "orc_scene_destroy": (None, [vp]),
            "orc_scene_trace": (C.c_int, [vp, vp, C.c_uint32, vp, vp, C.c_int32]),
            "orc_camera_ray": (None, [vp, C.c_float, C.c_float, vp, vp]),
            "orc_camera_ray_lens": (None, [vp, C.c_float, C.c_float, vp, vp, vp]),
            "orc_texture_sample": (None, [C.c_int32, C.c_int32, vp, C.c_float, C.c_float, vp]),
            "orc_texture_sample_dir": (None, [C.c_int32, C.c_int32, vp, vp, vp]),
            "orc_render": (C.c_int, [vp, C.POINTER(abi.RenderConfig), C.c_int32, C.c_int32, vp, vp, C.c_int32]),
            "orc_render_guides": (C.c_int, [vp, C.POINTER(abi.RenderConfig), vp, vp, vp]),
            "orc_tonemap": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp]),
            "orc_capture": (C.c_int, [C.c_int32, C.c_int32, vp, vp]),
            "orc_nlm": (C.c_int, [C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_int32, vp, vp, C.c_int32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def pcg_stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().orc_pcg_stream(seed, n, _p(out))
    return out


def pcg_srandom_stream(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().orc_pcg_srandom_stream(initstate, initseq, n, _p(out))
    return out


def triangle_hit(v0, v1, v2, ray):
    f = lambda x: np.ascontiguousarray(x, dtype=np.float32)  # noqa: E731
    r = np.ascontiguousarray(ray, dtype=abi.RAY_DTYPE).reshape(1)
    h = np.empty(1, dtype=abi.HIT_DTYPE)
    a, b, c = f(v0), f(v1), f(v2)
    hit = lib().orc_triangle_hit(_p(a), _p(b), _p(c), _p(r), _p(h))
    return bool(hit), h[0]


def tangent_space(n):
    n = np.ascontiguousarray(n, dtype=np.float32)
    t = np.empty(3, np.float32)
    b = np.empty(3, np.float32)
    lib().orc_tangent_space(_p(n), _p(t), _p(b))
    return t, b


def sample_lambert(u, v):
    o = np.empty(3, np.float32)
    lib().orc_sample_lambert(u, v, _p(o))
    return o


def shade(material: dict, inp, normal, state: int):
    m = abi.Material()
    m.type = material.get("type", abi.MAT_LAMBERT)
    m.diffuse[:] = list(material.get("diffuse", (0, 0, 0)))
    m.specular[:] = list(material.get("specular", (0, 0, 0)))
    m.emissive[:] = list(material.get("emissive", (0, 0, 0)))
    m.ior = material.get("ior", 1.5)
    m.power = material.get("power", 1.0)
    i = np.ascontiguousarray(inp, dtype=np.float32)
    n = np.ascontiguousarray(normal, dtype=np.float32)
    st = np.array([state], dtype=np.uint64)
    o = np.empty(3, np.float32)
    w = np.empty(3, np.float32)
    dice = lib().orc_shade(C.addressof(m), _p(i), _p(n), _p(st), _p(o), _p(w))
    return o, w, bool(dice), int(st[0])


def texture_sample(texels: np.ndarray, u: float, v: float) -> np.ndarray:
    """Texture::SampleColor(Vector2) on an H x W x 3 texel array."""
    t = np.ascontiguousarray(texels, dtype=np.float32)
    o = np.empty(3, np.float32)
    lib().orc_texture_sample(t.shape[1], t.shape[0], _p(t), u, v, _p(o))
    return o


def texture_sample_dir(texels: np.ndarray, d) -> np.ndarray:
    """Texture::SampleColor(Vector3): the latitude-longitude lookup."""
    t = np.ascontiguousarray(texels, dtype=np.float32)
    dd = np.ascontiguousarray(d, dtype=np.float32)
    o = np.empty(3, np.float32)
    lib().orc_texture_sample_dir(t.shape[1], t.shape[0], _p(t), _p(dd), _p(o))
    return o


def invert(m) -> np.ndarray:
    a = np.ascontiguousarray(m, dtype=np.float32).reshape(16)
    out = np.empty(16, np.float32)
    lib().orc_invert(_p(a), _p(out))
    return out.reshape(4, 4)


class OracleMesh:
    """Reference SAH tree over one mesh, with reference traversal and brute force."""

    def __init__(self, vertices: np.ndarray, kind: int = abi.BVH8):
        self.v = np.ascontiguousarray(vertices, dtype=abi.VERTEX_DTYPE)
        self.h = lib().orc_mesh_create(len(self.v), _p(self.v), kind)
        if not self.h:
            raise ValueError("orc_mesh_create failed")
        self.n = len(self.v) // 3

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def order(self) -> np.ndarray:
        o = np.empty(self.n, np.int32)
        lib().orc_mesh_order(self.h, _p(o))
        return o

    def tree(self):
        nn = lib().orc_mesh_tree(self.h, None, None, None, None, None, None, None)
        kind = np.empty(nn, np.int32)
        cb = np.empty(nn, np.int32)
        cc = np.empty(nn, np.int32)
        ch = np.empty(nn * 8 + 8, np.int32)
        lb = np.empty(nn, np.int32)
        lc = np.empty(nn, np.int32)
        boxes = np.empty((nn, 6), np.float32)
        lib().orc_mesh_tree(self.h, _p(kind), _p(cb), _p(cc), _p(ch), _p(lb), _p(lc), _p(boxes))
        return dict(kind=kind, child_begin=cb, child_count=cc, children=ch, begin=lb, count=lc, boxes=boxes)

    def trace(self, rays: np.ndarray, brute: bool = False, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        hits = np.empty(len(rays), dtype=abi.HIT_DTYPE)
        lib().orc_mesh_trace(self.h, _p(rays), len(rays), 1 if brute else 0, _p(hits), threads)
        return hits


class OracleScene:
    def __init__(self, desc: abi.SceneDesc):
        self.desc = desc
        cdesc, keep = desc.to_c()
        self.h = lib().orc_scene_create(C.byref(cdesc))
        del keep

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def trace(self, rays: np.ndarray, threads: int = 0):
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        hits = np.empty(len(rays), dtype=abi.HIT_DTYPE)
        shapes = np.empty(len(rays), dtype=np.int32)
        lib().orc_scene_trace(self.h, _p(rays), len(rays), _p(hits), _p(shapes), threads)
        return hits, shapes

    def camera_ray(self, x: float, y: float, state: int | None = None):
        """(origin, direction, state) of the camera ray through image-plane
        coordinates x, y in [-0.5, 0.5]; with a PCG state the thin lens is
        sampled from it (the state returned advanced)."""
        o = np.empty(3, np.float32)
        d = np.empty(3, np.float32)
        if state is None:
            lib().orc_camera_ray(self.h, x, y, _p(o), _p(d))
            return o, d, None
        st = np.array([state], np.uint64)
        lib().orc_camera_ray_lens(self.h, x, y, _p(st), _p(o), _p(d))
        return o, d, int(st[0])

    def render(self, width, height, spp, max_depth, seed=0, rr_start_depth=0, first_pass=0, rows=None,
               accum=None, threads: int = 0):
        if accum is None:
            accum = np.zeros((height, width, 3), np.float32)
        cfg = abi.RenderConfig(width, height, max_depth, spp, seed, rr_start_depth, first_pass, 0, 0)
        r0, r1 = (0, height) if rows is None else rows
        counts = np.zeros(3, np.uint64)
        lib().orc_render(self.h, C.byref(cfg), r0, r1, _p(accum), _p(counts), threads)
        return accum, counts

    def guides(self, width, height, seed=0, first_pass=0):
        cfg = abi.RenderConfig(width, height, 1, 1, seed, 0, first_pass, 0, 0)
        a = np.zeros((height, width, 3), np.float32)
        n = np.zeros((height, width, 3), np.float32)
        d = np.zeros((height, width), np.float32)
        lib().orc_render_guides(self.h, C.byref(cfg), _p(a), _p(n), _p(d))
        return a, n, d


def tonemap(rgb: np.ndarray, kind: int = abi.TONEMAP_ACES):
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    out = np.empty_like(rgb)
    ave = np.zeros(1, np.float32)
    mx = np.zeros(1, np.float32)
    lib().orc_tonemap(kind, w, h, _p(rgb), _p(out), _p(ave), _p(mx))
    return out, float(ave[0]), float(mx[0])


def capture(rgb: np.ndarray) -> np.ndarray:
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    out = np.empty(rgb.shape, np.uint8)
    lib().orc_capture(w, h, _p(rgb), _p(out))
    return out


def nlm(rgba: np.ndarray, coeff: float, search_radius: int = 6, patch_radius: int = 2, threads: int = 0):
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    out = np.empty_like(rgba)
    lib().orc_nlm(w, h, coeff, search_radius, patch_radius, _p(rgba), _p(out), threads)
    return out
