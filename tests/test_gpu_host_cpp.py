"""The C++ drop-in layer (salty_amd/host/s3d_gpu.h) end to end: a Cornell box
built with Mesh::Create / MaterialFactory / Scene / PathTracer::Run renders the
same image as the oracle given the same geometry, seed and pass count."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi
from salty_amd.abi import Camera, Light, Mesh, SceneDesc, Shape

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "salty_amd", "host", "s3d_cornell")


def scene_from_dump(path, w, h):
    raw = open(path, "rb").read()
    meshes, off = [], 0
    for _ in range(4):
        n = int(np.frombuffer(raw, np.uint32, 1, off)[0])
        off += 4
        meshes.append(np.frombuffer(raw, abi.VERTEX_DTYPE, n, off).copy())
        off += n * abi.VERTEX_DTYPE.itemsize
    d = SceneDesc()
    d.materials = [dict(type=abi.MAT_LAMBERT, diffuse=c) for c in
                   ((0.75, 0.75, 0.75), (0.75, 0.25, 0.25), (0.25, 0.75, 0.25))]
    d.materials.append(dict(type=abi.MAT_LAMBERT, diffuse=(0, 0, 0), emissive=(15.0, 15.0, 15.0)))
    d.meshes = [Mesh(v, material=i) for i, v in enumerate(meshes)]
    d.shapes = [Shape(abi.SHAPE_MESH, mesh=i) for i in range(4)]
    y, lo = np.float32(0.999), np.float32(0.375)
    d.lights = [Light(abi.LIGHT_QUAD, shape=3, first_triangle=0, triangle_count=2, corner=(lo, y, lo),
                      edge1=(0.25, 0.0, 0.0), edge2=(0.0, 0.0, 0.25))]
    fov = float(np.frombuffer(raw, np.float32, 1, off)[0])
    d.camera = Camera(position=(0.5, 0.5, 2.4), target=(0.5, 0.5, 0.0), fov=fov, aspect=w / h, near_clip=1.0)
    return d


def test_cpp_host_cornell_matches_oracle(tmp_path):
    w = h = 64
    passes = 4
    ppm, sums, geo = tmp_path / "out.ppm", tmp_path / "sum.bin", tmp_path / "geo.bin"
    out = subprocess.run([EXE, str(w), str(h), str(passes), str(ppm), str(sums), str(geo)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    img = np.fromfile(sums, dtype=np.float32).reshape(h, w, 3)
    ref, counts = O.OracleScene(scene_from_dump(geo, w, h)).render(w, h, spp=passes, max_depth=8, seed=0)
    rel = np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2))
    assert rel < 1e-4, rel
    assert f"passes={passes}" in out.stdout
    data = ppm.read_bytes()
    assert data.startswith(b"P6\n64 64\n255\n") and len(data) == len(b"P6\n64 64\n255\n") + w * h * 3
