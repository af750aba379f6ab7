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


def test_cpp_host_smd_mesh_matches_oracle(tmp_path):
    """Mesh::Create(filename) on an SMD file with a BMP colour map, lit by a
    sphere light, through Scene / PathTracer::Run, against the oracle."""
    import struct

    from _formats import write_bmp, write_smd
    from salty_amd import load_smd
    rng = np.random.default_rng(21)
    n = 12
    xs = np.linspace(-1, 1, n + 1)
    tris = []
    for i in range(n):
        for j in range(n):
            q = [(xs[i], xs[j]), (xs[i], xs[j + 1]), (xs[i + 1], xs[j + 1]), (xs[i + 1], xs[j])]
            for a, b, c in ((0, 1, 2), (0, 2, 3)):
                tris.append([[q[k][0], 0.1 * q[k][0] * q[k][1], q[k][1], 0, 1, 0, q[k][0], q[k][1]]
                             for k in (a, b, c)])
    pos = np.array(tris, np.float32)
    ids = rng.integers(0, 2, len(pos)).astype(np.int32)
    f3 = lambda *x: struct.pack("<3f", *x)  # noqa: E731
    mats = [(0, f3(0.9, 0.9, 0.9) + f3(0, 0, 0) + struct.pack("<i", 0)),
            (4, f3(0.3, 0.6, 0.3) + f3(0.5, 0.5, 0.5) + struct.pack("<f", 25.0) + f3(0, 0, 0)
             + struct.pack("<ii", -1, -1))]
    write_bmp(str(tmp_path / "t.bmp"), rng.integers(0, 256, (8, 8, 3), dtype=np.uint8))
    smd = str(tmp_path / "m.smd")
    write_smd(smd, (pos, ids), mats, ["t.bmp"])
    w = h = 48
    passes = 4
    ppm, sums, setup = tmp_path / "o.ppm", tmp_path / "s.bin", tmp_path / "setup.bin"
    out = subprocess.run([EXE, "--smd", smd, str(w), str(h), str(passes), str(ppm), str(sums), str(setup)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    img = np.fromfile(sums, dtype=np.float32).reshape(h, w, 3)
    p = np.fromfile(setup, dtype=np.float32)
    m = load_smd(smd)
    d = SceneDesc()
    d.materials = [dict(mt) for mt in m["materials"]] + [dict(type=abi.MAT_LAMBERT, emissive=(20.0, 20.0, 20.0))]
    d.textures = [__import__("salty_amd").load_bmp(m["texture_paths"][0])]
    d.meshes = [Mesh(m["vertices"], triangle_material=m["triangle_material"])]
    d.shapes = [Shape(abi.SHAPE_MESH, mesh=0),
                Shape(abi.SHAPE_SPHERE, material=2, center=tuple(float(x) for x in p[0:3]), radius=float(p[3]))]
    d.lights = [Light(abi.LIGHT_SHAPE, shape=1)]
    d.camera = Camera(position=tuple(float(x) for x in p[4:7]), target=tuple(float(x) for x in p[7:10]),
                      fov=float(p[10]), aspect=1.0, near_clip=1.0)
    d.environment = (0.05, 0.05, 0.05)
    ref, counts = O.OracleScene(d).render(w, h, spp=passes, max_depth=6, seed=0)
    rel = np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2))
    assert rel < 1e-4, rel
    assert img.mean() > 0.0
