"""Oracle (test infrastructure only): numpy restatements of the reference's
file readers, used to check the library's loaders.

* ``read_bmp``: LoadFromBMP (src/src/s3d_texture.cpp:1029-1077).
* ``read_hdr``: LoadFromHDR / ReadHdrHeader / ReadColor / ReadOldColors /
  RGBEToVec3 (src/src/s3d_texture.cpp:587-850).
* ``read_smd``: Mesh::LoadFromFile (src/src/s3d_mesh.cpp:250-465) with the SMD
  structs of s3d_mesh.cpp:27-177.

Where the reference reads past a buffer (a BMP texel it never converts, a run
before the first HDR texel, a material index past the table) these return
what the library documents: 0 for the unconverted texel, an error otherwise.
"""
from __future__ import annotations

import math
import struct

import numpy as np

MAT_LAMBERT, MAT_MIRROR, MAT_GLASS, MAT_PHONG, MAT_PLASTIC = range(5)


def read_bmp(path: str) -> np.ndarray:
    data = open(path, "rb").read()
    if len(data) < 54 or data[:2] != b"BM":
        raise ValueError("not a BMP")
    width, height = struct.unpack_from("<ii", data, 18)
    bits = struct.unpack_from("<H", data, 28)[0]
    if bits != 24 or width <= 0 or height <= 0:
        raise ValueError("unsupported BMP")
    n = width * height
    texels = np.zeros(n * 3, np.uint8)
    body = np.frombuffer(data, np.uint8, offset=54)[: n * 3]   # no OffBits, no row padding
    texels[: len(body)] = body
    out = np.zeros((n, 3), np.float32)
    bgr = texels.reshape(n, 3).astype(np.float32)
    # loop i < size - 3: the last texel is never converted
    out[: n - 1, 0] = bgr[: n - 1, 2] / np.float32(255.0)
    out[: n - 1, 1] = bgr[: n - 1, 1] / np.float32(255.0)
    out[: n - 1, 2] = bgr[: n - 1, 0] / np.float32(255.0)
    return out.reshape(height, width, 3)


class _Reader:
    def __init__(self, data: bytes, pos: int):
        self.d, self.p = data, pos

    def getc(self) -> int:
        if self.p >= len(self.d):
            self.p += 1
            return -1
        c = self.d[self.p]
        self.p += 1
        return c

    def eof(self) -> bool:
        return self.p > len(self.d)


def _old_line(r: _Reader, line: list, pos: int, count: int) -> None:
    shift = 0
    while pos < count:
        px = [r.getc() & 0xFF for _ in range(4)]
        if r.eof():
            raise ValueError("truncated HDR")
        if px[0] == 1 and px[1] == 1 and px[2] == 1:
            if pos == 0:
                raise ValueError("run before the first texel")
            if shift >= 32:
                raise ValueError("run count past 32 bits")
            rep = px[3] << shift
            if pos + rep > count:
                raise ValueError("run past the line")
            for _ in range(rep):
                line[pos] = list(line[pos - 1])
                pos += 1
            shift += 8
        else:
            line[pos] = px
            pos += 1
            shift = 0


def _line(r: _Reader, count: int) -> list:
    line = [[0, 0, 0, 0] for _ in range(count)]
    if count < 8 or count > 0x7FFF:
        _old_line(r, line, 0, count)
        return line
    i = r.getc()
    if i < 0:
        raise ValueError("truncated HDR")
    if i != 2:
        r.p -= 1
        _old_line(r, line, 0, count)
        return line
    g, b = r.getc(), r.getc()
    i = r.getc()
    if i < 0:
        raise ValueError("truncated HDR")
    if g != 2 or (b & 128):
        line[0] = [2, g & 0xFF, b & 0xFF, i]
        _old_line(r, line, 1, count)
        return line
    if (b << 8 | i) != count:
        raise ValueError("scanline width mismatch")
    for c in range(4):
        j = 0
        while j < count:
            code = r.getc()
            if code < 0:
                raise ValueError("truncated HDR")
            if code > 128:
                code &= 127
                val = r.getc() & 0xFF
                if j + code > count:
                    raise ValueError("run past the line")
                for _ in range(code):
                    line[j][c] = val
                    j += 1
            else:
                if j + code > count:
                    raise ValueError("run past the line")
                for _ in range(code):
                    line[j][c] = r.getc() & 0xFF
                    j += 1
    if r.eof():
        raise ValueError("truncated HDR")
    return line


def read_hdr(path: str) -> np.ndarray:
    data = open(path, "rb").read()
    if data[:2] != b"#?":
        raise ValueError("not a Radiance file")
    pos = 2

    def fgets():
        nonlocal pos
        if pos >= len(data):
            return None
        end = data.find(b"\n", pos)
        end = len(data) if end < 0 else end + 1
        s = data[pos:end]
        pos = end
        return s.decode("latin-1")

    valid = False
    while True:
        s = fgets()
        if s is None or s.startswith("\n"):
            break
        if s.startswith("#") or s.startswith("GAMMA=") or s.startswith("EXPOSURE="):
            continue
        if s == "FORMAT=32-bit_rle_rgbe\n":
            valid = True
    if not valid:
        raise ValueError("no FORMAT line")
    s = fgets() or ""
    parts = s.split()
    if len(parts) >= 4 and parts[0] == "-Y" and parts[2] == "+X":
        height, width = int(parts[1]), int(parts[3])
    elif len(parts) >= 4 and parts[0] == "+X" and parts[2] == "-Y":
        width, height = int(parts[1]), int(parts[3])
    else:
        raise ValueError("unsupported resolution line")
    r = _Reader(data, pos)
    out = np.zeros((height, width, 3), np.float32)
    for y in range(height):
        line = _line(r, width)
        for x, (rr, gg, bb, e) in enumerate(line):
            if e:
                f = math.ldexp(1.0, e - 136)     # RGBEToVec3, double then float
                out[y, x] = (np.float32(rr * f), np.float32(gg * f), np.float32(bb * f))
    return out


_MAT_SIZES = {0: 28, 1: 28, 2: 32, 3: 32, 4: 48}


def read_smd(path: str) -> dict:
    data = open(path, "rb").read()
    if data[:4] != b"SMD\0":
        raise ValueError("bad magic")
    version, dhs, ntri, nmat, ntex, _, _, _ = struct.unpack_from("<8I", data, 4)
    if version != 2 or dhs != 24 or ntri == 0 or nmat == 0:
        raise ValueError("bad SMD header")
    pos = 36
    folder = path[: path.rfind("/")] if "/" in path else ""
    paths = []
    for _ in range(ntex):
        name = data[pos:pos + 256].split(b"\0")[0].decode("latin-1").replace("\\", "/")
        paths.append(folder + "/" + name)
        pos += 256
    mats = []
    for _ in range(nmat):
        (kind,) = struct.unpack_from("<i", data, pos)
        pos += 4
        if kind not in _MAT_SIZES:
            raise ValueError("unknown material type")
        f = data[pos:pos + _MAT_SIZES[kind]]
        if len(f) != _MAT_SIZES[kind]:
            raise ValueError("truncated SMD")
        pos += _MAT_SIZES[kind]
        m = dict(diffuse=(0.0,) * 3, specular=(0.0,) * 3, ior=0.0, power=0.0)
        if kind in (0, 1):
            c = struct.unpack_from("<3f", f, 0)
            e = struct.unpack_from("<3f", f, 12)
            tex = struct.unpack_from("<i", f, 24)[0]
            m.update(type=MAT_LAMBERT if kind == 0 else MAT_MIRROR)
            m["diffuse" if kind == 0 else "specular"] = c
        elif kind in (2, 3):
            c = struct.unpack_from("<3f", f, 0)
            x = struct.unpack_from("<f", f, 12)[0]
            e = struct.unpack_from("<3f", f, 16)
            tex = struct.unpack_from("<i", f, 28)[0]
            m.update(type=MAT_GLASS if kind == 2 else MAT_PHONG, specular=c)
            m["ior" if kind == 2 else "power"] = x
        else:
            d = struct.unpack_from("<3f", f, 0)
            sp = struct.unpack_from("<3f", f, 12)
            x = struct.unpack_from("<f", f, 24)[0]
            e = struct.unpack_from("<3f", f, 28)
            tex = struct.unpack_from("<i", f, 40)[0]          # DiffuseMap; SpecularMap unused
            m.update(type=MAT_PLASTIC, diffuse=d, specular=sp, power=x)
        if tex >= ntex:
            raise ValueError("texture index out of range")
        m["emissive"] = e
        m["texture"] = tex if tex >= 0 else -1
        mats.append(m)
    need = pos + 100 * ntri
    if len(data) < need:
        raise ValueError("truncated SMD")
    tri = np.frombuffer(data, np.dtype([("v", np.float32, (3, 8)), ("mat", np.int32)]), count=ntri, offset=pos)
    if (tri["mat"] >= nmat).any():
        raise ValueError("material index out of range")
    vert = tri["v"].reshape(-1, 8)
    return dict(positions=vert[:, 0:3].copy(), normals=vert[:, 3:6].copy(), texcoords=vert[:, 6:8].copy(),
                triangle_material=np.where(tri["mat"] < 0, -1, tri["mat"]).astype(np.int32),
                materials=mats, texture_paths=paths, bytes_used=need)

