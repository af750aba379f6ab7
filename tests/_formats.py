"""Writers of the reference's file formats for tests (our own encoders, so that
files with known contents pin the loaders)."""
import struct

import numpy as np


def write_bmp(path, rgb_u8):
    h, w, _ = rgb_u8.shape
    body = rgb_u8[..., ::-1].reshape(-1).tobytes()        # B, G, R; rows as given; no padding
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, len(body), 2835, 2835, 0, 0)
    head = struct.pack("<HIHHI", 0x4D42, 54 + len(body), 0, 0, 54)
    open(path, "wb").write(head + info + body)


def write_smd(path, tris, mats, textures, version=2):
    """tris: (T, 3, 8) float32 + material ids; mats: list of (type, fields)."""
    body = b""
    for name in textures:
        body += name.encode().ljust(256, b"\0")
    for kind, fields in mats:
        body += struct.pack("<i", kind) + fields
    pos, mat_ids = tris
    for k in range(len(pos)):
        body += pos[k].astype("<f4").tobytes() + struct.pack("<i", int(mat_ids[k]))
    head = b"SMD\0" + struct.pack("<II6I", version, 24, len(pos), len(mats), len(textures), 100, 48, 256)
    open(path, "wb").write(head + body)


def mats():
    f3 = lambda *x: struct.pack("<3f", *x)  # noqa: E731
    return [
        (0, f3(0.5, 0.25, 0.125) + f3(0, 0, 0) + struct.pack("<i", 1)),                       # matte, map 1
        (1, f3(0.9, 0.8, 0.7) + f3(0, 0, 0) + struct.pack("<i", -1)),                          # mirror
        (2, f3(1, 1, 1) + struct.pack("<f", 1.5) + f3(0, 0, 0) + struct.pack("<i", 0)),        # dielectric
        (3, f3(0.3, 0.3, 0.3) + struct.pack("<f", 40.0) + f3(0, 0, 0) + struct.pack("<i", -1)),  # glossy
        (4, f3(0.2, 0.4, 0.6) + f3(0.5, 0.5, 0.5) + struct.pack("<f", 20.0) + f3(4, 5, 6)
         + struct.pack("<ii", 0, 1)),                                                          # plastic
    ]
