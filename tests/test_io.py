"""Loaders of the reference's file formats (host only, no GPU): BMP and
Radiance HDR images and SMD meshes.  Files written here with known contents
pin the formats (including the reference's quirks: no BMP row padding, the
last BMP texel left unconverted, old- and new-style RGBE run lengths); the
reference's own assets, where the reference tree is present, are checked
against the oracle's restatement of its readers."""
import os
import struct

import numpy as np
import pytest

from oracle import io as OIO
from _formats import mats as _mats, write_bmp as _write_bmp, write_smd as _write_smd

REF_RES = "/root/reference/src/project/res"


def _lib():
    import salty_amd
    return salty_amd


# ---------------------------------------------------------------- BMP ------
def test_bmp_known_texels(tmp_path):
    rng = np.random.default_rng(1)
    px = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)  # 7*3 bytes a row: a padded BMP would differ
    f = str(tmp_path / "a.bmp")
    _write_bmp(f, px)
    got = _lib().load_bmp(f)
    want = px.astype(np.float32) / np.float32(255.0)
    want[-1, -1] = 0.0                                     # never converted by the reference loop
    assert got.shape == (5, 7, 3)
    assert np.array_equal(got, want)
    assert np.array_equal(OIO.read_bmp(f), want)


def test_bmp_rejects(tmp_path):
    import salty_amd
    f = str(tmp_path / "b.bmp")
    _write_bmp(f, np.zeros((2, 2, 3), np.uint8))
    data = bytearray(open(f, "rb").read())
    data[28] = 32                                          # 32-bit: the reference accepts 24 only
    open(f, "wb").write(bytes(data))
    with pytest.raises(salty_amd.S3dgError):
        salty_amd.load_bmp(f)
    open(f, "wb").write(b"PK" + bytes(60))
    with pytest.raises(salty_amd.S3dgError):
        salty_amd.load_bmp(f)


@pytest.mark.skipif(not os.path.exists(REF_RES), reason="the reference's assets are not on this machine")
def test_reference_bmps_match_oracle():
    for name in ("texture/tile.bmp", "texture/brick.bmp", "mesh/dosei/textures/dosei_skin.bmp"):
        f = os.path.join(REF_RES, name)
        got = _lib().load_bmp(f)
        assert got.shape == (512, 512, 3)
        assert np.array_equal(got, OIO.read_bmp(f))


# ---------------------------------------------------------------- HDR ------
def _rgbe_bytes(rng, n):
    """n RGBE texels with a non-zero exponent."""
    t = rng.integers(0, 256, (n, 4), dtype=np.uint8)
    t[:, 3] = rng.integers(120, 140, n)
    return t


def _decode(t):
    out = np.zeros((len(t), 3), np.float32)
    for k, (r, g, b, e) in enumerate(t.astype(int)):
        if e:
            out[k] = (r * 2.0 ** (e - 136), g * 2.0 ** (e - 136), b * 2.0 ** (e - 136))
    return out


def _new_rle(line):
    """Adaptive run-length scanline: 2, 2, width hi, width lo, then per channel
    runs (128 + n, value) and literals (n, values)."""
    w = len(line)
    out = bytearray([2, 2, w >> 8, w & 255])
    for c in range(4):
        v = line[:, c]
        j = 0
        while j < w:
            run = 1
            while j + run < w and run < 127 and v[j + run] == v[j]:
                run += 1
            if run >= 3:
                out += bytes([128 + run, int(v[j])])
                j += run
            else:
                k = j
                while k < w and k - j < 128 and not (k + 2 < w and v[k] == v[k + 1] == v[k + 2]):
                    k += 1
                out += bytes([k - j]) + bytes(int(x) for x in v[j:k])
                j = k
    return bytes(out)


def _write_hdr(path, rows, header_res=None):
    h = len(rows)
    w = header_res or None
    head = b"#?RADIANCE\n# test\nGAMMA=1.0\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n"
    open(path, "wb").write(head + (w or b"-Y %d +X %d\n" % (h, rows.width)) + rows.payload)


class _Rows:
    def __init__(self, width, payload, n):
        self.width, self.payload, self.n = width, payload, n

    def __len__(self):
        return self.n


def test_hdr_flat_and_new_rle(tmp_path):
    rng = np.random.default_rng(2)
    w, h = 19, 4
    t = _rgbe_bytes(rng, w * h).reshape(h, w, 4)
    t[1, 3:11] = t[1, 3]                                   # runs for the encoder
    t[2, :, 3] = 0                                         # exponent 0: black
    payload = b"".join(_new_rle(t[y]) for y in range(h))
    f = str(tmp_path / "n.hdr")
    _write_hdr(f, _Rows(w, payload, h))
    want = _decode(t.reshape(-1, 4)).reshape(h, w, 3)
    got = _lib().load_hdr(f)
    assert np.array_equal(got, want)
    assert np.array_equal(OIO.read_hdr(f), want)
    # the same texels as flat RGBE (a width below 8 always reads the old form)
    t2 = _rgbe_bytes(rng, 5 * 3).reshape(3, 5, 4)
    f2 = str(tmp_path / "f.hdr")
    _write_hdr(f2, _Rows(5, t2.tobytes(), 3))
    assert np.array_equal(_lib().load_hdr(f2), _decode(t2.reshape(-1, 4)).reshape(3, 5, 3))


def test_hdr_old_runs(tmp_path):
    rng = np.random.default_rng(3)
    a, b = _rgbe_bytes(rng, 2)
    # a, then (1,1,1,3): three more a, then (1,1,1,1) continuing the run with
    # shift 8: 256 more a, then b -> width 1 + 3 + 256 + 1 = 261
    payload = bytes(a) + bytes([1, 1, 1, 3]) + bytes([1, 1, 1, 1]) + bytes(b)
    f = str(tmp_path / "o.hdr")
    _write_hdr(f, _Rows(261, payload, 1))
    got = _lib().load_hdr(f)
    want = _decode(np.array([a] * 260 + [b]))
    assert np.array_equal(got.reshape(-1, 3), want)
    assert np.array_equal(OIO.read_hdr(f).reshape(-1, 3), want)


def test_hdr_rejects(tmp_path):
    import salty_amd
    f = str(tmp_path / "x.hdr")
    open(f, "wb").write(b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x00\x00\x00\x00")
    with pytest.raises(salty_amd.S3dgError):
        salty_amd.load_hdr(f)
    # a run before the first texel
    open(f, "wb").write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 2\n\x01\x01\x01\x01\x01\x01\x01\x01")
    with pytest.raises(salty_amd.S3dgError):
        salty_amd.load_hdr(f)
    # truncated
    open(f, "wb").write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n\x05\x05\x05\x80")
    with pytest.raises(salty_amd.S3dgError):
        salty_amd.load_hdr(f)
    # nine consecutive zero-length repeat codes: the run count would be
    # shifted by 32 and more bits
    open(f, "wb").write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 2\n\x05\x05\x05\x80"
                        + b"\x01\x01\x01\x00" * 9 + b"\x06\x06\x06\x80")
    with pytest.raises(salty_amd.S3dgError):
        salty_amd.load_hdr(f)
    with pytest.raises(ValueError):
        OIO.read_hdr(f)


# ---------------------------------------------------------------- SMD ------
def test_smd_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    pos = rng.standard_normal((6, 3, 8)).astype(np.float32)
    ids = np.array([0, 1, 2, 3, 4, -1], np.int32)
    f = str(tmp_path / "m.smd")
    _write_smd(f, (pos, ids), _mats(), ["a.bmp", "sub\\b.bmp"])
    m = _lib().load_smd(f)
    v = m["vertices"]
    assert len(v) == 18
    assert np.array_equal(v["position"], pos[:, :, 0:3].reshape(-1, 3))
    assert np.array_equal(v["normal"], pos[:, :, 3:6].reshape(-1, 3))
    assert np.array_equal(v["texcoord"], pos[:, :, 6:8].reshape(-1, 2))
    assert m["triangle_material"].tolist() == [0, 1, 2, 3, 4, -1]
    assert m["texture_paths"] == [str(tmp_path / "a.bmp"), str(tmp_path / "sub/b.bmp")]
    types = [x["type"] for x in m["materials"]]
    assert types == [0, 1, 2, 3, 4]                        # Lambert, Mirror, Glass, Phong, Plastic
    mt = m["materials"]
    assert np.allclose(mt[0]["diffuse"], (0.5, 0.25, 0.125)) and mt[0]["texture"] == 1
    assert np.allclose(mt[1]["specular"], (0.9, 0.8, 0.7)) and mt[1]["texture"] == -1
    assert mt[2]["ior"] == 1.5 and mt[2]["texture"] == 0
    assert mt[3]["power"] == 40.0
    assert mt[4]["power"] == 20.0 and np.allclose(mt[4]["emissive"], (4, 5, 6)) and mt[4]["texture"] == 0
    o = OIO.read_smd(f)
    assert o["texture_paths"] == m["texture_paths"]
    assert [x["texture"] for x in o["materials"]] == [x["texture"] for x in mt]


def test_smd_rejects(tmp_path):
    import salty_amd
    rng = np.random.default_rng(5)
    pos = rng.standard_normal((2, 3, 8)).astype(np.float32)
    f = str(tmp_path / "r.smd")
    for kwargs, ids in ((dict(version=3), [0, 0]), ({}, [0, 7])):
        _write_smd(f, (pos, np.array(ids, np.int32)), _mats()[:1], ["t.bmp", "u.bmp"], **kwargs)
        with pytest.raises(salty_amd.S3dgError):
            salty_amd.load_smd(f)
    _write_smd(f, (pos, np.array([0, 0], np.int32)), _mats()[:1], ["t.bmp", "u.bmp"])
    data = open(f, "rb").read()
    open(f, "wb").write(data[:-10])                        # truncated
    with pytest.raises(salty_amd.S3dgError):
        salty_amd.load_smd(f)


@pytest.mark.skipif(not os.path.exists(REF_RES), reason="the reference's assets are not on this machine")
def test_reference_smds_match_oracle():
    for name, ntri in (("mesh/dosei/dosei.smd", 19328), ("mesh/paper_cup/paper_cup.smd", 1880)):
        f = os.path.join(REF_RES, name)
        m = _lib().load_smd(f)
        o = OIO.read_smd(f)
        assert len(m["vertices"]) == 3 * ntri
        assert o["bytes_used"] == os.path.getsize(f)          # every byte of the file accounted for
        assert np.array_equal(m["vertices"]["position"], o["positions"])
        assert np.array_equal(m["vertices"]["normal"], o["normals"])
        assert np.array_equal(m["vertices"]["texcoord"], o["texcoords"])
        assert np.array_equal(m["triangle_material"], o["triangle_material"])
        assert m["texture_paths"] == o["texture_paths"]
        for a, b in zip(m["materials"], o["materials"]):
            assert a["type"] == b["type"] and a["texture"] == b["texture"]
            assert np.array_equal(np.float32(a["diffuse"]), np.float32(b["diffuse"]))
        for p in m["texture_paths"]:
            assert _lib().load_bmp(p).shape[2] == 3
