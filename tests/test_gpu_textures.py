"""Textured materials and the IBL (TexturedMaterial, Texture::SampleColor,
Scene::SampleIBL) and SMD scenes on the GPU, against the oracle's
restatement; a texture of ones changes nothing bit for bit; scene sets carry
textures with their members."""
import numpy as np
import pytest

from _formats import write_bmp, write_smd
from oracle import oracle as O
from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu


def rel_mse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


def _render_both(ctx, desc, w, h, spp, depth, seed):
    import salty_amd
    sc = salty_amd.Scene(ctx, desc)
    gpu, st = sc.render(w, h, spp=spp, max_depth=depth, seed=seed)
    ref, counts = O.OracleScene(desc).render(w, h, spp=spp, max_depth=depth, seed=seed)
    return gpu, st, ref, counts


def test_textured_cornell_matches_oracle(ctx):
    desc = scenes.textured_cornell()
    gpu, st, ref, counts = _render_both(ctx, desc, 96, 96, 4, 8, 7)
    assert np.isfinite(gpu).all()
    assert rel_mse(gpu, ref) < 1e-4
    assert abs(int(st.extension_rays) - int(counts[0])) <= 1e-3 * counts[0]
    assert abs(int(st.shadow_rays) - int(counts[1])) <= 1e-3 * counts[1]
    # the textures and the IBL do show: the image differs from the plain box
    plain, _, _, _ = _render_both(ctx, scenes.cornell_box(), 96, 96, 4, 8, 7)
    assert rel_mse(gpu, plain) > 1e-2


def test_texture_of_ones_is_exact(ctx):
    """x * 1.0f is exact: every material textured with ones renders the
    untextured image bit for bit."""
    import salty_amd
    base = scenes.cornell_box()
    tex = scenes.cornell_box()
    tex.textures = [np.ones((3, 5, 3), np.float32)]
    for m in tex.materials:
        m["texture"] = 0
    a, sa = salty_amd.Scene(ctx, base).render(64, 48, spp=3, max_depth=6, seed=9)
    b, sb = salty_amd.Scene(ctx, tex).render(64, 48, spp=3, max_depth=6, seed=9)
    assert np.array_equal(a, b)
    assert (sa.extension_rays, sa.shadow_rays) == (sb.extension_rays, sb.shadow_rays)


def test_smd_scene_matches_oracle(ctx, tmp_path):
    """A mesh file written here (every material kind, a BMP colour map, a
    stored texture name with a Windows separator) through the SMD loader and
    the GPU path tracer, against the oracle."""
    rng = np.random.default_rng(12)
    # a bumpy 20 x 20 grid of 800 triangles with texture coordinates over [0, 2]
    n = 20
    xs = np.linspace(-1, 1, n + 1)
    gx, gz = np.meshgrid(xs, xs)
    gy = 0.15 * np.sin(3 * gx) * np.cos(2 * gz)
    tris = []
    for i in range(n):
        for j in range(n):
            a, b = (i, j), (i, j + 1)
            c, d = (i + 1, j + 1), (i + 1, j)
            for q in ((a, b, c), (a, c, d)):
                vs = []
                for (r, k) in q:
                    vs.append([gx[r, k], gy[r, k], gz[r, k], 0.0, 1.0, 0.0, gx[r, k] + 1.0, gz[r, k] + 1.0])
                tris.append(vs)
    pos = np.array(tris, np.float32)
    ids = rng.integers(0, 5, len(pos)).astype(np.int32)
    import struct
    f3 = lambda *x: struct.pack("<3f", *x)  # noqa: E731
    mats = [(0, f3(0.8, 0.8, 0.8) + f3(0, 0, 0) + struct.pack("<i", 0)),          # textured matte
            (1, f3(0.9, 0.9, 0.9) + f3(0, 0, 0) + struct.pack("<i", -1)),         # mirror
            (2, f3(1, 1, 1) + struct.pack("<f", 1.5) + f3(0, 0, 0) + struct.pack("<i", -1)),
            (3, f3(0.5, 0.5, 0.5) + struct.pack("<f", 30.0) + f3(0, 0, 0) + struct.pack("<i", 0)),
            (4, f3(0.3, 0.5, 0.2) + f3(0.4, 0.4, 0.4) + struct.pack("<f", 15.0) + f3(0, 0, 0)
             + struct.pack("<ii", 0, -1))]
    (tmp_path / "tex").mkdir()
    write_bmp(str(tmp_path / "tex" / "c.bmp"), rng.integers(0, 256, (9, 13, 3), dtype=np.uint8))
    f = str(tmp_path / "grid.smd")
    write_smd(f, (pos, ids), mats, ["tex\\c.bmp"])
    desc = scenes.smd_scene(f)
    assert len(desc.textures) == 1 and desc.textures[0].shape == (9, 13, 3)
    gpu, st, ref, counts = _render_both(ctx, desc, 80, 80, 4, 6, 3)
    assert np.isfinite(gpu).all()
    assert rel_mse(gpu, ref) < 1e-4
    assert abs(int(st.extension_rays) - int(counts[0])) <= 2e-3 * counts[0]


def test_scene_set_carries_textures(ctx):
    """Textures, texture indices and the IBL are rebased with each member of a
    scene set: every frame of the batch equals its single render."""
    import salty_amd
    descs = [scenes.textured_cornell(0), scenes.cornell_box(), scenes.textured_cornell(1)]
    members = [salty_amd.Scene(ctx, d) for d in descs]
    batch = salty_amd.SceneSet(ctx, members)
    frames, bufs = [], []
    for k, (w, h) in enumerate(((48, 40), (32, 32), (40, 48))):
        d = ctx.malloc(w * h * 12)
        ctx.copy_to_device(d, np.zeros(w * h * 3, np.float32))
        frames.append((k, w, h, 3, 5 + k, 0, d))
        bufs.append((d, w, h))
    batch.render_batch_device(frames, max_depth=6)
    for k, (d, w, h) in enumerate(bufs):
        got = ctx.download(d, (h, w, 3), np.float32)
        want, _ = members[k].render(w, h, spp=3, max_depth=6, seed=5 + k)
        assert np.array_equal(got, want.reshape(h, w, 3)), f"member {k}"
        ctx.free(d)
