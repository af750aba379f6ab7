"""Renders whose paths fit the device's resident threads trace each path in
one thread, in one launch; the per-bounce wavefront pipeline
(S3DG_RENDER_WAVEFRONT) gives the same sums bit for bit and the same ray
counts, for every material, light kind, camera, texture and instance layout,
and for scene-set batches."""
import numpy as np
import pytest

from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _render_batch(ctx, sset, frames, depth, flags=0):
    bufs, jobs = [], []
    for sc, w, h, spp, seed, fp in frames:
        d = ctx.malloc(w * h * 12)
        ctx.copy_to_device(d, np.zeros(w * h * 3, np.float32))
        bufs.append(d)
        jobs.append((sc, w, h, spp, seed, fp, d))
    st = sset.render_batch_device(jobs, depth, flags=flags)
    out = [ctx.download(d, (f[2], f[1], 3), np.float32) for d, f in zip(bufs, frames)]
    for d in bufs:
        ctx.free(d)
    return out, st


def _scenes():
    return {
        "cornell": scenes.cornell_box(),
        "cornell_variant": scenes.cornell_variant(3),
        "textured_ibl": scenes.textured_cornell(2),
        "mesh_all_materials_thin_lens": scenes.mesh_scene(big_level=4, small_level=2, ground=(40, 20), instances=16,
                                                          materials="all", lens_radius=0.15),
        "mesh_c2_materials": scenes.mesh_scene(big_level=4, small_level=2, ground=(40, 20), instances=8),
    }


@pytest.mark.parametrize("name", list(_scenes()))
@pytest.mark.parametrize("depth,rr", [(8, 0), (12, 3), (1, 0), (0, 0)])
def test_per_path_equals_wavefront(ctx, name, depth, rr):
    import salty_amd
    sc = salty_amd.Scene(ctx, _scenes()[name])
    w, h, spp = 72, 40, 3
    a, sa = sc.render(w, h, spp, depth, seed=21, rr_start_depth=rr, first_pass=5)
    b, sb = sc.render(w, h, spp, depth, seed=21, rr_start_depth=rr, first_pass=5, flags=abi.RENDER_WAVEFRONT)
    assert np.isfinite(a).all()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (sa.extension_rays, sa.shadow_rays, sa.paths) == (sb.extension_rays, sb.shadow_rays, sb.paths)
    if depth == 0:
        assert sa.extension_rays == 0 and not a.any()


def test_per_path_batch_equals_wavefront(ctx):
    import salty_amd
    descs = [scenes.cornell_box(), scenes.batch_scene(11, 3_000), scenes.cornell_variant(5)]
    # (scene, width, height, spp, seed, first_pass)
    frames = [(0, 64, 48, 3, 7, 0), (1, 80, 60, 5, 8, 2), (2, 48, 48, 4, 10, 1), (1, 40, 30, 6, 3, 0)]
    scs = [salty_amd.Scene(ctx, d) for d in descs]
    sset = salty_amd.SceneSet(ctx, scs)
    a, sa = _render_batch(ctx, sset, frames, 6)
    b, sb = _render_batch(ctx, sset, frames, 6, flags=abi.RENDER_WAVEFRONT)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert (sa.extension_rays, sa.shadow_rays) == (sb.extension_rays, sb.shadow_rays)


def test_per_path_repeated_renders_and_sizes(ctx):
    """Alternating sizes and modes reuse the captured launch graph only for
    identical launches; each render equals its wavefront twin."""
    import salty_amd
    sc = salty_amd.Scene(ctx, scenes.cornell_box())
    for w, h, spp in ((64, 64, 1), (96, 80, 2), (64, 64, 1), (256, 256, 2), (33, 17, 5)):
        a, sa = sc.render(w, h, spp, 8, seed=4)
        b, sb = sc.render(w, h, spp, 8, seed=4, flags=abi.RENDER_WAVEFRONT)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert (sa.extension_rays, sa.shadow_rays) == (sb.extension_rays, sb.shadow_rays)


def test_forced_per_path_with_several_paths_per_thread(ctx):
    """2 M paths: more than the per-path grid holds, so its threads take
    several paths each (grid-stride); still the wavefront's sums."""
    import salty_amd
    sc = salty_amd.Scene(ctx, scenes.cornell_box())
    w, h, spp = 512, 512, 8
    a, sa = sc.render(w, h, spp, 8, seed=6, flags=abi.RENDER_PER_PATH)
    b, sb = sc.render(w, h, spp, 8, seed=6, flags=abi.RENDER_WAVEFRONT)
    c, sc_ = sc.render(w, h, spp, 8, seed=6)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(c.view(np.uint32), b.view(np.uint32))
    assert (sa.extension_rays, sa.shadow_rays) == (sb.extension_rays, sb.shadow_rays) == \
        (sc_.extension_rays, sc_.shadow_rays)


def test_per_path_on_the_full_mesh_scene(ctx):
    """The 1080p benchmark scene's geometry at a small size: two-level trees
    deep enough that a traversal stack can outgrow its LDS part and continue
    in the global overflow area."""
    import salty_amd
    sc = salty_amd.Scene(ctx, scenes.mesh_scene())
    a, sa = sc.render(96, 54, 2, 12, seed=8)
    b, sb = sc.render(96, 54, 2, 12, seed=8, flags=abi.RENDER_WAVEFRONT)
    assert np.isfinite(a).all() and a.any()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (sa.extension_rays, sa.shadow_rays) == (sb.extension_rays, sb.shadow_rays)
