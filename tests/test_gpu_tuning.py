"""Kernel tuning changes speed only: every traversal variant and threshold
gives the same hits and the same render, and the autotuner leaves a valid
setting in the context."""
import numpy as np
import pytest

from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _settings():
    return [abi.Tuning(v, r, b, 0, c, 4) for v, r, b, c in ((0, 32, 16, 256), (1, 32, 16, 64), (2, 32, 16, 4096),
                                                         (1, 1, 1, 100), (1, 64, 65, 256), (2, 7, 40, 777))]


def test_variants_trace_identically(ctx):
    import salty_amd
    start = ctx.get_tuning()
    sc = salty_amd.Scene(ctx, scenes.triangle_soup(30_000, seed=8))
    rays = scenes.soup_rays(50_000, seed=9)
    try:
        ref = None
        for t in _settings():
            ctx.set_tuning(t)
            for mode in (abi.TRACE_CLOSEST, abi.TRACE_ANY):
                h = sc.trace(rays, mode)
                if mode == abi.TRACE_ANY:
                    continue
                if ref is None:
                    ref = h
                assert np.array_equal(h.view(np.uint32), ref.view(np.uint32))
    finally:
        ctx.set_tuning(start)


def test_variants_render_identically(ctx):
    import salty_amd
    start = ctx.get_tuning()
    sc = salty_amd.Scene(ctx, scenes.mesh_scene(big_level=4, small_level=2, ground=(40, 20), instances=8))
    try:
        ref = None
        for t in _settings():
            ctx.set_tuning(t)
            # the small render runs thread per path unless the wavefront pipeline is asked for
            for flags in (0, abi.RENDER_WAVEFRONT):
                img, st = sc.render(64, 48, spp=3, max_depth=6, seed=2, flags=flags)
                if ref is None:
                    ref = (img, st.extension_rays, st.shadow_rays)
                assert np.array_equal(img, ref[0]) and (st.extension_rays, st.shadow_rays) == ref[1:]
    finally:
        ctx.set_tuning(start)


def test_autotune_keeps_a_valid_setting(ctx):
    import salty_amd
    start = ctx.get_tuning()
    sc = salty_amd.Scene(ctx, scenes.cornell_box())
    try:
        cfg = abi.RenderConfig(128, 128, 6, 4, 1, 0, 0, 0, 0)
        t, ms = sc.autotune(cfg)
        assert ms > 0 and 0 <= t.traversal_variant < 3 and 1 <= t.refill <= 64 and 1 <= t.tri_batch <= 65
        now = ctx.get_tuning()
        assert (now.traversal_variant, now.refill, now.tri_batch) == (t.traversal_variant, t.refill, t.tri_batch)
    finally:
        ctx.set_tuning(start)
