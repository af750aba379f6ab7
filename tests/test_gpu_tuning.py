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


@pytest.mark.parametrize("rr_start", [0, 3])
def test_tail_launch_changes_nothing(ctx, rr_start):
    """The tail launch (remaining bounces thread per path once the
    extension queue is short) gives the pipeline's sums and ray counts bit
    for bit, wherever it takes over: never, after a few bounces, or early;
    also for several batches and for a scene set."""
    import salty_amd
    start = ctx.get_tuning()
    desc = scenes.mesh_scene(big_level=4, small_level=2, ground=(40, 20), instances=8, materials="all")
    sc = salty_amd.Scene(ctx, desc)
    try:
        outs = []
        for tail in (0, 50, 2000, 0xFFFFFFFF, 1 << 30):
            t = ctx.get_tuning()
            t.tail_paths = tail
            ctx.set_tuning(t)
            for flight in (0, 2):
                img, st = sc.render(96, 64, spp=5, max_depth=10, seed=4, rr_start_depth=rr_start,
                                    flags=abi.RENDER_WAVEFRONT, passes_in_flight=flight)
                outs.append((img, st.extension_rays, st.shadow_rays))
        for img, ne, ns in outs[1:]:
            assert np.array_equal(img, outs[0][0])
            assert (ne, ns) == outs[0][1:]
        # instrumented renders (no tail launch) agree too
        img, st = sc.render(96, 64, spp=5, max_depth=10, seed=4, rr_start_depth=rr_start,
                            flags=abi.RENDER_TIME_KERNELS)
        assert np.array_equal(img, outs[0][0]) and (st.extension_rays, st.shadow_rays) == outs[0][1:]
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
