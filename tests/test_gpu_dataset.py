"""The denoiser's training pairs rendered as scene sets equal those rendered
scene by scene, bit for bit (each frame of a batched render has the sum its
scene alone would give): noisy ladder, guides and reference."""
import numpy as np
import pytest

from salty_amd import denoise

pytestmark = pytest.mark.gpu


def test_batched_dataset_equals_per_scene(ctx):
    seeds = range(1000, 1005)
    kw = dict(width=32, height=24, levels=(1, 2, 4), spp_ref=8, max_depth=8)
    a = denoise.make_dataset(ctx, seeds, **kw)
    b = denoise.make_dataset_batched(ctx, seeds, group=2, **kw)   # three groups, the last one short
    for x, y in ((a.noisy, b.noisy), (a.guides, b.guides), (a.target, b.target)):
        assert x.shape == y.shape
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.isfinite(b.target).all() and b.target.max() > 0


def test_device_features_equal_host_features(ctx):
    """The denoiser input assembled on the device from a render's buffers
    (sum / spp, albedo, normal, depth) equals the host assembly bit for bit,
    and a DeviceTimer reads a positive time for it."""
    import salty_amd
    from salty_amd import scenes
    w, h, spp = 40, 24, 3
    fr = denoise.FrameRenderer(ctx, scenes.cornell_variant(1001), w, h, 8, seed=5)
    try:
        a, n, d, guide_ms = fr.guides()
        img, _ = fr.mean(spp)
        host = denoise.features(img, a, n, d)
        d_f = ctx.malloc(w * h * 10 * 4)
        timer = salty_amd.DeviceTimer(ctx)
        timer.start()
        ctx.features_device(1, h, w, fr.d_acc, spp, fr.d_alb, fr.d_nrm, fr.d_dep, d_f)
        ms = timer.stop()
        timer.close()
        dev = ctx.download(d_f, (h, w, 10), np.float32)
        ctx.free(d_f)
    finally:
        fr.close()
    assert np.array_equal(dev.view(np.uint32), host.view(np.uint32))
    assert guide_ms > 0.0 and ms > 0.0
