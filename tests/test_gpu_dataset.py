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
