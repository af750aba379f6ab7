"""s3dg_tonemap_device_noalloc (reduction scratch in a file-scope device
array) against s3dg_tonemap_device (scratch the context allocates), through
the C ABI, for two calls in a row, the second on a smaller image: float and
8-bit outputs identical."""
import ctypes as C

import numpy as np
import pytest

from salty_amd import abi

pytestmark = pytest.mark.gpu


def _run(ctx, fn, kind, img):
    h, w = img.shape[:2]
    d_in = ctx.upload(img)
    d_out = ctx.malloc(img.nbytes)
    d_8 = ctx.malloc(img.size)
    try:
        ctx.check(fn(ctx.handle, kind, w, h, C.c_void_p(d_in), C.c_void_p(d_out), C.c_void_p(d_8)), fn.__name__)
        ctx.synchronize()
        return ctx.download(d_out, img.shape, np.float32), ctx.download(d_8, img.shape, np.uint8)
    finally:
        for p in (d_in, d_out, d_8):
            ctx.free(p)


@pytest.mark.parametrize("kind", [abi.TONEMAP_REINHARD, abi.TONEMAP_UNCHARTED2, abi.TONEMAP_ACES])
def test_noalloc_tonemap_matches_allocating_entry_point(ctx, kind):
    rng = np.random.default_rng(kind)
    big = rng.gamma(1.0, 0.8, size=(1080, 1920, 3)).astype(np.float32)     # 8,100 blocks -> 1,024 partials
    small = rng.gamma(2.0, 0.3, size=(37, 53, 3)).astype(np.float32)       # 8 partials: stale ones must not count
    lib = ctx.lib
    for img in (big, small):
        want = _run(ctx, lib.s3dg_tonemap_device, kind, img)
        got = _run(ctx, lib.s3dg_tonemap_device_noalloc, kind, img)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])


def test_noalloc_tonemap_rejects_oversized_images(ctx):
    rc = ctx.lib.s3dg_tonemap_device_noalloc(ctx.handle, abi.TONEMAP_ACES, 1 << 15, 1 << 14, C.c_void_p(16),
                                             None, None)
    assert rc == abi.ERR_INVALID
