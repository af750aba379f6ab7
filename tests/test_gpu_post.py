"""Tonemapping (ToneMapper::Map, s3d_tonemapper.cpp:206-353), capture
(PathTracer::Capture, s3d_pt.cpp:137-176) and non-local means (FilterNLM,
s3d_denoiser.cpp:19-95) on the GPU against the oracle.

The log-average luminance is a sequential float sum in the reference (and the
oracle) and a fixed-order FP64 reduction on the GPU, so the two differ by the
float sum's own rounding (relative 1e-5 at these sizes): the tonemapped
floats are compared to 1e-4 relative, the 8-bit outputs exactly except where
a value sits on a rounding boundary (at most one step, on a handful of
pixels)."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi

pytestmark = pytest.mark.gpu


def _hdr(h, w, seed):
    rng = np.random.default_rng(seed)
    img = rng.gamma(0.8, 0.7, size=(h, w, 3)).astype(np.float32)
    if h * w > 4:
        img[0, 0] = 0.0                        # Lw below FLT_EPSILON: the log term uses eps
        img[-1, -1] = (40.0, 25.0, 60.0)       # a highlight far above the key
        img[h // 2, w // 2] = (0.0, 1e-9, 0.0)
    return img


@pytest.mark.parametrize("kind", [abi.TONEMAP_REINHARD, abi.TONEMAP_UNCHARTED2, abi.TONEMAP_ACES])
@pytest.mark.parametrize("shape", [(1, 1), (37, 53), (256, 256), (9, 1031)])
def test_tonemap_and_capture_match_oracle(ctx, kind, shape):
    img = _hdr(*shape, seed=shape[0] * 7 + kind)
    out, out8 = ctx.tonemap(img, kind)
    ref, _, _ = O.tonemap(img, kind)
    assert np.isfinite(out).all()
    assert np.allclose(out, ref, rtol=1e-4, atol=1e-7)
    ref8 = O.capture(ref)
    diff = np.abs(out8.astype(np.int32) - ref8.astype(np.int32))
    assert diff.max() <= 1 and np.count_nonzero(diff) <= max(2, diff.size // 1000)


def test_capture_of_exact_values(ctx):
    """Values the tonemapper leaves unchanged on their way to capture are
    encoded exactly like the oracle's Capture: Reinhard of a constant image
    maps every pixel to the same value, so the 8-bit image is one colour."""
    img = np.full((8, 8, 3), 0.18, np.float32)
    out, out8 = ctx.tonemap(img, abi.TONEMAP_REINHARD)
    ref, _, _ = O.tonemap(img, abi.TONEMAP_REINHARD)
    assert np.allclose(out, ref, rtol=1e-6)
    assert np.array_equal(out8, O.capture(ref))
    assert len(np.unique(out8.reshape(-1, 3), axis=0)) == 1


# patch radius 2 runs as box sums over 60 x 8 pixel tiles (search radii 6
# and 7 compiled in, others at run time, up to 10); other patch radii run
# per pixel; the shapes cover one tile, partial tiles and many tiles
@pytest.mark.parametrize("rs,rp,coeff", [(6, 2, 0.1), (7, 2, 0.05), (3, 1, 0.4), (6, 2, 1e-6), (10, 2, 0.2),
                                         (2, 2, 0.3)])
@pytest.mark.parametrize("shape", [(48, 40), (5, 3), (17, 70), (70, 130)])
def test_nlm_matches_oracle(ctx, rs, rp, coeff, shape):
    rng = np.random.default_rng(rs * 100 + shape[1])
    h, w = shape
    base = np.linspace(0, 1, h * w, dtype=np.float32).reshape(h, w, 1)
    img = np.concatenate([base + rng.normal(scale=0.08, size=(h, w, 3)).astype(np.float32),
                          np.ones((h, w, 1), np.float32)], -1).astype(np.float32)
    got = ctx.nlm(img, coeff, rs, rp)
    want = O.nlm(img, coeff, rs, rp)
    assert np.isfinite(got).all()
    assert np.allclose(got, want, rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("rs", [6, 7])
def test_nlm_1080p_matches_oracle_on_crops(ctx, rs):
    """NLM at 1920x1080 (4,320 bands of 60 x 8 pixels, one per workgroup).  An
    output pixel depends only on the input within rs + 2 texels, clamped at
    the image's edges, so the oracle run on a crop reproduces the pixels at
    least rs + 2 texels inside it and, where the crop reaches the image's
    edge, up to that edge: the two corners and a window across tile seams
    (columns 60k, rows 8k) are compared."""
    h, w = 1080, 1920
    rng = np.random.default_rng(42 + rs)
    base = np.linspace(0, 1, h * w, dtype=np.float32).reshape(h, w, 1)
    img = np.concatenate([base + rng.normal(scale=0.08, size=(h, w, 3)).astype(np.float32),
                          rng.uniform(0.5, 1.0, (h, w, 1)).astype(np.float32)], -1).astype(np.float32)
    got = ctx.nlm(img, 0.1, rs, 2)
    assert np.isfinite(got).all()
    m = rs + 2
    # (crop rows, crop columns, kept rows and columns inside the crop)
    for y0, y1, x0, x1 in ((0, 70, 0, 130), (h - 70, h, w - 130, w), (300, 400, 900, 1040)):
        want = O.nlm(np.ascontiguousarray(img[y0:y1, x0:x1]), 0.1, rs, 2)
        ky0 = 0 if y0 == 0 else m
        ky1 = (y1 - y0) if y1 == h else (y1 - y0) - m
        kx0 = 0 if x0 == 0 else m
        kx1 = (x1 - x0) if x1 == w else (x1 - x0) - m
        assert np.allclose(got[y0 + ky0:y0 + ky1, x0 + kx0:x0 + kx1], want[ky0:ky1, kx0:kx1], rtol=1e-4, atol=1e-6), \
            (rs, y0, x0)
