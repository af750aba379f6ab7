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


@pytest.mark.parametrize("rs,rp,coeff", [(6, 2, 0.1), (7, 2, 0.05), (3, 1, 0.4), (6, 2, 1e-6)])
@pytest.mark.parametrize("shape", [(48, 40), (5, 3), (17, 70)])
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
