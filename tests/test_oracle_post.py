"""Oracle tonemapping, capture and NLM against closed forms and an independent
vectorised float64 restatement of the published non-local-means filter."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi


def test_log_average_matches_float64():
    rng = np.random.default_rng(0)
    img = rng.gamma(1.0, 0.5, size=(40, 50, 3)).astype(np.float32)
    img[3, 4] = 0.0                      # Lw < FLT_EPSILON -> eps
    _, ave, mx = O.tonemap(img, abi.TONEMAP_ACES)
    y = img.astype(np.float64) @ np.array([0.299, 0.587, 0.114])
    y = np.where(y < np.finfo(np.float32).eps, 1e-5, y)
    want = math.exp(np.mean(np.log(1e-5 + y)))
    assert abs(ave / want - 1) < 1e-5
    assert abs(mx - y.max()) <= 1e-6 * y.max()


def test_aces_limits():
    """ACESFilm(x) = x(a x f/g + b) / (x f/g (c x f + d) + e): 0 at 0, slope
    b/e = 0.03/0.14 near 0, saturates to 1 for large x; monotone in between."""
    x = np.zeros((1, 4, 3), np.float32)
    ave = 1.0                            # one bright pixel fixes the key
    img = np.full((1, 1, 3), 0.0, np.float32)
    out, _, _ = O.tonemap(np.zeros((2, 2, 3), np.float32), abi.TONEMAP_ACES)
    assert np.all(out == 0.0)
    ramp = np.logspace(-4, 3, 200, dtype=np.float32)
    img = np.stack([ramp, ramp, ramp], -1)[None]
    out, ave, _ = O.tonemap(img, abi.TONEMAP_ACES)
    v = out[0, :, 0]
    assert np.all(np.diff(v) >= 0) and v[-1] == 1.0
    small = ramp[0] * 0.27 / ave
    assert abs(v[0] / small - 0.03 / 0.14) < 1e-3


def test_reinhard_maps_the_key_and_white():
    rng = np.random.default_rng(1)
    img = rng.gamma(2.0, 0.3, size=(16, 16, 3)).astype(np.float32)
    out, ave, mx = O.tonemap(img, abi.TONEMAP_REINHARD)
    L = img * (0.18 / ave)
    w2 = (mx * 0.18 / ave) ** 2
    assert np.allclose(out, L * (1 + L / w2) / (1 + L), rtol=1e-5)


def test_capture_uses_the_nonstandard_srgb_curve():
    """Capture encodes pow(1.055 x, 1/2.4) - 0.055 above 0.0031308 (the
    standard curve is 1.055 x^(1/2.4) - 0.055): 0.5 -> 181 (standard: 188);
    white 1.0 -> 1.055^(1/2.4) - 0.055 = 0.96756 -> 247 (standard: 255);
    12.92 x below 0.0031308; clamps to [0, 1]."""
    img = np.array([[[0.5, 0.002, 1.5], [-1.0, 1.0, 0.0]]], np.float32)
    out = O.capture(img)
    assert out.tolist() == [[[181, 7, 247], [0, 247, 0]]]


def nlm_reference(img, coeff, rs, rp):
    """Non-local means (Buades, Coll & Morel 2005) with the reference's
    parameters: patch distance = sum over the (2rp+1)^2 patch of squared
    RGBA differences, weight exp(-max(d2 - 2 sigma^2, 0) / h^2), h = sigma =
    max(1e-4, coeff), clamp-to-edge sampling.  Vectorised in float64."""
    h = max(1e-4, coeff)
    H, W, _ = img.shape
    pad = rs + rp
    P = np.pad(img.astype(np.float64), ((pad, pad), (pad, pad), (0, 0)), mode="edge")
    out = np.zeros((H, W, 4))
    wsum = np.zeros((H, W))

    def window(dy, dx):
        return P[pad + dy:pad + dy + H, pad + dx:pad + dx + W]

    for oy in range(-rs, rs + 1):
        for ox in range(-rs, rs + 1):
            d2 = np.zeros((H, W))
            for py in range(-rp, rp + 1):
                for px in range(-rp, rp + 1):
                    d2 += ((window(py, px) - window(oy + py, ox + px)) ** 2).sum(-1)
            w = np.exp(-np.maximum(d2 - 2 * h * h, 0) / (h * h))
            out += w[..., None] * window(oy, ox)
            wsum += w
    return out / wsum[..., None]


@pytest.mark.parametrize("coeff,rs,rp", [(0.2, 6, 2), (0.05, 3, 1), (0.5, 7, 2)])
def test_nlm_matches_vectorised_float64(coeff, rs, rp):
    rng = np.random.default_rng(2)
    base = np.linspace(0, 1, 24)[None, :, None] * np.ones((20, 24, 4))
    img = (base + rng.normal(scale=0.1, size=base.shape)).astype(np.float32)
    got = O.nlm(img, coeff, rs, rp)
    want = nlm_reference(img, coeff, rs, rp)
    assert np.allclose(got, want, rtol=2e-4, atol=2e-5)


def test_nlm_flat_image_is_identity():
    img = np.full((9, 7, 4), 0.3, np.float32)
    assert np.allclose(O.nlm(img, 0.1), img, rtol=2e-5)
