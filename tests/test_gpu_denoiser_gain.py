"""What the trained network gains over the methods without it.

The shipped weights (salty_amd/weights/unet_cornell.npz, trained on seeds
1000-2999) on held-out Cornell-style scenes (seed 0, the canonical box, and
seeds 4 and 5), rendered at 256x256, depth 8: from a 4-spp input plus the
guide buffers the U-Net reaches relMSE <= 2e-3 against a 65,536-spp ground
truth, below NLM on the same input at its best strength (FilterNLM's weights,
src/src/s3d_denoiser.cpp:19-95, search radius 7, patch radius 2, swept on
radiance and on log radiance) and below a plain 256-spp render, which costs
more device time than the whole denoised path.
"""
import pytest

from salty_amd import denoise

pytestmark = pytest.mark.gpu

SEEDS = (0, 4, 5)


@pytest.fixture(scope="module")
def rows(ctx):
    import salty_amd
    net = salty_amd.Denoiser(ctx, seed=0)
    assert denoise.load_weights(net), "shipped weights missing"
    out = denoise.evaluate_heldout(ctx, net, SEEDS, spp_in=4, gt_spp=65536, plain_spp=256)
    net.close()
    for r in out:
        print(r)
    return out


def test_learned_reaches_target_from_4spp(rows):
    for r in rows:
        assert r["learned_relmse"] <= 2e-3, r
        assert r["learned_relmse"] < 0.05 * r["input_relmse"], r


def test_learned_beats_nlm_at_its_best_strength(rows):
    for r in rows:
        # the sweep reaches a strength at which NLM does filter its input
        assert r["nlm_relmse"] < r["input_relmse"], r
        assert r["learned_relmse"] < r["nlm_relmse"], r


def test_learned_beats_plain_rendering_at_more_device_time(rows):
    for r in rows:
        learned_ms = r["input_ms"] + r["guides_ms"] + r["inference_ms"]
        assert r["plain_ms"] > learned_ms, r
        assert r["learned_relmse"] < r["plain_relmse"], r
