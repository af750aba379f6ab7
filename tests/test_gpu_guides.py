"""Denoiser guide buffers (albedo, shading normal, hit distance of the
camera ray; zero where it misses) on the GPU against the oracle's, for
plain, textured, instanced and thin-lens scenes."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import scenes

pytestmark = pytest.mark.gpu


def _cases():
    return {
        "cornell": scenes.cornell_box(),
        "textured_ibl": scenes.textured_cornell(1),
        "mesh_all_materials_thin_lens": scenes.mesh_scene(big_level=4, small_level=2, ground=(40, 20), instances=12,
                                                          materials="all", lens_radius=0.15),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_guides_match_oracle(ctx, name):
    import salty_amd
    from salty_amd import abi
    desc = _cases()[name]
    w, h, seed = 80, 48, 5
    sc = salty_amd.Scene(ctx, desc)
    d_a, d_n, d_d = ctx.malloc(w * h * 12), ctx.malloc(w * h * 12), ctx.malloc(w * h * 4)
    try:
        sc.guides_device(abi.RenderConfig(w, h, 1, 1, seed, 0, 0, 0, 0), d_a, d_n, d_d)
        a = ctx.download(d_a, (h, w, 3), np.float32)
        n = ctx.download(d_n, (h, w, 3), np.float32)
        d = ctx.download(d_d, (h, w), np.float32)
    finally:
        for p in (d_a, d_n, d_d):
            ctx.free(p)
    ra, rn, rd = O.OracleScene(desc).guides(w, h, seed=seed)
    # same hits: depth and normals agree to the last bits but for the
    # transcendental functions of a thin-lens sample or a texture lookup
    assert np.array_equal(d > 0, rd > 0)
    assert np.allclose(d, rd, rtol=1e-5, atol=1e-6)
    assert np.allclose(n, rn, atol=1e-5)
    assert np.allclose(a, ra, atol=1e-5)
    exact = np.mean(d == rd)
    assert exact > 0.99, exact
