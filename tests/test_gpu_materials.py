"""Every material of the reference (Lambert, Mirror, Glass, Phong, Plastic)
and the ThinLensCamera on a scene of the mesh-scene family (instanced
icospheres over a Plastic ground), GPU against the oracle with the same
counter-keyed streams."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu


def rel_mse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


@pytest.mark.parametrize("lens", [0.0, 0.15])
def test_all_materials_and_thin_lens_match_oracle(ctx, lens):
    import salty_amd
    desc = scenes.mesh_scene(big_level=4, small_level=2, ground=(40, 20), instances=16, materials="all",
                             lens_radius=lens)
    kinds = {m["type"] for m in desc.materials}
    assert {abi.MAT_LAMBERT, abi.MAT_MIRROR, abi.MAT_GLASS, abi.MAT_PHONG, abi.MAT_PLASTIC} <= kinds
    assert desc.camera.type == (abi.CAMERA_THINLENS if lens > 0 else abi.CAMERA_PINHOLE)
    w, h = 112, 64
    sc = salty_amd.Scene(ctx, desc)
    gpu, st = sc.render(w, h, spp=4, max_depth=8, seed=13)
    ref, counts = O.OracleScene(desc).render(w, h, spp=4, max_depth=8, seed=13)
    assert np.isfinite(gpu).all()
    assert rel_mse(gpu, ref) < 1e-4
    # same streams: paths diverge only on last-ulp differences of the
    # transcendental functions (powf in Phong/Plastic, cos/sin)
    assert abs(int(st.extension_rays) - int(counts[0])) <= 2e-3 * counts[0]
    assert abs(int(st.shadow_rays) - int(counts[1])) <= 2e-3 * counts[1]
