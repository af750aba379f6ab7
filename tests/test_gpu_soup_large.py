"""Batch traversal at the full 4 M-triangle soup size: closest hits of a ray
subset are bit-identical to the oracle's reference traversal, and over the
whole batch any-hit agrees with closest-hit (hit iff closest-hit finds a hit,
any-hit distance within [tmin, tmax] and not below the closest one)."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu

TRIS = 4 << 20
RAYS = 1 << 20


def test_full_size_soup(ctx):
    import salty_amd
    desc = scenes.triangle_soup(TRIS, seed=0)
    rays = scenes.soup_rays(RAYS, seed=1)
    sc = salty_amd.Scene(ctx, desc)
    closest = sc.trace(rays, abi.TRACE_CLOSEST)
    anyh = sc.trace(rays, abi.TRACE_ANY)
    hit = closest["prim"] >= 0
    assert 0.05 < hit.mean() < 0.95
    assert np.array_equal(anyh["prim"] >= 0, hit)
    assert np.all(anyh["t"][hit] >= closest["t"][hit])
    assert np.all(anyh["t"][hit] >= rays["tmin"][hit])
    rng = np.random.default_rng(2)
    sub = np.sort(rng.choice(RAYS, 2048, replace=False))
    ref = O.OracleMesh(desc.meshes[0].vertices, abi.BVH8).trace(rays[sub])
    got = closest[sub]
    assert np.array_equal(got["prim"], ref["prim"])
    h = ref["prim"] >= 0
    for f in ("t", "u", "v"):
        assert np.array_equal(got[f][h].view(np.uint32), ref[f][h].view(np.uint32))
