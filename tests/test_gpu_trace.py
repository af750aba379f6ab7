"""Traversal parity: GPU closest/any hit against the oracle's reference
traversal (exact slab tests, no clipping, depth-first order)."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _check_closest(gpu, ref):
    assert np.array_equal(gpu["prim"], ref["prim"]), f"{(gpu['prim'] != ref['prim']).sum()} prim mismatches"
    hit = ref["prim"] >= 0
    # same arithmetic, same operands: t and barycentrics carry the same bits
    assert np.array_equal(gpu["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32))
    assert np.array_equal(gpu["u"][hit].view(np.uint32), ref["u"][hit].view(np.uint32))
    assert np.array_equal(gpu["v"][hit].view(np.uint32), ref["v"][hit].view(np.uint32))


@pytest.mark.parametrize("kind", [abi.BVH8, abi.BVH4, abi.BVH2])
def test_soup_closest_matches_reference(ctx, kind):
    import salty_amd
    desc = scenes.triangle_soup(50_000, seed=3)
    desc.meshes[0].bvh_kind = kind
    rays = scenes.soup_rays(100_000, seed=4)
    sc = salty_amd.Scene(ctx, desc)
    gpu = sc.trace(rays, abi.TRACE_CLOSEST)
    ref = O.OracleMesh(desc.meshes[0].vertices, kind).trace(rays)
    _check_closest(gpu, ref)
    assert (ref["prim"] >= 0).mean() > 0.2


def test_soup_any_hit_agrees_with_closest(ctx):
    import salty_amd
    desc = scenes.triangle_soup(50_000, seed=5)
    rays = scenes.soup_rays(100_000, seed=6)
    sc = salty_amd.Scene(ctx, desc)
    closest = sc.trace(rays, abi.TRACE_CLOSEST)
    anyh = sc.trace(rays, abi.TRACE_ANY)
    assert np.array_equal(anyh["prim"] >= 0, closest["prim"] >= 0)
    hit = anyh["prim"] >= 0
    assert np.all(anyh["t"][hit] >= closest["t"][hit])
    assert np.all(anyh["t"][hit] >= rays["tmin"][hit])


def test_cornell_camera_rays(ctx):
    import salty_amd
    desc = scenes.cornell_box()
    rays = scenes.camera_rays(desc, 128, 128)
    sc = salty_amd.Scene(ctx, desc)
    gpu, gs = sc.trace(rays, abi.TRACE_CLOSEST, with_shapes=True)
    ref, rs = O.OracleScene(desc).trace(rays)
    _check_closest(gpu, ref)
    assert np.array_equal(gs, rs)
