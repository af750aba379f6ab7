"""The tree the kernels traverse is their own (binned SAH on three axes, small
leaves); which hit wins a tie is still the reference tree's depth-first
order.  Every leaf size, and the reference-shaped tree itself, must give the
oracle's hits bit for bit, duplicated triangles (exact ties) included, and the
same renders."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu

LEAF_SIZES = (0, 1, 2, 4, 8)


def _with_leaf_size(ctx, leaf_size, build):
    start = ctx.get_tuning()
    t = ctx.get_tuning()
    t.leaf_size = leaf_size
    ctx.set_tuning(t)
    try:
        return build()
    finally:
        ctx.set_tuning(start)


def _duplicated_soup(n, copies, seed, kind):
    desc = scenes.triangle_soup(n, seed=seed)
    v = desc.meshes[0].vertices.reshape(n, 3)
    # copy c of triangle i sits at index c * n + i: every hit is an exact tie
    # among the copies, won by the first in the reference tree's order
    desc.meshes[0].vertices = np.concatenate([v] * copies).reshape(-1)
    desc.meshes[0].bvh_kind = kind
    return desc


@pytest.mark.parametrize("kind", [abi.BVH8, abi.BVH2])
def test_leaf_sizes_match_reference_with_ties(ctx, kind):
    import salty_amd
    desc = _duplicated_soup(30_000, 2, seed=21, kind=kind)
    rays = scenes.soup_rays(40_000, seed=22)
    ref = O.OracleMesh(desc.meshes[0].vertices, kind).trace(rays)
    assert (ref["prim"] >= 0).mean() > 0.1
    # the tie goes to either copy, as the reference tree orders them
    assert (ref["prim"] >= 30_000).sum() > 100 and ((ref["prim"] >= 0) & (ref["prim"] < 30_000)).sum() > 100
    for ls in LEAF_SIZES:
        sc = _with_leaf_size(ctx, ls, lambda: salty_amd.Scene(ctx, desc))
        gpu = sc.trace(rays, abi.TRACE_CLOSEST)
        bad = int((gpu["prim"] != ref["prim"]).sum())
        assert bad == 0, f"leaf_size {ls}: {bad} prim mismatches"
        hit = ref["prim"] >= 0
        assert np.array_equal(gpu["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32))
        anyh = sc.trace(rays, abi.TRACE_ANY)
        assert np.array_equal(anyh["prim"] >= 0, hit)


def test_leaf_sizes_render_identically(ctx):
    import salty_amd
    desc = scenes.mesh_scene(big_level=4, small_level=2, ground=(40, 20), instances=8)
    ref = None
    for ls in LEAF_SIZES:
        sc = _with_leaf_size(ctx, ls, lambda: salty_amd.Scene(ctx, desc))
        img, st = sc.render(80, 60, spp=3, max_depth=6, seed=4)
        if ref is None:
            ref = (img, st.extension_rays, st.shadow_rays)
        assert np.array_equal(img, ref[0]), f"leaf_size {ls}"
        assert (st.extension_rays, st.shadow_rays) == ref[1:]
