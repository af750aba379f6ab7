"""Instance::UpdateMatrix: new world matrices for a scene's instances rebuild
its top-level tree over the unchanged mesh trees.  After every update the
scene traces and renders exactly like a scene created with the new matrices,
and like the oracle's restatement of it."""
import copy
import math

import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _rot_y(deg):
    a = math.radians(deg)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = math.cos(a), -math.sin(a), math.sin(a), math.cos(a)
    return m


def _frame(desc, inst_ids, t):
    """World matrices of frame t: each instance spins and drifts (row vectors:
    rotate, then translate)."""
    worlds = []
    for k, i in enumerate(inst_ids):
        base = desc.shapes[i].world
        w = _rot_y(17.0 * t + 40.0 * k) @ base
        w = w @ scenes.translate(0.3 * math.sin(t + k), 0.1 * t, -0.25 * t)
        worlds.append(w.astype(np.float32))
    return worlds


def test_updates_equal_fresh_scenes(ctx):
    import salty_amd
    desc = scenes.mesh_scene(big_level=3, small_level=2, ground=(30, 16), instances=10)
    ids = [i for i, s in enumerate(desc.shapes) if s.type == abi.SHAPE_INSTANCE]
    assert len(ids) == 10
    live = salty_amd.Scene(ctx, desc)
    rays = scenes.soup_rays(30_000, seed=5)
    rays["org"] = rays["org"] * 8.0 - np.array([4.0, 0.5, 4.0], np.float32)
    live_worlds = {i: desc.shapes[i].world for i in ids}
    for t in range(1, 4):
        worlds = _frame(desc, ids, t)
        moved = copy.deepcopy(desc)
        for i, w in zip(ids, worlds):
            moved.shapes[i].world = w
        # only some instances change on odd frames
        sel = ids if t % 2 == 0 else ids[::3]
        live.update_instances(sel, [worlds[ids.index(i)] for i in sel])
        if t % 2 == 1:
            for i in ids:
                if i not in sel:
                    moved.shapes[i].world = live_worlds[i]
        live_worlds = {i: moved.shapes[i].world for i in ids}
        fresh = salty_amd.Scene(ctx, moved)
        a, sa = live.trace(rays, abi.TRACE_CLOSEST, with_shapes=True)
        b, sb = fresh.trace(rays, abi.TRACE_CLOSEST, with_shapes=True)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(sa, sb), f"frame {t}"
        ia, sta = live.render(64, 40, spp=2, max_depth=5, seed=t)
        ib, stb = fresh.render(64, 40, spp=2, max_depth=5, seed=t)
        assert np.array_equal(ia, ib), f"frame {t}"
        assert (sta.extension_rays, sta.shadow_rays) == (stb.extension_rays, stb.shadow_rays)
        fresh.close()
    ref, _ = O.OracleScene(moved).render(64, 40, spp=2, max_depth=5, seed=3)
    assert float(np.mean((ia - ref) ** 2 / (ref ** 2 + 1e-2))) < 1e-4


def test_update_rejects_non_instances(ctx):
    import salty_amd
    desc = scenes.mesh_scene(big_level=2, small_level=1, ground=(8, 4), instances=2)
    sc = salty_amd.Scene(ctx, desc)
    mesh_shape = next(i for i, s in enumerate(desc.shapes) if s.type == abi.SHAPE_MESH)
    with pytest.raises(salty_amd.S3dgError):
        sc.update_instances([mesh_shape], [np.eye(4, dtype=np.float32)])
    inst = next(i for i, s in enumerate(desc.shapes) if s.type == abi.SHAPE_INSTANCE)
    with pytest.raises(salty_amd.S3dgError):
        sc.update_instances([inst], [np.zeros((4, 4), np.float32)])     # singular
    # the scene is unchanged by the rejected calls
    img, _ = sc.render(32, 24, spp=1, max_depth=3, seed=1)
    ref, _ = salty_amd.Scene(ctx, desc).render(32, 24, spp=1, max_depth=3, seed=1)
    assert np.array_equal(img, ref)
