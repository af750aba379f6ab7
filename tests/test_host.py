"""Host-side checks that need no GPU: the C-ABI library loads and exports every
entry point include/s3dg.h declares, and the product's host SAH builder
reproduces the oracle's reference tree exactly (same depth-first permutation)."""
import ctypes
import os
import re

import numpy as np
import pytest

import salty_amd
from oracle import oracle as O
from salty_amd import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "s3dg.h")).read()
    return sorted(set(re.findall(r"\b(s3dg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(salty_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    assert set(declared_symbols()) <= set(salty_amd._SIGNATURES)


@pytest.mark.parametrize("kind", [abi.BVH8, abi.BVH4, abi.BVH2])
@pytest.mark.parametrize("n,seed", [(9, 0), (100, 1), (5000, 2), (30000, 3)])
def test_host_builder_matches_oracle_tree(kind, n, seed):
    v = scenes.triangle_soup(n, seed=seed).meshes[0].vertices
    order, nodes = salty_amd.build_order(v, kind)
    om = O.OracleMesh(v, kind)
    assert np.array_equal(order, om.order())
    assert nodes == len(om.tree()["kind"])


def test_host_builder_matches_oracle_on_structured_meshes():
    from salty_amd.scenes import _grid, _mesh_vertices, icosphere
    gv, gf, gn = _grid(-10, 10, -10, 10, 60, 30)
    v1 = _mesh_vertices(gv, gf, gn)
    iv, iff = icosphere(4)
    v2 = _mesh_vertices(iv, iff)
    for v in (v1, v2, scenes.cornell_box().meshes[0].vertices):
        order, nodes = salty_amd.build_order(v, abi.BVH8)
        om = O.OracleMesh(v, abi.BVH8)
        assert np.array_equal(order, om.order())
        assert nodes == len(om.tree()["kind"])


def test_context_without_device_fails_loudly():
    import salty_amd as s
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a device is present")
    with pytest.raises(s.S3dgError):
        s.Context(0)
