"""Oracle geometry: Triangle::IsHit arithmetic, the reference SAH trees and the
reference traversal, pinned by closed forms, FP64 recomputation, tree
invariants and brute force."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes


def ray(o, d, tmin=1e-3, tmax=1e12):
    r = np.zeros(1, dtype=abi.RAY_DTYPE)
    r["org"] = o
    r["dir"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def test_triangle_centroid_hit_closed_form():
    hit, h = O.triangle_hit([0, 0, 0], [3, 0, 0], [0, 3, 0], ray([1, 1, 2], [0, 0, -1]))
    assert hit
    assert h["t"] == np.float32(2.0)
    assert h["u"] == np.float32(1.0 / 3.0) and h["v"] == np.float32(1.0 / 3.0)


def test_triangle_edges_and_vertices_are_rejected():
    # beta or gamma exactly 0, or beta+gamma exactly 1, is not a hit (strict tests)
    for o in ([1.5, 0.0, 1.0], [0.0, 1.5, 1.0], [1.5, 1.5, 1.0], [0.0, 0.0, 1.0]):
        hit, _ = O.triangle_hit([0, 0, 0], [3, 0, 0], [0, 3, 0], ray(o, [0, 0, -1]))
        assert not hit, o


def test_triangle_parallel_and_range():
    tri = ([0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert not O.triangle_hit(*tri, ray([0.2, 0.2, 1], [1, 0, 0]))[0]            # |div| <= FLT_EPSILON
    assert not O.triangle_hit(*tri, ray([0.2, 0.2, 5e-4], [0, 0, -1]))[0]        # t < F_HIT_MIN
    assert O.triangle_hit(*tri, ray([0.2, 0.2, 5e-4], [0, 0, -1], tmin=1e-4))[0]
    assert not O.triangle_hit(*tri, ray([0.2, 0.2, 5.0], [0, 0, -1], tmax=4.0))[0]
    # |div| threshold is absolute: a tiny triangle is missed even head-on
    assert not O.triangle_hit([0, 0, 0], [1e-4, 0, 0], [0, 1e-4, 0], ray([2e-5, 2e-5, 1], [0, 0, -1]))[0]


def test_triangle_t_and_barycentrics_match_fp64():
    rng = np.random.default_rng(0)
    n_hit = 0
    for _ in range(3000):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        p = (v[0] * 0.3 + v[1] * 0.3 + v[2] * 0.4 + rng.normal(scale=0.3, size=3)).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        o = (p - 3.0 * d).astype(np.float32)
        hit, h = O.triangle_hit(v[0], v[1], v[2], ray(o, d))
        # FP64 Moller-Trumbore on the same float inputs
        v0, v1, v2 = (x.astype(np.float64) for x in v)
        e0, e1 = v1 - v0, v2 - v0
        s1 = np.cross(d.astype(np.float64), e1)
        div = s1 @ e0
        q = o.astype(np.float64) - v0
        b = (q @ s1) / div
        s2 = np.cross(q, e0)
        g = (d.astype(np.float64) @ s2) / div
        t = (e1 @ s2) / div
        inside = b > 1e-6 and g > 1e-6 and b + g < 1 - 1e-6 and abs(div) > 1e-6
        if inside:
            assert hit
        if hit:
            n_hit += 1
            assert abs(h["t"] - t) <= 1e-5 * abs(t) + 1e-7
            assert abs(h["u"] - b) <= 1e-5 * abs(b) + 1e-6
            assert abs(h["v"] - g) <= 1e-5 * abs(g) + 1e-6
    assert n_hit > 500


def _tree_invariants(mesh, n, kind):
    t = mesh.tree()
    order = mesh.order()
    assert sorted(order.tolist()) == list(range(n))           # a permutation
    kinds = t["kind"]
    seen = np.zeros(n, dtype=np.int64)
    verts = mesh.v["position"].reshape(-1, 3, 3)
    lo = verts.min(axis=1)
    hi = verts.max(axis=1)
    for i in range(len(kinds)):
        b, c = t["begin"][i], t["count"][i]
        box = t["boxes"][i]
        prims = order[b:b + c]
        # node box is exactly the union of its prims' boxes
        assert np.array_equal(box[:3], lo[prims].min(axis=0)) and np.array_equal(box[3:], hi[prims].max(axis=0))
        if kinds[i] == 0:
            seen[b:b + c] += 1
            continue
        assert kinds[i] in ((8, 4) if kind == 8 else (kind,))
        kids = t["children"][t["child_begin"][i]:t["child_begin"][i] + t["child_count"][i]]
        assert len(kids) == kinds[i]
        # children tile the parent's range in order
        starts = t["begin"][kids]
        counts = t["count"][kids]
        assert starts[0] == b and np.all(starts[1:] == starts[:-1] + counts[:-1]) and counts.sum() == c
        assert np.all(counts > 0)
    assert np.all(seen == 1)                                     # every prim in exactly one leaf
    # leaf sizes: BVH8 leaves from CreateNode hold <= 8 unless a split failed
    return t


@pytest.mark.parametrize("kind", [abi.BVH8, abi.BVH4, abi.BVH2])
def test_tree_invariants(kind):
    d = scenes.triangle_soup(3000, seed=1)
    m = O.OracleMesh(d.meshes[0].vertices, kind)
    t = _tree_invariants(m, 3000, kind)
    if kind == abi.BVH8:
        assert 8 in t["kind"]


def test_bvh8_small_mesh_is_one_leaf_and_fallback_to_bvh4():
    d = scenes.triangle_soup(8, seed=2)
    m = O.OracleMesh(d.meshes[0].vertices, abi.BVH8)
    t = m.tree()
    assert len(t["kind"]) == 1 and t["kind"][0] == 0 and t["count"][0] == 8
    # 9..64 triangles cannot give 8 groups of more than 8 (SplitSAH refuses <= 8):
    # BVH8::Create falls back to BVH4::Create (s3d_bvh8.cpp:94-110)
    d = scenes.triangle_soup(40, seed=2)
    t = O.OracleMesh(d.meshes[0].vertices, abi.BVH8).tree()
    assert 8 not in t["kind"] and 4 in t["kind"]


@pytest.mark.parametrize("kind", [abi.BVH8, abi.BVH4, abi.BVH2])
def test_reference_traversal_equals_brute_force(kind):
    d = scenes.triangle_soup(4000, seed=7)
    m = O.OracleMesh(d.meshes[0].vertices, kind)
    rays = scenes.soup_rays(20000, seed=8)
    a = m.trace(rays)
    b = m.trace(rays, brute=True)
    assert np.array_equal(a["prim"], b["prim"])
    assert np.array_equal(a["t"], b["t"])
    assert (a["prim"] >= 0).mean() > 0.02


def test_structured_mesh_traversal_equals_brute_force():
    # axis-aligned grid triangles give flat boxes: the slab-test corner cases
    from salty_amd.scenes import _grid, _mesh_vertices
    gv, gf, gn = _grid(-1, 1, -1, 1, 40, 40)
    m = O.OracleMesh(_mesh_vertices(gv, gf, gn), abi.BVH8)
    rng = np.random.default_rng(3)
    r = np.zeros(20000, dtype=abi.RAY_DTYPE)
    r["org"] = np.c_[rng.uniform(-1.2, 1.2, 20000), rng.uniform(0.1, 2.0, 20000), rng.uniform(-1.2, 1.2, 20000)]
    d = rng.normal(size=(20000, 3))
    d[:, 1] = -np.abs(d[:, 1])
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmin"] = 1e-3
    r["tmax"] = 1e12
    a = m.trace(r)
    b = m.trace(r, brute=True)
    assert np.array_equal(a["prim"], b["prim"])
    assert (a["prim"] >= 0).mean() > 0.2


def test_invert_is_exact_inverse_for_rigid_and_reproduces_precedence_quirk():
    # rigid transform (det = 1): the reference formula is a true inverse
    c, s = np.cos(0.3), np.sin(0.3)
    m = np.array([[c, 0, -s, 0], [0, 1, 0, 0], [s, 0, c, 0], [1.5, -2.0, 0.25, 1]], dtype=np.float32)
    assert np.allclose(O.invert(m) @ m, np.eye(4), atol=1e-6)
    # uniform scale 2 (det = 8): only the last product of each cofactor is
    # divided by det (operator precedence, s3d_math.h:1805-1839).  For
    # diag(2,2,2,1) every diagonal cofactor's surviving product is the first
    # one: entry 11 = 2*2*1 = 4 (a true inverse would give 0.5), entry 44 = 8.
    m = np.diag([2.0, 2.0, 2.0, 1.0]).astype(np.float32)
    inv = O.invert(m)
    assert inv[0, 0] == 4.0 and inv[1, 1] == 4.0 and inv[2, 2] == 4.0 and inv[3, 3] == 8.0
