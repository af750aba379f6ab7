"""Traversal edge cases, GPU against the oracle bit for bit: axis-aligned
(flat) triangles on an integer grid hit exactly on vertices and shared edges
(where Triangle::IsHit's strict edge tests reject both neighbours), rays along
the coordinate axes (infinite reciprocals in the slab tests), grazing rays in
the plane and leaving it at slopes down to 1e-7, ray ends exactly at a hit
distance, and scenes far from the origin or tiny."""
import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes
from salty_amd.abi import Camera, Mesh, SceneDesc, Shape

pytestmark = pytest.mark.gpu


def _grid_desc(n=24, offset=(0.0, 0.0, 0.0), scale=1.0):
    """An n x n grid of unit quads in the plane z = 0 and the same grid turned
    into the plane x = 0, plus a third copy of the first quad row (duplicates)."""
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = (i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)
            for q in ((a, b, c), (a, c, d)):
                tris.append([(p[0], p[1], 0.0) for p in q])
                tris.append([(0.0, p[0], p[1]) for p in q])
    for i in range(n):
        a, b, c = (i, 0), (i + 1, 0), (i + 1, 1)
        tris.append([(p[0], p[1], 0.0) for p in (a, b, c)])
    p = (np.array(tris, np.float64) * scale + np.array(offset)).astype(np.float32)
    d = SceneDesc()
    d.meshes = [Mesh(scenes._tri_vertices(p), material=0)]
    d.shapes = [Shape(abi.SHAPE_MESH, mesh=0)]
    d.materials = [dict(type=abi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))]
    d.camera = Camera(position=(0, 0, 5), target=(0, 0, 0))
    return d, p


def _edge_rays(n, offset, scale):
    rng = np.random.default_rng(3)
    o, dirs, tmin, tmax = [], [], [], []
    off = np.array(offset, np.float64)
    # straight down (-z) onto vertices, edge midpoints, cell centres and diagonals
    for fx in (0.0, 0.5, 0.25, 0.75):
        for fy in (0.0, 0.5, 0.25):
            for i in range(0, n, 3):
                for j in range(0, n, 3):
                    o.append((i + fx, j + fy, 3.0))
                    dirs.append((0.0, 0.0, -1.0))
                    tmin.append(1e-3)
                    tmax.append(1e12)
    # along -x onto the x = 0 grid, exactly on its edges
    for i in range(n + 1):
        for j in range(0, n + 1, 2):
            o.append((2.0, float(i), float(j)))
            dirs.append((-1.0, 0.0, 0.0))
            tmin.append(1e-3)
            tmax.append(1e12)
    # grazing rays in the planes z = 0 and x = 0, and random directions
    for k in range(400):
        a = rng.uniform(0, 2 * np.pi)
        o.append((rng.uniform(0, n), rng.uniform(0, n), 0.0))
        dirs.append((np.cos(a), np.sin(a), 0.0))
        tmin.append(1e-3)
        tmax.append(1e12)
    # rays leaving the plane z = 0 almost parallel to it, from on, just above
    # and just below it (the case the slab test's rounding slack must keep
    # exact with flat leaf boxes)
    for k in range(600):
        a = rng.uniform(0, 2 * np.pi)
        s = rng.choice([0.0, 1e-7, -1e-7, 1e-6, -1e-6, 1e-5, -1e-5, 1e-3, -1e-3])
        d = np.array([np.cos(a), np.sin(a), s])
        o.append((rng.uniform(0, n), rng.uniform(0, n), rng.choice([0.0, 1e-7, -1e-7, 1e-6, -1e-6])))
        dirs.append(tuple(d / np.linalg.norm(d)))
        tmin.append(1e-3)
        tmax.append(1e12)
    for k in range(3000):
        d = rng.normal(size=3)
        o.append(tuple(rng.uniform(-1, n + 1, 3)))
        dirs.append(tuple(d / np.linalg.norm(d)))
        tmin.append(1e-3)
        tmax.append(1e12)
    # ray ends exactly at the hit distance (3.0 above the plane): t == tmax is
    # kept, t == tmin is kept, tmax just below t misses
    for tm, tx in ((1e-3, 3.0), (3.0, 1e12), (1e-3, float(np.nextafter(np.float32(3.0), np.float32(0.0))))):
        for i in range(1, n, 5):
            o.append((i + 0.3, 0.6, 3.0))
            dirs.append((0.0, 0.0, -1.0))
            tmin.append(tm)
            tmax.append(tx)
    r = np.zeros(len(o), abi.RAY_DTYPE)
    r["org"] = (np.array(o) * scale + off).astype(np.float32)
    r["dir"] = np.array(dirs, np.float32)
    r["tmin"] = np.array(tmin, np.float32) * np.float32(scale)
    r["tmax"] = np.minimum(np.array(tmax, np.float64) * scale, 3.4e38).astype(np.float32)
    return r


@pytest.mark.parametrize("offset,scale", [((0.0, 0.0, 0.0), 1.0), ((1e5, -2e5, 3e4), 1.0),
                                          ((0.0, 0.0, 0.0), 1e-3), ((5e3, 5e3, 5e3), 40.0)])
@pytest.mark.parametrize("leaf_size", [0, 2])
def test_edge_cases_match_oracle(ctx, offset, scale, leaf_size):
    import salty_amd
    n = 24
    desc, _ = _grid_desc(n, offset, scale)
    rays = _edge_rays(n, offset, scale)
    start = ctx.get_tuning()
    t = ctx.get_tuning()
    t.leaf_size = leaf_size
    ctx.set_tuning(t)
    try:
        sc = salty_amd.Scene(ctx, desc)
    finally:
        ctx.set_tuning(start)
    gpu = sc.trace(rays, abi.TRACE_CLOSEST)
    om = O.OracleMesh(desc.meshes[0].vertices, abi.BVH8)
    ref = om.trace(rays)
    bad = np.nonzero(gpu["prim"] != ref["prim"])[0]
    assert len(bad) == 0, f"{len(bad)} mismatches, first {bad[:5]}: gpu {gpu['prim'][bad[:5]]} ref {ref['prim'][bad[:5]]}"
    hit = ref["prim"] >= 0
    assert 0.2 < hit.mean() < 0.95
    for f in ("t", "u", "v"):
        assert np.array_equal(gpu[f][hit].view(np.uint32), ref[f][hit].view(np.uint32))
    # the brute-force argmin agrees with the reference traversal here
    assert np.array_equal(om.trace(rays, brute=True)["prim"], ref["prim"])
    anyh = sc.trace(rays, abi.TRACE_ANY)
    assert np.array_equal(anyh["prim"] >= 0, hit)
