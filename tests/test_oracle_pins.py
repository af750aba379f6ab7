"""Pins of oracle functions against hand-worked cases and closed forms:
the reference SAH tree shape (s3d_bvh8.cpp / s3d_bvh4.cpp / s3d_bvh2.cpp),
the Uncharted2 and ACES operators with the reference's constants
(s3d_tonemapper.cpp), Sphere::IsHit (s3d_sphere.cpp:66-92) and the instance
ray transform (s3d_instance.cpp:65-72).

Every expected value here is derived in the comments from the reference's
source as read, not computed by the oracle's code."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi
from salty_amd.abi import Camera, Light, Mesh, SceneDesc, Shape

# ----------------------------------------------------------------- SAH tree --
# Test triangles: triangle k has the box x in [c_k - A, c_k + A], y in
# [-A, A], z = 0 (vertices (c-A, -A, 0), (c+A, -A, 0), (c, A, 0)), so its box
# centre, which is what the reference bins (Triangle::GetCenter returns the
# box centre, s3d_triangle.cpp:132), is (c_k, 0, 0).  All centres lie on the
# x axis, so the centroid box has zero y and z extent and x is the split axis
# (GetLongestAxis, s3d_bvh8.cpp:61-70).  With z extent 0 a box's surface area
# is 2*X*Y and Y = 2A for every group, so a split's cost is
# 1 + (n0*X0 + n1*X1) / Xb with X the x extent of a group of boxes.  Every
# coordinate is a dyadic rational, so centres, offsets and 16*offset are
# exact in float.


def _tris(centres, a):
    n = len(centres)
    v = np.zeros(3 * n, dtype=abi.VERTEX_DTYPE)
    p = v["position"].reshape(n, 3, 3)
    c = np.asarray(centres, dtype=np.float64)
    p[:, 0] = np.stack([c - a, -a + 0 * c, 0 * c], -1)
    p[:, 1] = np.stack([c + a, -a + 0 * c, 0 * c], -1)
    p[:, 2] = np.stack([c, a + 0 * c, 0 * c], -1)
    v["normal"] = (0.0, 0.0, 1.0)
    return v


def _shape(tree, node):
    """Nested (kind, ...) description of a subtree: ("leaf", count) or
    (arity, [children])."""
    k = int(tree["kind"][node])
    if k == 0:
        return ("leaf", int(tree["count"][node]))
    b, c = tree["child_begin"][node], tree["child_count"][node]
    return (k, [_shape(tree, int(ch)) for ch in tree["children"][b:b + c]])


# Ten triangles; s = 16 c is the value the root split bins (its centroid box
# is [0, 1], so offset = c).  Input index: s
#   0: 14.5  1: 0.75  2: 1.5  3: 0  4: 16  5: 1.0  6: 14  7: 0.25  8: 15  9: 0.5
TEN_S = [14.5, 0.75, 1.5, 0.0, 16.0, 1.0, 14.0, 0.25, 15.0, 0.5]
TEN_A = 1.0 / 1024.0
# Derivation (x extents in units of 1/1024, 2A = 2 units, s = 1 is 64 units):
#
# Root, 16 buckets: b0 = {0, .25, .5, .75}, b1 = {1, 1.5}, b14 = {14, 14.5},
# b15 = {15, 16 (16 -> bucket 15)}; Xb = 1026.
#   split 0:      4 * 50  + 6 * 930 -> 1 + 5780/1026 = 6.63
#   splits 1..13: 6 * 98  + 4 * 130 -> 1 + 1108/1026 = 2.08  (equal sets: equal floats)
#   split 14:     8 * 930 + 2 * 66  -> 1 + 7572/1026 = 8.38
# The first strict minimum is split 1 (s3d_bvh8.cpp:305-315), 2.08 < 10.  The
# partition predicate compares the UNTRUNCATED 16*offset with the bucket
# (s3d_bvh8.cpp:321-334): s <= 1, so s = 1.5, binned into bucket 1, goes to
# the right side (an int-truncated predicate would keep it left).
# std::partition (two-ended swaps) on [0..9], pred true for {1, 3, 5, 7, 9}:
#   0 F | 9 T -> swap -> [9,1,2,...,8,0];  1 T, 2 F | 8 F, 7 T -> swap ->
#   [9,1,7,3,4,5,6,2,8,0];  3 T, 4 F | 6 F, 5 T -> swap -> [9,1,7,3,5,4,6,2,8,0];
#   first meets last: mid = 5.
#
# Left [9,1,7,3,5] = s {.5, .75, .25, 0, 1}: centroid box [0, 1/16], so
# 16*offset = 16 s: 0, 4, 8, 12, 16 -> buckets 0, 4, 8, 12, 15; Xb = 66.
#   splits 0..3:   1 * 2  + 4 * 50 -> 1 + 202/66 = 4.06
#   splits 4..7:   2 * 18 + 3 * 34 -> 1 + 138/66 = 3.09
#   splits 8..11:  3 * 34 + 2 * 18 -> 1 + 138/66 = 3.09  (the same float sum)
#   splits 12..14: 4 * 50 + 1 * 2  -> 4.06
# First strict minimum: split 4 (a "<=" minimum would take 11).  Predicate
# 16 s <= 4: {0, .25} = triangles {3, 7}.  Partition [9,1,7,3,5]:
#   9 F | 5 F, 3 T -> swap -> [3,1,7,9,5];  1 F | 7 T -> swap -> [3,7,1,9,5];
#   mid = 2.
#
# Right [4,6,2,8,0] = s {16, 14, 1.5, 15, 14.5}: centroid box [3/32, 1]
# (extent 29/32), 16*offset = 16 (32c - 3)/29: 16 -> 16 (bucket 15),
# 14 -> 13.79 (13), 1.5 -> 0 (0), 15 -> 14.90 (14), 14.5 -> 14.34 (14);
# Xb = 930.
#   splits 0..12: 1 * 2 + 4 * 130   -> 1 + 522/930  = 1.56
#   split 13:     2 * 834 + 3 * 98  -> 1 + 1962/930 = 3.11
#   split 14:     4 * 866 + 1 * 2   -> 4.73
# Split 0; predicate 16*offset <= 0: only s = 1.5 (offset exactly 0) =
# triangle 2.  Partition [4,6,2,8,0]: 4 F | 0 F, 8 F, 2 T -> swap ->
# [2,6,4,8,0]; 6 F meets last: mid = 1.
#
# BVH4::Create(10) (s3d_bvh4.cpp:370-423): root split, both halves (5 > 4)
# split, four children of 2, 3, 1, 4 triangles -> BVH4::Create(<= 4) = Leaf.
TEN_ORDER = [3, 7, 1, 9, 5, 2, 6, 4, 8, 0]


def test_sah_hand_worked_bvh4():
    m = O.OracleMesh(_tris(np.array(TEN_S) / 16.0, TEN_A), abi.BVH4)
    assert m.order().tolist() == TEN_ORDER
    t = m.tree()
    assert _shape(t, len(t["kind"]) - 1) == (4, [("leaf", 2), ("leaf", 3), ("leaf", 1), ("leaf", 4)])


def test_sah_hand_worked_bvh8_falls_back_to_bvh4():
    """BVH8::Create(10): the root split (count 10 > 8, the same SAH as above)
    partitions in place, then Split of the 5-triangle left half returns false
    (count <= 8), so CreateNode(10) -- 10 > 8 -- builds BVH4::Create(10) over
    the partitioned array (s3d_bvh8.cpp:94-110, 401-405).  BVH4 repeats the
    root split on an already partitioned array (no swaps) and ends as above."""
    m = O.OracleMesh(_tris(np.array(TEN_S) / 16.0, TEN_A), abi.BVH8)
    assert m.order().tolist() == TEN_ORDER
    t = m.tree()
    assert _shape(t, len(t["kind"]) - 1) == (4, [("leaf", 2), ("leaf", 3), ("leaf", 1), ("leaf", 4)])


def test_sah_hand_worked_bvh2():
    """BVH2::Create, 12 buckets, no minimum count inside the split
    (s3d_bvh2.cpp:174-274).  Root: 12 c = 0.75 s: 0, .1875, .375, .5625, .75
    (bucket 0), 1.125 (bucket 1), 10.5, 10.875 (10), 11.25, 12 (11).  Costs as
    for 16 buckets: split 0 6.63, splits 1..9 2.08, split 10 8.38 -> split 1;
    predicate 12 c <= 1 sends 1.125 (triangle 2) right: the same sides, so the
    same partition.  Left: 12 * 16 c = 12 s -> 0, 3, 6, 9, 12: splits 3..5
    and 6..8 tie at 3.09 -> split 3 -> {3, 7} left.  Right: 12*offset = 0
    (triangle 2), 10.34, 10.76, 11.17, 12 -> split 0 -> {2} left.  Children
    of <= 4 triangles are leaves."""
    m = O.OracleMesh(_tris(np.array(TEN_S) / 16.0, TEN_A), abi.BVH2)
    assert m.order().tolist() == TEN_ORDER
    t = m.tree()
    assert _shape(t, len(t["kind"]) - 1) == (2, [(2, [("leaf", 2), ("leaf", 3)]), (2, [("leaf", 1), ("leaf", 4)])])


def test_sah_hand_worked_bvh8_eight_children():
    """36 triangles in 8 groups of identical triangles at x = u * {0, 1, 15, 16,
    48, 49, 63, 64}, u = 1/64, A = 1/4096, group sizes 4, 5, 5, 4, 3, 6, 6, 3.
    Input order: the groups from last to first, each group in index order.

    Root (centroid box [0, 1], 16 offset = 16 x): groups at 0, .25, 3.75, 4
    (buckets 0, 0, 3, 4) and 12, 12.25, 15.75, 16 (12, 12, 15, 15).  In
    units of 1/4096 (Xb = 4098): splits 4..11 give the halves, 1 + (18 * 1026
    + 18 * 1026) / 4098 = 10.01, the least (splits 0..2: 21.8, split 3: 20.8,
    splits 12..14: 21.8); the first is 4, and 16 x <= 4 holds exactly for the
    left groups (3.75 and 4 are <= 4).  Each half's centroid box spans 16 u,
    so 16 offset = 0, 1, 15, 16 (Xb = 1026): splits 1..14 cost 1 + (9 * 66 +
    9 * 66) / 1026 = 2.16, split 0 14.1; first: 1 (value 1 <= 1 stays left).
    Each quarter is two groups, 16 offset = 0 and 16: splits 0..14 all cost
    1 + 18 * 2 / 66 = 1.55 < 9, split 0.  All seven splits succeed, so
    the node is a BVH8 whose eight children are the groups (<= 8: leaves),
    in the order idx0..idx7 of s3d_bvh8.cpp:444-478.

    Order: the root partition sees the right half's 18 triangles first and
    the left half's last; the two-ended partition swaps position i with
    35 - i for i = 0..17, reversing the array.  The reversed array is already
    partitioned for every lower split (no further swaps), so the depth-first
    order is the input reversed."""
    u, a = 1.0 / 64.0, 1.0 / 4096.0
    pos = [0, 1, 15, 16, 48, 49, 63, 64]
    sizes = [4, 5, 5, 4, 3, 6, 6, 3]
    centres = []
    for g in reversed(range(8)):
        centres += [pos[g] * u] * sizes[g]
    m = O.OracleMesh(_tris(centres, a), abi.BVH8)
    assert m.order().tolist() == list(range(35, -1, -1))
    t = m.tree()
    # children in group order 0..7 (input was 7..0)
    assert _shape(t, len(t["kind"]) - 1) == (8, [("leaf", s) for s in sizes])
    # a stable partition would have kept each group in index order
    assert m.order()[:4].tolist() == [35, 34, 33, 32]


# ------------------------------------------------------------------ tonemap --
# One-pixel images: Y = 0.299 r + 0.587 g + 0.114 b (RGBToY,
# s3d_tonemapper.cpp:20,31), aveLw = exp(log(1e-5 + Y)) = Y + 1e-5
# (ComputeLogarithmicAverage with one pixel), so each channel's operator
# input is exact up to the float rounding of logf/expf.
#
# ACES (key 0.27, ACESFilm s3d_tonemapper.cpp:117-130, 320-353), x = 0.27 c /
# (Y + 1e-5), f/g = 0.665406 / 12:
#   (1, 0, 0): x = 0.27 / 0.29901 = 0.902980 -> x(2.51 x f/g + 0.03) /
#              (x f/g (2.43 x f + 0.59) + 0.14) = 0.579331
#   (0, 1, 0): x = 0.27 / 0.58701 = 0.459958 -> 0.248505
#   (0, 0, 1): x = 2.368213 -> 1.1820 before Saturate -> 1
#   (0.2, 0.1, 1): Y = 0.2325, x = 0.232248, 0.116124, 1.161240 ->
#              0.0949572, 0.0369672, 0.744473
# Uncharted2 (key 0.18, s3d_tonemapper.cpp:100-112, 280-315):
# out = U(2 x) / U(11.2), x = 0.18 c / (Y + 1e-5), U(x) = (x (0.15 x + 0.1 *
# 0.5) + 0.2 * 0.01) / (x (0.15 x + 0.5) + 0.2 * 0.3) - 0.01 / 0.3 (E = 0.01
# in the reference; Hable's published curve has E = 0.02, which gives
# 0.349704 for the first case); U(11.2) = 0.758381:
#   (1, 0, 0): 2x = 1.203973 -> 0.375325
#   (0, 1, 0): 2x = 0.613277 -> 0.233696
#   (0, 0, 1): 2x = 3.157618 -> 0.652479
#   (0.2, 0.1, 1): 2x = 0.309664, 0.154832, 1.548321 -> 0.139366, 0.0807657, 0.440969
# A zero channel maps to U(0) / U(11.2) = (0.002 / 0.06 - 0.01 / 0.3) / ... = 0.
TONEMAP_CASES = [
    (abi.TONEMAP_ACES, (1, 0, 0), (0.5793313, 0.0, 0.0)),
    (abi.TONEMAP_ACES, (0, 1, 0), (0.0, 0.2485053, 0.0)),
    (abi.TONEMAP_ACES, (0, 0, 1), (0.0, 0.0, 1.0)),
    (abi.TONEMAP_ACES, (0.2, 0.1, 1), (0.09495717, 0.03696720, 0.7444729)),
    (abi.TONEMAP_UNCHARTED2, (1, 0, 0), (0.3753251, 0.0, 0.0)),
    (abi.TONEMAP_UNCHARTED2, (0, 1, 0), (0.0, 0.2336963, 0.0)),
    (abi.TONEMAP_UNCHARTED2, (0, 0, 1), (0.0, 0.0, 0.6524790)),
    (abi.TONEMAP_UNCHARTED2, (0.2, 0.1, 1), (0.1393663, 0.08076570, 0.4409692)),
]


@pytest.mark.parametrize("kind,pixel,want", TONEMAP_CASES)
def test_tonemap_operators_hand_evaluated(kind, pixel, want):
    img = np.array(pixel, np.float32).reshape(1, 1, 3)
    out, ave, _ = O.tonemap(img, kind)
    y = 0.299 * pixel[0] + 0.587 * pixel[1] + 0.114 * pixel[2]
    assert abs(ave / (y + 1e-5) - 1) < 2e-6
    assert np.allclose(out[0, 0], want, rtol=2e-6, atol=1e-7), (out[0, 0], want)


# ------------------------------------------------------------------- sphere --
def _sphere_scene(spheres):
    d = SceneDesc()
    d.materials = [dict(type=abi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))]
    d.shapes = [Shape(abi.SHAPE_SPHERE, material=0, center=c, radius=r) for c, r in spheres]
    d.lights = []
    d.camera = Camera(position=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), fov=0.7, aspect=1.0)
    return d


def _rays(org, dirs, tmin=1e-3, tmax=1e12):
    r = np.zeros(len(org), dtype=abi.RAY_DTYPE)
    r["org"] = org
    r["dir"] = dirs
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def test_sphere_roots_match_fp64():
    """Sphere::IsHit returns the nearer root above F_HIT_MIN: from outside the
    near root b - sqrt(D4), from inside the far root b + sqrt(D4), where
    b = (c - o).d and D4 = b^2 - |c - o|^2 + r^2; misses when D4 < 0 or both
    roots are below F_HIT_MIN.  Checked against the float64 roots of
    |o + t d - c| = r on random rays."""
    c, rad = np.array([0.3, -0.2, 0.1]), 1.25
    osc = O.OracleScene(_sphere_scene([(tuple(c), rad)]))
    rng = np.random.default_rng(7)
    n = 4000
    org = rng.uniform(-3, 3, (n, 3))
    org[: n // 4] = c + rng.uniform(-0.6, 0.6, (n // 4, 3))          # a quarter start inside
    tgt = c + rng.normal(scale=1.0, size=(n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org32, d32 = org.astype(np.float32), d.astype(np.float32)
    hits, shapes = osc.trace(_rays(org32, d32))
    o64, d64 = org32.astype(np.float64), d32.astype(np.float64)
    po = c.astype(np.float32).astype(np.float64) - o64
    aa = (d64 * d64).sum(1)
    b = (po * d64).sum(1)
    disc = b * b - aa * ((po * po).sum(1) - np.float32(rad) ** 2)
    inside = (po * po).sum(1) < rad * rad
    checked = 0
    for i in range(n):
        if disc[i] < -1e-4:
            assert hits[i]["prim"] == -1
            continue
        if disc[i] < 1e-4:
            continue                                    # tangent within rounding: either answer
        t1 = (b[i] - math.sqrt(disc[i])) / aa[i]
        t2 = (b[i] + math.sqrt(disc[i])) / aa[i]
        want = t1 if t1 > 1e-3 else (t2 if t2 > 1e-3 else None)
        if want is None:
            assert hits[i]["prim"] == -1
            continue
        assert hits[i]["prim"] == -2 and shapes[i] == 0
        assert abs(hits[i]["t"] - want) <= 2e-5 * want + 1e-5, (i, hits[i]["t"], want)
        if inside[i]:
            assert want == t2
        checked += 1
    assert checked > 2000


def test_sphere_tangent_origin_on_surface_and_behind():
    osc = O.OracleScene(_sphere_scene([((0.0, 0.0, 0.0), 1.0)]))
    r = _rays(np.array([[0.0, 1.0 - 2 ** -10, -5.0],     # passes just inside the rim
                        [0.0, 1.0 + 2 ** -10, -5.0],     # just outside
                        [0.0, 0.0, -1.0],                # on the surface, looking in: far side
                        [0.0, 0.0, 3.0],                 # sphere behind the origin
                        [0.0, 0.0, -3.0]], np.float32),
              np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32))
    hits, _ = osc.trace(r)
    assert hits["prim"].tolist() == [-2, -1, -2, -1, -2]
    # rim ray: t = 5 - sqrt(1 - y^2), y = 1 - 2^-10; D4 = b^2 - |po|^2 + r^2
    # cancels 25 against 25.998 in float (ulp 2e-6 of D4 = 0.00195), so the
    # root carries an error of about 2e-5
    assert abs(hits[0]["t"] - (5.0 - math.sqrt(1 - (1 - 2 ** -10) ** 2))) < 5e-5
    # origin on the surface: t1 = 0 < F_HIT_MIN, so the far root t2 = 2
    assert hits[2]["t"] == np.float32(2.0)
    assert hits[4]["t"] == np.float32(2.0)


def test_later_sphere_wins_a_tie():
    """Spheres reject only dist > record.distance (s3d_sphere.cpp:83), so of
    identical spheres the last in shape order keeps the hit; triangles reject
    t >= distance, so the first keeps it."""
    osc = O.OracleScene(_sphere_scene([((0.0, 0.0, 0.0), 1.0)] * 3))
    hits, shapes = osc.trace(_rays(np.array([[0.1, 0.2, -4.0]], np.float32), np.array([[0, 0, 1]], np.float32)))
    assert hits[0]["prim"] == -2 and shapes[0] == 2
    # a smaller sphere in front still wins by distance whatever its order
    osc = O.OracleScene(_sphere_scene([((0.0, 0.0, -2.0), 0.5), ((0.0, 0.0, 0.0), 1.0)]))
    hits, shapes = osc.trace(_rays(np.array([[0.0, 0.0, -4.0]], np.float32), np.array([[0, 0, 1]], np.float32)))
    assert shapes[0] == 0 and hits[0]["t"] == np.float32(1.5)


# ----------------------------------------------------------------- instance --
def _rot(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis /= np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(ang), math.sin(ang)
    R = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    return R


def test_rigid_instances_hit_like_pretransformed_meshes():
    """Instance::IsHit moves the ray to local space with the inverse world
    matrix and renormalises the direction (s3d_instance.cpp:65-72).  For a
    rigid world matrix the local ray is the world ray seen from the object,
    so an instance hits the same triangles at the same distances (world
    units: a rigid transform keeps lengths) as the mesh transformed into the
    world; only the rounding differs."""
    rng = np.random.default_rng(3)
    from salty_amd import scenes
    src = scenes.triangle_soup(400, seed=5).meshes[0].vertices
    R = _rot(rng.normal(size=3), 0.7)
    T = np.array([0.4, -1.1, 2.3])
    world = np.eye(4)
    world[:3, :3] = R.T          # row-vector convention: p' = p W, W[:3,:3] = R^T
    world[3, :3] = T
    moved = src.copy()
    p = src["position"].astype(np.float64)
    moved["position"] = p @ R.T + T
    cam = Camera(position=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), fov=0.7, aspect=1.0)
    mat = [dict(type=abi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))]
    inst = SceneDesc(meshes=[Mesh(src)], shapes=[Shape(abi.SHAPE_INSTANCE, mesh=0, world=world.astype(np.float32))],
                     materials=mat, lights=[], camera=cam)
    flat = SceneDesc(meshes=[Mesh(moved)], shapes=[Shape(abi.SHAPE_MESH, mesh=0)], materials=mat, lights=[],
                     camera=cam)
    n = 3000
    # aim at points inside random triangles (world space)
    tri = p.reshape(-1, 3, 3)
    k = rng.integers(len(tri), size=n)
    uv = rng.uniform(0.05, 0.45, (n, 2))
    local = tri[k, 0] + uv[:, :1] * (tri[k, 1] - tri[k, 0]) + uv[:, 1:] * (tri[k, 2] - tri[k, 0])
    tgt = local @ R.T + T
    org = tgt + rng.normal(size=(n, 3)) * 2.0
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = _rays(org.astype(np.float32), d.astype(np.float32))
    hi, _ = O.OracleScene(inst).trace(rays)
    hf, _ = O.OracleScene(flat).trace(rays)
    both = (hi["prim"] >= 0) & (hf["prim"] >= 0)
    assert both.sum() > 300
    same = hi["prim"][both] == hf["prim"][both]
    assert same.mean() > 0.99                      # edge-grazing rays may differ
    ti, tf = hi["t"][both][same], hf["t"][both][same]
    assert np.all(np.abs(ti - tf) <= 1e-4 * tf + 1e-5)
    assert np.mean((hi["prim"] >= 0) == (hf["prim"] >= 0)) > 0.99


def test_scaled_instance_reports_local_distance():
    """For world = diag(2, 2, 2, 1) the reference's inverse is diag(4, 4, 4, 8)
    (its cofactor precedence quirk), which still maps positions to p / 2 after
    TransformCoord's division by w; TransformNormal gives 4 d, and UnitVector
    renormalises it, so the local ray is the world ray at half scale and the
    hit distance is reported in local units: t_world / 2 (without the
    renormalisation it would be t_world / 8)."""
    from salty_amd import scenes
    rng = np.random.default_rng(4)
    src = scenes.triangle_soup(300, seed=6).meshes[0].vertices
    world = np.diag([2.0, 2.0, 2.0, 1.0]).astype(np.float32)
    big = src.copy()
    big["position"] = src["position"] * 2.0
    cam = Camera(position=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), fov=0.7, aspect=1.0)
    mat = [dict(type=abi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))]
    inst = SceneDesc(meshes=[Mesh(src)], shapes=[Shape(abi.SHAPE_INSTANCE, mesh=0, world=world)],
                     materials=mat, lights=[], camera=cam)
    flat = SceneDesc(meshes=[Mesh(big)], shapes=[Shape(abi.SHAPE_MESH, mesh=0)], materials=mat, lights=[],
                     camera=cam)
    n = 2000
    tri = big["position"].astype(np.float64).reshape(-1, 3, 3)
    k = rng.integers(len(tri), size=n)
    uv = rng.uniform(0.05, 0.45, (n, 2))
    tgt = tri[k, 0] + uv[:, :1] * (tri[k, 1] - tri[k, 0]) + uv[:, 1:] * (tri[k, 2] - tri[k, 0])
    org = tgt + rng.normal(size=(n, 3)) * 3.0
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = _rays(org.astype(np.float32), d.astype(np.float32))
    hi, _ = O.OracleScene(inst).trace(rays)
    hf, _ = O.OracleScene(flat).trace(rays)
    both = (hi["prim"] >= 0) & (hf["prim"] >= 0) & (hi["prim"] == hf["prim"])
    assert both.sum() > 0.9 * n
    assert np.all(np.abs(hi["t"][both] - hf["t"][both] / 2) <= 1e-5 * hf["t"][both])
