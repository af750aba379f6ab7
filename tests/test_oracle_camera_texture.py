"""Oracle cameras and textures, pinned by closed forms and the properties the
reference's code implies: the pinhole frame and its field of view
(PinholeCamera::Update / GetRay, s3d_camera.h:63-100), the thin lens focusing
every lens sample on one point (ThinLensCamera, s3d_camera.h:137-268),
bilinear texel lookup with its abs(i % size) wrap and the latitude-longitude
map (Texture::SampleColor, s3d_texture.cpp:925-972)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi, scenes


def _scene_with_camera(cam):
    desc = scenes.cornell_box()
    desc.camera = cam
    return O.OracleScene(desc)


def test_pinhole_axis_field_of_view_and_row_order():
    fov, aspect = 0.5, 1.5
    sc = _scene_with_camera(abi.Camera(position=(0, 0, 0), target=(0, 0, -5), fov=fov, aspect=aspect))
    o, d, _ = sc.camera_ray(0.0, 0.0)
    assert np.allclose(o, 0) and np.allclose(d, [0, 0, -1], atol=1e-7)
    # image-plane edges: half-angles atan(fov / 2) vertically and
    # atan(fov * aspect / 2) horizontally (near plane at distance 1)
    _, top, _ = sc.camera_ray(0.0, -0.5)
    _, right, _ = sc.camera_ray(0.5, 0.0)
    assert top[1] > 0                                   # row 0 (y = -0.5) looks up: CY points down
    assert abs(math.atan2(top[1], -top[2]) - math.atan(fov / 2)) < 1e-6
    assert right[0] > 0
    assert abs(math.atan2(right[0], -right[2]) - math.atan(fov * aspect / 2)) < 1e-6


def test_pinhole_general_frame_is_orthogonal():
    """For any position, target and up vector the image plane spans
    unit(dir x up) * fov * aspect and unit(dir x CX) * fov at distance near."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        pos = rng.normal(size=3) * 3
        tgt = pos + rng.normal(size=3)
        up = rng.normal(size=3)
        fov, aspect, near = 0.3 + rng.random(), 0.5 + rng.random(), 0.5 + rng.random()
        sc = _scene_with_camera(abi.Camera(position=tuple(pos), target=tuple(tgt), upward=tuple(up), fov=fov,
                                           aspect=aspect, near_clip=near))
        fwd = (tgt - pos) / np.linalg.norm(tgt - pos)
        cx = np.cross(fwd, up)
        cx = cx / np.linalg.norm(cx) * fov * aspect
        cy = np.cross(fwd, cx)
        cy = cy / np.linalg.norm(cy) * fov
        for x, y in ((0.3, -0.2), (-0.5, 0.5), (0.1, 0.45)):
            o, d, _ = sc.camera_ray(x, y)
            want = fwd * near + cx * x + cy * y
            want /= np.linalg.norm(want)
            assert np.allclose(o, pos, atol=1e-5)
            assert np.allclose(d, want, atol=2e-5)


def test_thin_lens_focuses_every_lens_sample_on_one_point():
    """Lens samples start uniformly on the disc of the lens radius around the
    camera (two draws each) and all pass through the pinhole ray's point on
    the focus plane z = -focal (focal = distance to the target)."""
    focal, radius = 4.0, 0.3
    sc = _scene_with_camera(abi.Camera(position=(0, 0, 0), target=(0, 0, -focal), fov=0.6, aspect=1.0,
                                       type=abi.CAMERA_THINLENS, lens_radius=radius))
    mul, inc, mask = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    for x, y in ((0.0, 0.0), (0.4, -0.3), (-0.2, 0.5)):
        _, d0, _ = sc.camera_ray(x, y)
        focus = d0 * (focal / abs(d0[2]))
        state, r2 = 17, []
        for _ in range(4000):
            o, d, s1 = sc.camera_ray(x, y, state)
            assert s1 == (((state * mul + inc) & mask) * mul + inc) & mask      # two draws
            state = s1
            assert o[2] == 0 and math.hypot(o[0], o[1]) <= radius * (1 + 1e-6)
            assert np.linalg.norm(np.cross(focus - o, d)) < 2e-5 * np.linalg.norm(focus - o)
            r2.append(o[0] ** 2 + o[1] ** 2)
        # uniform on the disc: E[r^2] = R^2 / 2
        assert abs(np.mean(r2) - radius ** 2 / 2) < 4 * radius ** 2 * math.sqrt(1 / 12 / len(r2))


def _linear_texture(w, h):
    x = np.arange(w, dtype=np.float32)[None, :, None]
    y = np.arange(h, dtype=np.float32)[:, None, None]
    return (np.float32([0.25, -0.5, 1.0]) * x + np.float32([0.75, 0.125, -0.25]) * y
            + np.float32([2.0, 3.0, 5.0])).astype(np.float32)


def _linear_value(w, h, u, v):
    fx, fy = u * (w - 1), v * (h - 1)
    return np.array([0.25, -0.5, 1.0]) * fx + np.array([0.75, 0.125, -0.25]) * fy + np.array([2.0, 3.0, 5.0])


def test_bilinear_reproduces_texels_and_linear_functions():
    w, h = 7, 5
    tex = _linear_texture(w, h)
    for i in range(w):
        for j in range(h):
            assert np.allclose(O.texture_sample(tex, i / (w - 1), j / (h - 1)), tex[j, i], atol=1e-5)
    rng = np.random.default_rng(1)
    for u, v in rng.random((200, 2)):
        assert np.allclose(O.texture_sample(tex, u, v), _linear_value(w, h, u, v), atol=1e-4)


def test_wrap_is_abs_of_remainder():
    """Coordinates outside [0, 1]: texel index i maps to abs(i % size), which
    mirrors index -1 to 1 (not to size - 1) and wraps size to 0."""
    w, h = 6, 4
    rng = np.random.default_rng(2)
    tex = rng.random((h, w, 3)).astype(np.float32)
    got = O.texture_sample(tex, -0.5 / (w - 1), 0.0)
    assert np.allclose(got, 0.5 * tex[0, 1] + 0.5 * tex[0, 0], atol=1e-6)
    got = O.texture_sample(tex, 1.0 + 0.5 / (w - 1), 0.0)
    assert np.allclose(got, 0.5 * tex[0, w - 1] + 0.5 * tex[0, 0], atol=1e-6)
    got = O.texture_sample(tex, 0.0, -1.0 / (h - 1))
    assert np.allclose(got, tex[1, 0], atol=1e-6)


@pytest.mark.parametrize("d,u,v", [
    ((0.0, 1.0, 0.0), 0.0, 0.0),
    ((0.0, -1.0, 0.0), 0.0, 1.0),
    ((1.0, 0.0, 0.0), 0.0, 0.5),              # z == 0: the longitude is left at 0
    ((0.0, 0.0, 1.0), 0.0, 0.5),              # x == 0: likewise
    ((1.0, 0.0, 1.0), 0.125, 0.5),
    ((-1.0, 0.0, 1.0), 0.375, 0.5),
    ((-1.0, 0.0, -1.0), 0.625, 0.5),
    ((1.0, 0.0, -1.0), 0.875, 0.5),
    ((1.0, 1.0, 1.0), 0.125, math.acos(1 / math.sqrt(3)) / math.pi),
])
def test_latitude_longitude_lookup(d, u, v):
    w, h = 9, 5
    tex = _linear_texture(w, h)
    dd = np.array(d, np.float32)
    dd /= np.linalg.norm(dd)
    assert np.allclose(O.texture_sample_dir(tex, dd), _linear_value(w, h, u, v), atol=2e-4)
