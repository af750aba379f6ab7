"""Oracle materials and sampling, pinned by closed forms and statistics the
reference's code implies (not by re-typing its formulas)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import abi


def test_tangent_space_is_orthonormal_and_matches_duff_special_cases():
    t, b = O.tangent_space([0.0, 0.0, 1.0])
    assert np.array_equal(t, [1, 0, 0]) and np.array_equal(b, [0, 1, 0])
    t, b = O.tangent_space([0.0, 0.0, -1.0])
    assert np.array_equal(t, [1, 0, 0]) and np.array_equal(b, [0, -1, 0])
    rng = np.random.default_rng(0)
    for _ in range(500):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        t, b = O.tangent_space(n)
        m = np.stack([t, b, n]).astype(np.float64)
        assert np.allclose(m @ m.T, np.eye(3), atol=2e-6)
        assert np.allclose(np.cross(t, b), n, atol=2e-6)        # right-handed frame


def test_sample_lambert_closed_forms():
    assert np.allclose(O.sample_lambert(1.0, 0.3), [0, 0, 1])
    assert np.allclose(O.sample_lambert(0.0, 0.25), [0, 1, 0], atol=1e-6)
    assert np.allclose(O.sample_lambert(0.6, 0.0), [0.8, 0, 0.6], atol=1e-6)


LAMBERT = dict(type=abi.MAT_LAMBERT, diffuse=(0.6, 0.3, 0.2))


def test_lambert_samples_the_hemisphere_uniformly_with_uncompensated_roulette():
    """Lambert::Shade draws SampleLambert(u, v) with z = u: directions are
    uniform over the hemisphere (E[cos] = 1/2, not the 2/3 of cosine
    sampling), the weight is the albedo itself, and dice = rand >= max(albedo)."""
    n = np.array([0.0, 0.0, 1.0], np.float32)
    inp = np.array([0.3, 0.1, -0.9], np.float32)
    inp /= np.linalg.norm(inp)
    state = 12345
    cos, dice = [], 0
    for _ in range(20000):
        o, w, d, state = O.shade(LAMBERT, inp, n, state)
        assert np.array_equal(w, np.float32([0.6, 0.3, 0.2]))
        assert abs(np.linalg.norm(o) - 1) < 1e-5 and o[2] >= 0
        cos.append(o[2])
        dice += d
    cos = np.array(cos)
    assert abs(cos.mean() - 0.5) < 4 * math.sqrt(1 / 12 / len(cos))
    p = 1 - 0.6
    assert abs(dice / 20000 - p) < 4 * math.sqrt(p * (1 - p) / 20000)


def test_lambert_flips_normal_towards_the_incoming_ray():
    n = np.array([0.0, 0.0, 1.0], np.float32)
    o, _, _, _ = O.shade(LAMBERT, np.float32([0, 0, 1]), n, 99)   # ray arrives from below
    assert o[2] <= 0


def test_mirror_reflects_exactly():
    n = np.array([0.0, 1.0, 0.0], np.float32)
    i = np.array([0.6, -0.8, 0.0], np.float32)
    o, w, d, st = O.shade(dict(type=abi.MAT_MIRROR, specular=(0.9, 0.8, 0.7)), i, n, 5)
    assert np.allclose(o, [0.6, 0.8, 0.0], atol=1e-7) and not d and st == 5
    assert np.array_equal(w, np.float32([0.9, 0.8, 0.7]))


GLASS = dict(type=abi.MAT_GLASS, specular=(1.0, 1.0, 1.0), ior=1.5)


def test_glass_normal_incidence_weights_include_the_half_probability_factor():
    """At normal incidence entering glass: R0 = Re = 0.04, P = 0.25 + 0.5 Re = 0.27,
    reflect weight Re / (P * 0.5) = 0.2963, refract weight (1-Re) / ((1-P) * 0.5)
    = 2.6301 (the reference divides by an extra 0.5)."""
    n = np.array([0.0, 0.0, 1.0], np.float32)
    seen = set()
    state = 1
    for _ in range(200):
        o, w, d, state = O.shade(GLASS, np.float32([0, 0, -1]), n, state)
        assert not d
        if o[2] > 0:
            assert np.allclose(w, 0.04 / (0.27 * 0.5), rtol=1e-5)
            seen.add("reflect")
        else:
            assert np.allclose(o, [0, 0, -1], atol=1e-6)
            assert np.allclose(w, 0.96 / (0.73 * 0.5), rtol=1e-5)
            seen.add("refract")
    assert seen == {"reflect", "refract"}


def test_glass_obeys_snell_and_total_internal_reflection():
    n = np.array([0.0, 0.0, 1.0], np.float32)
    th = math.radians(40)
    inp = np.float32([math.sin(th), 0, -math.cos(th)])
    state = 3
    for _ in range(100):
        o, w, d, state = O.shade(GLASS, inp, n, state)
        if o[2] < 0:          # refracted into the glass
            sin_t = math.hypot(o[0], o[1])
            assert abs(sin_t - math.sin(th) / 1.5) < 1e-5
    # leaving the glass at 60 degrees (beyond the 41.8 degree critical angle)
    th = math.radians(60)
    inp = np.float32([math.sin(th), 0, math.cos(th)])    # travelling outwards along +z
    o, w, d, _ = O.shade(GLASS, inp, n, 7)
    assert o[2] < 0 and np.allclose(w, 1.0)               # reflected back inside, weight = specular


def test_shade_consumes_the_documented_number_of_draws():
    n = np.float32([0, 0, 1])
    inp = np.float32([0, 0, -1])

    def draws(mat):
        s0 = 777
        _, _, _, s1 = O.shade(mat, inp, n, s0)
        ref = O.pcg_stream(0, 1)   # warm the library
        # count PCG steps by stepping a copy of the state with the published LCG
        mul, inc, mask = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
        s, k = s0, 0
        while s != s1:
            s = (s * mul + inc) & mask
            k += 1
            assert k < 16
        return k

    assert draws(LAMBERT) == 3
    assert draws(dict(type=abi.MAT_MIRROR, specular=(1, 1, 1))) == 0
    assert draws(GLASS) == 1
    assert draws(dict(type=abi.MAT_PHONG, specular=(0.5, 0.5, 0.5), power=20.0)) == 3
    assert draws(dict(type=abi.MAT_PLASTIC, diffuse=(0.5, 0.5, 0.5), specular=(0.5, 0.5, 0.5), power=20)) == 4
