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


def _normal_incidence_samples(mat, count, state=4242):
    n = np.float32([0, 0, 1])
    inp = np.float32([0, 0, -1])
    out = []
    for _ in range(count):
        o, w, d, state = O.shade(mat, inp, n, state)
        out.append((o, w, d))
    return out


@pytest.mark.parametrize("power", [1.0, 10.0])
def test_phong_lobe_at_normal_incidence(power):
    """At normal incidence the Phong lobe is centred on the normal, with
    cos(theta) = (1 - v)^(1/(n+1)): E[cos] = (n+1)/(n+2); the weight is the
    specular colour times cos; dice = rand >= max(specular)."""
    spec = (0.5, 0.7, 0.2)
    samples = _normal_incidence_samples(dict(type=abi.MAT_PHONG, specular=spec, power=power), 20000)
    cos = np.array([o[2] for o, _, _ in samples], np.float64)
    assert (cos >= 0).all()
    want = (power + 1) / (power + 2)
    sd = math.sqrt((power + 1) / (power + 3) - want ** 2)   # E[cos^2] = (n+1)/(n+3)
    assert abs(cos.mean() - want) < 4 * sd / math.sqrt(len(cos))
    w = np.array([w for _, w, _ in samples], np.float64)
    assert np.allclose(w.mean(0), np.array(spec) * want, rtol=0.02)
    p = 1 - 0.7
    dice = np.mean([d for _, _, d in samples])
    assert abs(dice - p) < 4 * math.sqrt(p * (1 - p) / len(samples))


def test_phong_high_power_is_a_mirror():
    n = np.float32([0, 0, 1])
    inp = np.float32([0.6, 0.0, -0.8])
    mirror = np.float32([0.6, 0.0, 0.8])
    state = 5
    for _ in range(200):
        o, w, _, state = O.shade(dict(type=abi.MAT_PHONG, specular=(1, 1, 1), power=1e6), inp, n, state)
        assert np.linalg.norm(o - mirror) < 1e-2
        assert abs(w[0] - 0.8) < 1e-2


def test_plastic_lobes_at_normal_incidence():
    """Schlick with R0 = 1/2: at normal incidence R = P = 1/2, so the diffuse
    lobe (uniform hemisphere) is taken half the time with weight exactly the
    diffuse colour, the specular lobe otherwise with weight specular * cos,
    E[cos] = (n+1)/(n+2); both lobes roll their dice against the diffuse
    threshold (the reference copies it into the specular one)."""
    D, S, power = (0.3, 0.0, 0.0), (0.0, 0.9, 0.0), 3.0
    samples = _normal_incidence_samples(dict(type=abi.MAT_PLASTIC, diffuse=D, specular=S, power=power), 20000)
    diffuse = [s for s in samples if s[1][0] > 0]
    specular = [s for s in samples if s[1][0] == 0]
    assert len(diffuse) + len(specular) == len(samples)
    frac = len(diffuse) / len(samples)
    assert abs(frac - 0.5) < 4 * math.sqrt(0.25 / len(samples))
    assert all(np.array_equal(w, np.float32(D)) for _, w, _ in diffuse)
    cos_d = np.array([o[2] for o, _, _ in diffuse])
    assert abs(cos_d.mean() - 0.5) < 4 * math.sqrt(1 / 12 / len(cos_d))       # uniform hemisphere
    cos_s = np.array([o[2] for o, _, _ in specular])
    assert abs(cos_s.mean() - (power + 1) / (power + 2)) < 0.01
    assert np.allclose([w[1] / o[2] for o, w, _ in specular], 0.9, rtol=1e-5)
    for group in (diffuse, specular):
        rate = np.mean([d for _, _, d in group])
        assert abs(rate - 0.7) < 4 * math.sqrt(0.21 / len(group))           # 1 - max(diffuse)


def test_plastic_grazing_incidence_prefers_the_diffuse_lobe_by_schlick():
    n = np.float32([0, 0, 1])
    c = 0.05
    inp = np.float32([math.sqrt(1 - c * c), 0.0, -c])
    R = 0.5 + 0.5 * (1 - c) ** 5
    P = (R + 0.5) / 2
    state, hits, count = 31, 0, 20000
    for _ in range(count):
        o, w, _, state = O.shade(dict(type=abi.MAT_PLASTIC, diffuse=(0.4, 0, 0), specular=(0, 0.5, 0), power=2.0),
                                 inp, n, state)
        if w[0] > 0:
            hits += 1
            assert abs(w[0] - 0.4 * R / P) < 1e-5
    assert abs(hits / count - P) < 4 * math.sqrt(P * (1 - P) / count)


def test_phong_and_plastic_draw_counts():
    n = np.float32([0, 0, 1])
    inp = np.float32([0, 0, -1])
    mul, inc, mask = 6364136223846793005, 1442695040888963407, (1 << 64) - 1

    def draws(mat):
        s0 = 991
        _, _, _, s1 = O.shade(mat, inp, n, s0)
        s, k = s0, 0
        while s != s1:
            s = (s * mul + inc) & mask
            k += 1
            assert k < 16
        return k

    assert draws(dict(type=abi.MAT_PHONG, specular=(1, 1, 1), power=5.0)) == 3
    assert draws(dict(type=abi.MAT_PLASTIC, diffuse=(0.5, 0.5, 0.5), specular=(0.5, 0.5, 0.5), power=5.0)) == 4
