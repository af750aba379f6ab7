"""Random-number streams of the oracle against published known answers.

* PCG32 XSH-RR: the pcg32-demo output of the published PCG library
  (O'Neill 2014) for pcg32_srandom_r(42, 54) pins the LCG multiplier, the
  output permutation and the step the reference's PCG class uses
  (src/include/s3d_math.h:2769-2817).
* Philox4x32-10: the Random123 known-answer vectors (Salmon et al., SC'11)
  pin the counter-based generator that keys every path's stream.
"""
import numpy as np

from oracle import oracle as O


def test_pcg32_published_demo_vector():
    got = O.pcg_srandom_stream(42, 54, 6)
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert [int(x) for x in got] == want


def test_pcg_reference_seeding_is_state_plus_increment_then_one_step():
    # SetSeed(s): state = s + inc, then one GetAsU32.  With inc fixed, the
    # stream of SetSeed(s) equals the published generator started from state
    # s + inc after one step -- computed here with Python integers.
    mul, inc = 6364136223846793005, 1442695040888963407
    mask = (1 << 64) - 1

    def out(old):
        xs = (((old >> 18) ^ old) >> 27) & 0xffffffff
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xffffffff

    for seed in (0, 1, 3141592, 2 ** 63 + 12345):
        state = (seed + inc) & mask
        state = (state * mul + inc) & mask      # the step SetSeed takes
        want = []
        for _ in range(8):
            want.append(out(state))
            state = (state * mul + inc) & mask
        assert [int(x) for x in O.pcg_stream(seed, 8)] == want


def test_pcg_f32_is_u32_over_2_pow_32_and_can_reach_one():
    lib = O.lib()
    assert lib.orc_pcg_f32(0) == 0.0
    assert lib.orc_pcg_f32(0xffffffff) == 1.0          # float(0xffffffff) rounds to 2^32
    assert lib.orc_pcg_f32(1 << 31) == 0.5
    x = 123456789
    assert lib.orc_pcg_f32(x) == float(np.float32(x)) / 4294967296.0


def test_philox_known_answers():
    assert [int(x) for x in O.philox([0, 0, 0, 0], [0, 0])] == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert [int(x) for x in O.philox([0xffffffff] * 4, [0xffffffff] * 2)] == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert [int(x) for x in O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                     [0xa4093822, 0x299f31d0])] == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_path_seed_is_philox_of_pixel_pass_scene():
    lib = O.lib()
    seed = 0x1234_5678_9abc_def0
    o = O.philox([17, 3, 2, 0], [seed & 0xffffffff, seed >> 32])
    assert lib.orc_path_seed(seed, 17, 3, 2) == int(o[0]) | (int(o[1]) << 32)
    # distinct counters give distinct streams
    seeds = {lib.orc_path_seed(0, p, s, 0) for p in range(64) for s in range(16)}
    assert len(seeds) == 64 * 16
