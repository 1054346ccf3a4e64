"""The oracle's RNG and trigonometry against published / library values (CPU)."""
import math

import numpy as np

from helios_amd import scenes
from oracle import oracle as O


def test_tea_published_vector():
    # TEA (Wheeler & Needham 1994), 32 rounds, key 0, plaintext 0 -> 41ea3a0a 94baa940
    v0, v1 = O.tea(0, 0, 32, (0, 0, 0, 0))
    assert (v0, v1) == (0x41EA3A0A, 0x94BAA940)


def test_tea16_is_tea_with_helios_key():
    key = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)  # R/shaders/common/random.glsl:17-18
    for a, b in [(0, 0), (1, 2), (123456789, 987654321), (0xFFFFFFFF, 7)]:
        assert O.tea16(a, b) == O.tea(a, b, 16, key)[0]


def test_tea16_avalanche():
    # a one-bit change of the index flips about half of the 32 output bits
    flips = [bin(O.tea16(i, 99) ^ O.tea16(i ^ 1, 99)).count("1") for i in range(0, 2000, 2)]
    assert 14 < np.mean(flips) < 18


def test_lcg_numerical_recipes_sequence():
    # Numerical Recipes "quick and dirty" generator ranqd1 from idum = 0
    expect = [0x3C6EF35F, 0x47502932, 0xD1CCF6E9, 0xAAF95334, 0x6252E503, 0x9F2EC686]
    x, got = 0, []
    for _ in expect:
        x = O.lcg_next(x)
        got.append(x)
    assert got == expect


def test_rnd_24bit_uniform():
    vals = np.array(O.rnd_stream(12345, 20000), np.float64)
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert np.all(vals * 16777216.0 == np.round(vals * 16777216.0))  # exact 24-bit values
    assert abs(vals.mean() - 0.5) < 0.01


def test_sincos_2pi_against_libm():
    worst = 0.0
    for r in np.linspace(0.0, 1.0, 200001, endpoint=False, dtype=np.float32):
        s, c = O.sincos_2pi(float(r))
        a = 2.0 * math.pi * float(r)
        worst = max(worst, abs(s - math.sin(a)), abs(c - math.cos(a)))
    assert worst < 2.5e-7


def test_scene_generator_uses_the_same_sincos():
    for r in np.linspace(0.0, 1.0, 4001, endpoint=False, dtype=np.float32):
        assert O.sincos_2pi(float(r)) == scenes.sincos_2pi(float(r))
