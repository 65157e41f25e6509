"""Pins the oracle's random-number stream and trigonometry
(assets/shader/utils/glsl/rendering_utils.glsl:361-392, ray_tracing_utils.glsl:178-182)."""
import math

import numpy as np


def tea_published(v0, v1, key, rounds):
    """Tiny Encryption Algorithm (Wheeler & Needham 1994), encryption rounds."""
    M = 0xFFFFFFFF
    s = 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & M
        v0 = (v0 + ((((v1 << 4) & M) + key[0]) ^ ((v1 + s) & M) ^ ((v1 >> 5) + key[1]))) & M
        v1 = (v1 + ((((v0 << 4) & M) + key[2]) ^ ((v0 + s) & M) ^ ((v0 >> 5) + key[3]))) & M
    return v0, v1


def test_tea_published_vector():
    # the published all-zero TEA known answer pins the round function used below
    assert tea_published(0, 0, [0, 0, 0, 0], 32) == (0x41EA3A0A, 0x94BAA940)


def test_tea16_matches_published_algorithm(oracle):
    key = [0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E]
    rng = np.random.default_rng(1)
    for v0, v1 in [(0, 0), (1, 1234), (0xFFFFFFFF, 7)] + [tuple(int(x) for x in rng.integers(0, 2**32, 2)) for _ in range(200)]:
        assert oracle.tea16(v0, v1) == tea_published(v0, v1, key, 16)[0]


def test_lcg_numerical_recipes_sequence(oracle):
    # Numerical Recipes quick-and-dirty generator from seed 0:
    # 1013904223, 1196435762, 3519870697, 2868466484
    expect = [1013904223, 1196435762, 3519870697, 2868466484]
    f = oracle.rnd_stream(0, 4)
    got = [float(x) for x in f]
    want = [(e >> 9) / 2.0**23 for e in expect]
    assert got == want


def test_rnd_is_exact_mantissa_map(oracle):
    seed = 0xDEADBEEF
    f = oracle.rnd_stream(seed, 10000)
    s = seed
    for i in range(10000):
        s = (1664525 * s + 1013904223) & 0xFFFFFFFF
        assert f[i] == np.float32((s >> 9) * 2.0**-23)
    assert f.min() >= 0.0 and f.max() < 1.0


def test_sincos_turns_accuracy(oracle):
    us = np.concatenate([np.linspace(0, 1, 4097, endpoint=False, dtype=np.float32),
                         np.random.default_rng(0).random(4000, dtype=np.float32)])
    worst = 0.0
    for u in us:
        s, c = oracle.sincos_turns(float(u))
        th = 2.0 * math.pi * float(u)
        worst = max(worst, abs(s - math.sin(th)), abs(c - math.cos(th)))
    assert worst < 3e-7
    assert oracle.sincos_turns(0.0) == (0.0, 1.0)
    s, c = oracle.sincos_turns(0.25)
    assert (s, c) == (1.0, 0.0) or (abs(s - 1) < 1e-7 and abs(c) < 1e-7)
