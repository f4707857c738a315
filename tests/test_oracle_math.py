"""Pins of the oracle's RNG and math (CPU only).

tea (Random.glsl:7-21) is TEA encipherment (Wheeler & Needham 1994) with 16
cycles and the shader's key; it is checked through a parameterised TEA whose
32-cycle form reproduces the published all-zero test vector.  lcg/rnd
(Random.glsl:26-39) follow the Numerical Recipes LCG.  The oracle's own
sin/cos/exp2 are checked against float64 numpy within 2 ulp.
"""
import numpy as np

import oracle

M32 = 0xFFFFFFFF
SHADER_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)


def tea_encipher(v0, v1, key, cycles):
    s = 0
    for _ in range(cycles):
        s = (s + 0x9E3779B9) & M32
        v0 = (v0 + ((((v1 << 4) & M32) + key[0]) ^ ((v1 + s) & M32) ^ ((v1 >> 5) + key[1]))) & M32
        v1 = (v1 + ((((v0 << 4) & M32) + key[2]) ^ ((v0 + s) & M32) ^ ((v0 >> 5) + key[3]))) & M32
    return v0, v1


def test_tea_reference_vector_pins_the_cipher():
    # Published TEA test vector: key 0, plaintext 0 -> 41ea3a0a 94baa940 (32 cycles).
    assert tea_encipher(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)


def test_tea_oracle_matches_cipher():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 2**32, 500, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, 500, dtype=np.uint64).astype(np.uint32)
    a[:4] = [0, 1, 0xFFFFFFFF, 12345]
    b[:4] = [0, 0, 0xFFFFFFFF, 678]
    got = oracle.tea(a, b)
    exp = [tea_encipher(int(x), int(y), SHADER_KEY, 16)[0] for x, y in zip(a, b)]
    assert got.tolist() == exp


def test_lcg_rnd_sequence():
    vals, state = oracle.rnd_sequence(0, 4)
    # Numerical Recipes LCG from 0: 1013904223, 1196435762, 3519870697, 2868466484
    seq = [1013904223, 1196435762, 3519870697, 2868466484]
    assert state == seq[-1]
    exp = [(x & 0xFFFFFF) / 16777216.0 for x in seq]
    assert vals.tolist() == exp
    big, _ = oracle.rnd_sequence(987654321, 100000)
    assert big.min() >= 0.0 and big.max() < 1.0
    assert abs(big.mean() - 0.5) < 0.01


def _ulp_err(got, exact):
    exact32 = exact.astype(np.float32)
    ulp = np.spacing(np.abs(exact32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / np.maximum(ulp, np.spacing(np.float32(1e-30)))


def test_sincos_within_two_ulp():
    x = np.concatenate([np.linspace(0, 2 * np.pi, 200001), np.linspace(-4, 4, 100001)]).astype(np.float32)
    s, c = oracle.sincos(x)
    xd = x.astype(np.float64)
    es, ec = np.sin(xd), np.cos(xd)
    # absolute error relative to 1 (the reduction is Cody-Waite, exact for |k| <= 4)
    assert np.max(np.abs(s - es)) < 2.5e-7
    assert np.max(np.abs(c - ec)) < 2.5e-7
    big = np.abs(es) > 1e-3
    assert np.max(_ulp_err(s[big], es[big])) <= 2.0
    big = np.abs(ec) > 1e-3
    assert np.max(_ulp_err(c[big], ec[big])) <= 2.0
    s0, c0 = oracle.sincos(np.zeros(1))
    assert s0[0] == 0.0 and c0[0] == 1.0


def test_exp2_within_two_ulp():
    y = np.linspace(-20, 8, 300001).astype(np.float32)
    got = oracle.exp2(y)
    assert np.max(_ulp_err(got, np.exp2(y.astype(np.float64)))) <= 2.0
    ints = np.arange(-20, 9).astype(np.float32)
    assert (oracle.exp2(ints) == np.exp2(ints)).all()
    assert oracle.exp2(np.array([-200.0], np.float32))[0] == 0.0
