"""Oracle sampling pinned to the reference's own known-answer tests.

Vectors: Tests/Renderer/SampleStreamTest.cpp:20-50, Tests/Renderer/SobolSamplingTest.cpp:22-47,
Tests/Renderer/OwenScramblingTest.cpp:30-81, Cie1931Sensor.cpp:167-176 (table pins).
"""
import ctypes as C

import numpy as np
import pytest

U64 = 2**64 - 1


def idx(abi, px=0, py=0, s=0, seed=0):
    return abi.SampleStreamIndex(px, py, s, seed)


SAMPLE_STREAM_KATS = [  # (index, dimension, transport numerator, reference numerator)
    ((0, 0, 0, 0), 0, 0x2F37B1, 0x05E6F6339B5611),
    ((17, 29, 0, 0x0123456789ABCDEF), 0, 0xDE27E5, 0x1BC4FCA4264791),
    ((17, 29, 4095, 0x0123456789ABCDEF), 1, 0xDC4661, 0x1B88CC28476D40),
    ((0xFFFFFFFF, 0xFFFFFFFF, U64, U64), U64, 0x8F4FD2, 0x11E9FA55578958),
]


@pytest.mark.parametrize("index,dimension,transport,reference", SAMPLE_STREAM_KATS)
def test_sample_stream_known_answers(oracle, abi, index, dimension, transport, reference):
    i = idx(abi, *index)
    bits = oracle.lib.bfo_indexed_bits(C.byref(i), dimension)
    assert bits >> 40 == transport
    assert bits >> 11 == reference
    assert oracle.lib.bfo_sample_1d(abi.SAMPLER_INDEXED_HASH, C.byref(i), dimension) == np.float32(transport * 2.0**-24)


def test_sample_stream_is_random_access(oracle, abi):
    i = idx(abi, 113, 47, 91, 0xA5A5F00D12345678)
    forward = [oracle.lib.bfo_indexed_bits(C.byref(i), d) for d in range(512)]
    for order in range(512):
        d = (order * 4051 + 17) % 512
        assert oracle.lib.bfo_indexed_bits(C.byref(i), d) == forward[d]


def test_mix_bits_is_splitmix64_finalizer(oracle):
    # closed form: the SplitMix64 output function (Steele et al. 2014) on a few words
    def mix(v):
        v ^= v >> 30
        v = (v * 0xBF58476D1CE4E5B9) & U64
        v ^= v >> 27
        v = (v * 0x94D049BB133111EB) & U64
        return v ^ (v >> 31)

    for v in (0, 1, 0x9E3779B97F4A7C15, U64, 0x0123456789ABCDEF):
        assert oracle.lib.bfo_mix_bits(v) == mix(v)


PUBLISHED_FIRST_TEN = [(0, 0, 0), (8, 8, 8), (12, 4, 4), (4, 12, 12), (6, 6, 10), (14, 14, 2), (10, 2, 14), (2, 10, 6),
                       (3, 5, 15), (11, 13, 7)]


def test_sobol_published_first_ten_points(oracle):
    for s, row in enumerate(PUBLISHED_FIRST_TEN):
        for d in range(3):
            assert oracle.lib.bfo_sobol_bits(s, d) == row[d] << 60


@pytest.mark.parametrize("sample,dimension,bits", [
    (0x0123456789ABCDEF, 63, 0x013E8C6349603C80),
    (0x0123456789ABCDEF, 255, 0xB321991D6689DC80),
    (0x0123456789ABCDEF, 1023, 0xE4BCA0B1ED509A80),
    (U64, 1, U64),
])
def test_sobol_known_answers(oracle, sample, dimension, bits):
    assert oracle.lib.bfo_sobol_bits(sample, dimension) == bits


def test_sobol_is_a_0_m_2_net_in_dimensions_0_1(oracle):
    # dyadic-net property: the first 2^m points put exactly one point in every 2^a x 2^(m-a) box
    m = 8
    for a in range(m + 1):
        seen = set()
        for s in range(2**m):
            x = oracle.lib.bfo_sobol_bits(s, 0) >> (64 - a) if a else 0
            y = oracle.lib.bfo_sobol_bits(s, 1) >> (64 - (m - a)) if m - a else 0
            seen.add((x, y))
        assert len(seen) == 2**m


@pytest.mark.parametrize("index,dimension,bits", [
    ((0, 0, 0, 0), 0, 0x92A8B3FFD11450D8),
    ((0, 0, 1, 0), 0, 0x05425D9FC8F958EF),
    ((0, 0, 2, 0x0123456789ABCDEF), 0, 0x2D5F4AEE875D131D),
    ((0, 0, 0x0123456789ABCDEF, 0x0123456789ABCDEF), 63, 0x952C2015E62C23E3),
    ((113, 47, 0x0123456789ABCDEF, 0xA5A5F00D12345678), 63, 0xA694114D8B9547EF),
    ((0, 0, 0x0123456789ABCDEF, 0xFEDCBA9876543210), 255, 0x210EA3A92DEE7EFC),
])
def test_owen_known_answers(oracle, abi, index, dimension, bits):
    i = idx(abi, *index)
    assert oracle.lib.bfo_owen_bits(C.byref(i), dimension) == bits


def test_owen_scramble_preserves_stratification(oracle, abi):
    # a nested scramble is a bijection on every dyadic level: 256 points stay one per 1/256 bin
    i = idx(abi, 5, 9, 0, 1234)
    bins = set()
    for s in range(256):
        i.sample_index = s
        bins.add(oracle.lib.bfo_owen_bits(C.byref(i), 7) >> 56)
    assert len(bins) == 256


def test_cie_table_pins(oracle):
    assert oracle.lib.bfo_cie_fingerprint() == 0x504F4B16E8279104
    integral = oracle.lib.bfo_cie_y_integral()
    assert 106.8569149166 < integral < 106.8569149170


def test_xyz_to_rgb_of_white_point(oracle):
    # D65 white (X, Y, Z) = (0.95047, 1, 1.08883) maps to RGB ~ (1, 1, 1) through Color.hpp:77-87
    xyz = (C.c_float * 3)(0.95047, 1.0, 1.08883)
    rgb = (C.c_float * 3)()
    oracle.lib.bfo_xyz_to_rgb(xyz, rgb)
    assert np.allclose(list(rgb), [1.0, 1.0, 1.0], atol=2e-4)


def test_oracle_dump_matches_the_frozen_canonical_fingerprint(oracle, abi):
    """Tests/CUDA/SampleStreamTest.cpp:401-414: the canonical text of the reference's dump requests
    (indexed bits, binary32 sample, bounce dimensions, manifest), computed from the oracle, hashes
    to the reference's frozen FNV-1a-64. A wrong salt, sample shift or dimension offset changes it."""
    import sample_dump as sd

    reqs = sd.requests()
    results = []
    for px, py, s, seed, dim, bounce in reqs:
        i = abi.SampleStreamIndex(px, py, s, seed)
        bits = oracle.lib.bfo_indexed_bits(C.byref(i), dim)
        sample = oracle.lib.bfo_sample_1d(abi.SAMPLER_INDEXED_HASH, C.byref(i), dim)
        dims = (C.c_uint64 * 10)()
        status = oracle.lib.bfo_bounce_dimensions(bounce, dims)
        results.append((status, bits, sd.float_bits(sample), list(dims)))
    text = sd.canonical_text(sd.manifest_values(), reqs, results)
    assert sd.fnv1a64(text) == sd.FROZEN_CANONICAL_DUMP_FNV1A, text
    # the known bits and samples of the first four requests (:455-474)
    assert [r[1] for r in results[:4]] == [0x2F37B19CDAB08AA0, 0xDE27E521323C8AFF, 0xDC4661423B6A00FA, 0x8F4FD2AABC4AC602]
    assert [r[2] for r in results[:4]] == [0x3E3CDEC4, 0x3F5E27E5, 0x3F5C4661, 0x3F0F4FD2]
    assert results[4][0] == 1 and results[4][3] == [0] * 10
    # a planted off-by-one in the bounce map changes the fingerprint
    planted = [(st, b, f, [d + 1 if d else 0 for d in dims]) for st, b, f, dims in results]
    assert sd.fnv1a64(sd.canonical_text(sd.manifest_values(), reqs, planted)) != sd.FROZEN_CANONICAL_DUMP_FNV1A
