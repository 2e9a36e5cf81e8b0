"""The power-weighted light sampler (Renderer/Source/LightSampler.cpp) on the CPU: the library's CDF
builder (blackframe_hip_light_power_cdf, host code) and the oracle's restatement (CDF, sample,
probability), both pinned by the reference's own expectations in Tests/Renderer/Light/
LightSamplerTest.cpp:
  - four-lane powers (4,0,0,0), 0, (0,4,0,0), (0,0,8,0), (0,0,0,16): probabilities 1/8, 0, 1/8,
    1/4, 1/2 exactly, 4096 stratified samples counted 512/0/512/1024/2048 (:95-146);
  - boundary classification: 0 and prev(1/8) -> 0, 1/8 -> 2, 1/4 -> 3, 1/2 -> 4, prev(1) -> 4
    (:147-167); a single powered light among zeros takes all the probability (:169-190);
  - the indexed sample stream of pixel (37, 91), seed 0xD1B54A32D192ED03, dimension 4 (bounce 0's
    light selection), 65 536 samples within 8 sigma + 1 of the pdf (:192-228);
  - non-dyadic powers 1, 2, 3, 5, 7: stratified counts within 1 of p N, probabilities summing to
    1 within 8 ulp (:230-268);
  - scale invariance and range: (1,2,4) and (8,16,32) give identical probabilities; FLT_MAX and
    the smallest subnormal each give 1/2, 1/2 (:285-327);
  - rejections: no lights, black, negative / NaN / inf lanes (invalid argument); a tiny middle or
    terminal weight and unscalable lanes (resource exhausted) (:329-372).
The two builders agree bit for bit on seeded random powers (zeros, ties, wide ranges).
"""
import numpy as np
import pytest

F32 = np.float32
INVALID, EXHAUSTED = 1, 3  # bf_status


def lanes(a, b, c, d):
    return [a, b, c, d]


def filled(v):
    return [v] * 4


WEIGHTED = [lanes(4, 0, 0, 0), lanes(0, 0, 0, 0), lanes(0, 4, 0, 0), lanes(0, 0, 8, 0), lanes(0, 0, 0, 16)]


def product_cdf(hip, powers):
    return hip.light_power_cdf(np.array(powers, dtype=np.float32).reshape(-1, 4))


def product_status(hip, powers):
    try:
        product_cdf(hip, powers)
        return 0
    except hip.XpuError as e:
        return e.status


def test_exact_power_frequencies(oracle, hip):
    s = oracle.power_sampler(WEIGHTED)
    assert s.status == 0
    expected = [0.125, 0.0, 0.125, 0.25, 0.5]
    counts = [0] * 5
    for k in range(4096):
        u = (F32(k) + F32(0.5)) / F32(4096)
        index, probability = s.sample(u)
        counts[index] += 1
        assert probability == expected[index] and s.probability(index) == probability
    assert counts == [512, 0, 512, 1024, 2048]
    assert [s.probability(i) for i in range(5)] == expected and sum(expected) == 1.0
    assert np.array_equal(product_cdf(hip, WEIGHTED), s.cdf)


def test_boundary_classification_skips_zero_power_slots(oracle):
    s = oracle.power_sampler(WEIGHTED)
    prev = lambda x: np.nextafter(F32(x), F32(0))
    cases = [(0.0, 0), (prev(0.125), 0), (0.125, 2), (0.25, 3), (0.5, 4), (prev(1.0), 4)]
    for u, index in cases:
        assert s.sample(u)[0] == index, u
    assert s.probability(1) == 0.0


@pytest.mark.parametrize("supported", [0, 1, 2])
def test_single_supported_light_takes_all(oracle, hip, supported):
    powers = [filled(0.0)] * 3
    powers[supported] = filled(1.0)
    s = oracle.power_sampler(powers)
    assert s.sample(0.0)[0] == supported and s.sample(np.nextafter(F32(1), F32(0)))[0] == supported
    assert [s.probability(i) for i in range(3)] == [1.0 if i == supported else 0.0 for i in range(3)]
    assert np.array_equal(product_cdf(hip, powers), s.cdf)


def test_indexed_stream_frequencies_match_the_pdf(oracle, abi):
    import ctypes as C

    s = oracle.power_sampler(WEIGHTED)
    counts = np.zeros(5, dtype=np.int64)
    n = 65536
    for k in range(n):
        index = abi.SampleStreamIndex(37, 91, k, 0xD1B54A32D192ED03)
        u = oracle.lib.bfo_sample_1d(0, C.byref(index), 4)  # indexed hash, bounce 0 light selection
        counts[s.sample(u)[0]] += 1
    for count, p in zip(counts, [0.125, 0.0, 0.125, 0.25, 0.5]):
        sigma = np.sqrt(n * p * (1.0 - p))
        assert abs(count - n * p) <= 8.0 * sigma + 1.0


def test_non_dyadic_frequencies_match_the_stored_pdf(oracle, hip):
    powers = [filled(v) for v in (1.0, 2.0, 3.0, 5.0, 7.0)]
    s = oracle.power_sampler(powers)
    n = 65536
    counts = np.zeros(5, dtype=np.int64)
    for k in range(n):
        counts[s.sample((F32(k) + F32(0.5)) / F32(n))[0]] += 1
    probabilities = np.array([s.probability(i) for i in range(5)], dtype=np.float64)
    assert np.all(np.abs(counts - probabilities * n) <= 1.0)
    assert abs(probabilities.sum() - 1.0) <= 8.0 * np.finfo(np.float32).eps
    assert np.array_equal(product_cdf(hip, powers), s.cdf)


def test_scale_invariance_and_range(oracle, hip):
    base = oracle.power_sampler([filled(v) for v in (1.0, 2.0, 4.0)])
    scaled = oracle.power_sampler([filled(v) for v in (8.0, 16.0, 32.0)])
    assert [base.probability(i) for i in range(3)] == [scaled.probability(i) for i in range(3)]
    big = np.finfo(np.float32).max
    tiny = np.float32(1e-45)  # the smallest subnormal
    for v in (big, tiny):
        s = oracle.power_sampler([filled(v), filled(v)])
        assert s.status == 0 and s.probability(0) == 0.5 and s.probability(1) == 0.5
        assert np.array_equal(product_cdf(hip, [filled(v), filled(v)]), s.cdf)


def test_rejections_carry_the_reference_status(oracle, hip):
    eps = np.finfo(np.float32).eps
    cases = [([], INVALID), ([filled(0.0), filled(0.0)], INVALID)]
    for bad in (-1.0, np.nan, np.inf):
        p = [filled(1.0), filled(2.0)]
        p[1] = [2.0, 2.0, bad, 2.0]
        cases.append((p, INVALID))
    tiny = float(F32(eps) * F32(eps))
    cases += [([filled(1.0), filled(tiny), filled(1.0)], EXHAUSTED), ([filled(1.0), filled(tiny)], EXHAUSTED),
              ([lanes(float(np.finfo(np.float32).max), 0, 0, 0), lanes(1e-45, 0, 0, 0)], EXHAUSTED)]
    for powers, status in cases:
        if powers:
            assert oracle.power_sampler(powers).status == status, powers
        assert product_status(hip, powers) == status, powers


def test_product_and_oracle_builders_agree(oracle, hip):
    rng = np.random.default_rng(8)
    for trial in range(200):
        n = int(rng.integers(1, 40))
        powers = rng.lognormal(0.0, 4.0, size=(n, 4)).astype(np.float32)
        powers[rng.random((n, 4)) < 0.3] = 0.0
        if trial % 5 == 0:
            powers[rng.integers(0, n)] = powers[0]  # ties
        if not np.any(powers > 0):
            powers[0, 0] = 1.0
        s = oracle.power_sampler(powers)
        status = product_status(hip, powers)
        assert status == s.status
        if s.status == 0:
            assert np.array_equal(product_cdf(hip, powers), s.cdf)
