"""The device sample-stream dumps.

  - blackframe_hip_launch_sample_stream_dump (the reference's contract): the canonical text of the
    reference's requests computed from the device results and manifest hashes to the reference's
    frozen FNV-1a-64 (Tests/CUDA/SampleStreamTest.cpp:35,401-414); replaying the requests and
    reversing their order gives the same results (:484-499); bad schemas, null pointers and empty
    requests are rejected before any launch (:416-450, CPU test below);
  - blackframe_hip_launch_sample_dump: both samplers (indexed hash, Sobol-Owen) equal the oracle's
    sample_1d bit for bit on seeded indices and dimensions, including the raster dimensions.
"""
import ctypes as C

import numpy as np
import pytest

import sample_dump as sd


def request_array(abi, reqs):
    arr = (abi.SampleDumpRequest * len(reqs))()
    for k, (px, py, s, seed, dim, bounce) in enumerate(reqs):
        arr[k].sample_stream = abi.SampleStreamIndex(px, py, s, seed)
        arr[k].dimension, arr[k].bounce_index = dim, bounce
    return arr


def as_tuples(results):
    b = lambda d: [d.light_selection, d.light_u, d.light_v, d.bsdf_component, d.bsdf_u, d.bsdf_v, d.medium_distance,
                   d.medium_phase_u, d.medium_phase_v, d.russian_roulette]
    return [(r.status, r.indexed_bits, sd.float_bits(r.sample), b(r.bounce_dimensions)) for r in results]


def manifest_tuple(m):
    p, o = m.primary_dimensions, m.bounce_offsets
    return (m.schema_version, (p.camera_raster_x, p.camera_raster_y, p.lens_u, p.lens_v, p.time, p.wavelength),
            m.first_bounce_dimension, m.dimensions_per_bounce, m.maximum_bounce_index,
            (o.light_selection, o.light_u, o.light_v, o.bsdf_component, o.bsdf_u, o.bsdf_v, o.medium_distance,
             o.medium_phase_u, o.medium_phase_v, o.russian_roulette))


def test_dump_rejects_bad_arguments_before_any_launch(hip, abi):
    lib = hip.library()
    fake = C.c_void_p(0x10000)  # never dereferenced: every case is rejected on the host
    call = lib.blackframe_hip_launch_sample_stream_dump
    for schema in (0, 2):
        assert call(schema, fake, 11, fake, fake) == 1
    assert call(1, None, 11, fake, fake) == 1
    assert call(1, fake, 11, None, fake) == 1
    assert call(1, fake, 11, fake, None) == 1
    assert call(1, fake, 0, fake, fake) == 1


@pytest.mark.gpu
def test_device_dump_matches_the_frozen_canonical_fingerprint(hip, abi):
    reqs = sd.requests()
    manifest, results = hip.sample_stream_dump_requests(request_array(abi, reqs))
    m = manifest_tuple(manifest)
    assert m == sd.manifest_values() and manifest.reserved == 0
    assert all(r.reserved == 0 and r.reserved_tail == 0 for r in results)
    text = sd.canonical_text(m, reqs, as_tuples(results))
    assert sd.fnv1a64(text) == sd.FROZEN_CANONICAL_DUMP_FNV1A, text
    # replay and reversed request order
    _, again = hip.sample_stream_dump_requests(request_array(abi, reqs))
    _, reversed_results = hip.sample_stream_dump_requests(request_array(abi, reqs[::-1]))
    assert as_tuples(again) == as_tuples(results)
    assert as_tuples(reversed_results)[::-1] == as_tuples(results)


@pytest.mark.gpu
@pytest.mark.parametrize("sampler", [0, 1])
def test_device_samples_equal_the_oracle(hip, abi, oracle, sampler):
    rng = np.random.default_rng(17 + sampler)
    n = 4096
    indices = (abi.SampleStreamIndex * n)()
    dims = np.empty(n, dtype=np.uint64)
    for k in range(n):
        indices[k] = abi.SampleStreamIndex(int(rng.integers(0, 4096)), int(rng.integers(0, 4096)),
                                           int(rng.integers(0, 1 << 20)), int(rng.integers(0, 1 << 63)))
        dims[k] = [sd.RASTER_X, sd.RASTER_Y, 3, int(rng.integers(0, 700))][k % 4]
    got = hip.sample_stream_dump(sampler, indices, dims)
    want = np.array([oracle.lib.bfo_sample_1d(sampler, C.byref(indices[k]), int(dims[k])) for k in range(n)],
                    dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
