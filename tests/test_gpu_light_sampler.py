"""Power-weighted light selection in the transport on the GPU, against the oracle with the same
power CDF bound to its scene:

  - the microfacet Cornell box with a point, a spot and a directional light beside its area light,
    uneven powers: paths, queues and film bit-exact, and the film differs from uniform selection's;
  - the 32x32 icosphere grid's 64 emissive spheres with seeded powers at 48x27 x 2 spp: bit-exact;
  - unbiasedness: the power-selected and the uniform film of the same scene agree in mean within
    their Monte Carlo error at 64 spp;
  - an emitter given zero power fails the paths whose BSDF samples hit it, on both sides alike
    (the reference's scalar MIS refuses a light sampler without support for a hit emitter).
"""
import numpy as np
import pytest

from test_gpu_punctual import lit_cornell
from test_gpu_transport import check_parity

pytestmark = pytest.mark.gpu

CORNELL_POWERS = [[4.0, 4.0, 4.0, 4.0], [1.0, 2.0, 3.0, 4.0], [0.5, 0.5, 0.5, 0.5], [16.0, 12.0, 8.0, 16.0]]


def render_both(bf, oracle, hip, scene, cam, spp, powers, seed=0, failures=False, **opts):
    frame = bf.scenes.frame_desc(cam, spp, seed=seed)
    options = bf.scenes.transport_options(**opts)
    options.record_queues = 1
    if failures:
        options.failure_policy = bf.abi.FAILURES_COUNT
    expected = oracle.scene(scene).set_light_powers(powers).render(frame, options, threads=8)
    xs = hip.XpuScene(scene).set_light_powers(powers)
    ws = hip.XpuWavefrontTransportWorkspace(cam.width * cam.height * spp)
    got = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, want_paths=True, want_counters=True)
    return expected, got


def test_cornell_power_selection_bit_exact(bf, oracle, hip):
    scene, cam = lit_cornell(bf)(32, 32)
    expected, got = render_both(bf, oracle, hip, scene, cam, 4, CORNELL_POWERS)
    check_parity(expected, got)
    uniform = oracle.scene(scene).render(bf.scenes.frame_desc(cam, 4), bf.scenes.transport_options(), threads=8)
    assert not np.array_equal(uniform["film"], expected["film"])


def test_icosphere_grid_power_selection_bit_exact(bf, oracle, hip):
    scene, cam = bf.scenes.icosphere_grid(48, 27)
    lights = oracle.scene(scene).light_count()
    assert lights == 64
    powers = np.random.default_rng(64).uniform(0.1, 10.0, size=(lights, 4)).astype(np.float32)
    expected, got = render_both(bf, oracle, hip, scene, cam, 2, powers, mis=bf.abi.MIS_POWER, max_depth=6,
                                failures=True)
    assert got["report"].failed_paths == expected["report"].failed_paths
    if expected["report"].failed_paths == 0:
        check_parity(expected, got)
    else:
        a = np.frombuffer(bytes(memoryview(got["paths"]).cast("B")), dtype=np.uint32)
        b = np.frombuffer(bytes(memoryview(expected["paths"]).cast("B")), dtype=np.uint32)
        assert np.array_equal(a, b)


def test_power_selection_is_unbiased(bf, hip):
    scene, cam = lit_cornell(bf)(24, 24)
    frame = bf.scenes.frame_desc(cam, 64, seed=5)
    options = bf.scenes.transport_options()
    ws = hip.XpuWavefrontTransportWorkspace(24 * 24 * 64)
    means = []
    for powers in (None, CORNELL_POWERS):
        xs = hip.XpuScene(scene).set_light_powers(powers)
        got = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, want_paths=True)
        radiance = np.frombuffer(bytes(memoryview(got["paths"]).cast("B")), dtype=np.float32).reshape(-1, 16)[:, 0:4]
        means.append((radiance.mean(), radiance.std() / np.sqrt(len(radiance))))
    (m0, e0), (m1, e1) = means
    assert abs(m0 - m1) <= 5.0 * np.hypot(e0, e1), means


def test_zero_power_emitter_fails_its_bsdf_hits_alike(bf, oracle, hip):
    scene, cam = lit_cornell(bf)(16, 16)
    powers = [row[:] for row in CORNELL_POWERS]
    powers[3] = [0.0, 0.0, 0.0, 0.0]  # the area light: never selected, yet hit by BSDF samples
    expected, got = render_both(bf, oracle, hip, scene, cam, 4, powers, failures=True)
    assert expected["report"].failed_paths > 0
    assert got["report"].failed_paths == expected["report"].failed_paths
    assert got["report"].first_failure_slot == expected["report"].first_failure_slot
    assert got["report"].first_failure_code == expected["report"].first_failure_code
