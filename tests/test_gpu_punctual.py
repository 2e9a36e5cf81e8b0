"""Punctual lights on the GPU, bit-exact against the oracle: the microfacet Cornell box with a
point, a spot and a directional light beside its area light, and the reference's rich test scene
(three punctual lights, two emissive instances, a parent chain) ingested from its blob."""
import numpy as np
import pytest

from soa_scenes import soa, scene_from_canonical
from test_gpu_transport import check_parity, run_both
from test_scene_soa import rich_blob, rich_description

pytestmark = pytest.mark.gpu


def unit32(v):
    v = np.asarray(v, dtype=np.float64)
    return tuple(float(x) for x in (v / np.linalg.norm(v)).astype(np.float32))


def lit_cornell(bf):
    def factory(w, h):
        s, cam = bf.scenes.cornell(w, h, materials="mixed")
        s.add_point_light((0.3, 0.6, 1.2), 1.5, (1e-6, 1e-6, 1e-6))
        s.add_spot_light((-0.5, 0.9, 1.6), unit32((0.3, -1.0, -0.2)), 0.3, 0.6, 4.0, (1e-6, 1e-6, 1e-6))
        s.add_directional_light(unit32((0.1, -0.3, -1.0)), 0.8)
        s.finalize()
        return s, cam
    return factory


@pytest.mark.parametrize("opts", [dict(), dict(mis=1, max_depth=12, sampler=0)])
def test_cornell_with_punctual_lights_bit_exact(bf, oracle, hip, opts):
    print(f"punctual Cornell {opts}", flush=True)
    check_parity(*run_both(bf, oracle, hip, 32, 32, 4, scene_factory=lit_cornell(bf), **dict(opts)))


def test_rich_reference_scene_from_its_blob(bf, oracle, hip):
    """make_rich_scene (SceneSoATest.cpp:202-333), whose blob hash is frozen, rendered from the
    blob on the GPU and from its Scene-builder twin by the oracle."""
    canon = soa.frame_scene(rich_description())
    blob = rich_blob()
    wl = rich_description()["environment"]["wavelengths"]
    twin = scene_from_canonical(bf, canon)
    cam = bf.scenes.look_at_camera((3.0, 2.5, 12.0), (3.0, 2.5, 2.0), 32, 32, 0.9)
    frame = bf.scenes.frame_desc(cam, 4, seed=3)
    options = bf.scenes.transport_options(wavelength_mode=0, packet=[float(w[0]) for w in wl], mis=1)
    options.record_queues = 1
    print("rich scene from its blob", flush=True)
    expected = oracle.scene(twin).render(frame, options, threads=8)
    ws = hip.XpuWavefrontTransportWorkspace(32 * 32 * 4)
    got = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene.from_soa(blob), frame, options, want_paths=True,
                                            want_counters=True)
    check_parity(expected, got)
    assert np.count_nonzero(got["film"].host()["film"]) > 100
