"""HIP closest-hit / any-hit traversal is bit-identical to the oracle's restatement of
CudaSceneClosestHit.cu on the same rays (status, t, barycentrics, normal, ids, and the canonical
node/triangle/instance visit counts)."""
import numpy as np
import pytest

from rays import random_rays, result_bytes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scene_name,n,seed,grazing", [("cornell", 20000, 1, 0.0), ("cornell", 20000, 2, 0.15),
                                                        ("grid", 20000, 3, 0.0)])
def test_closest_hit_matches_oracle(bf, abi, oracle, hip, scene_name, n, seed, grazing):
    if scene_name == "cornell":
        scene, _ = bf.scenes.cornell()
        center, radius = (0.0, 0.0, 1.0), 2.5
    else:
        scene, _ = bf.scenes.icosphere_grid(grid=4, subdivisions=2)
        center, radius = (0.0, 0.0, 0.0), 8.0
    rays = random_rays(abi, n, seed, center=center, radius=radius, grazing=grazing)
    expected, expected_ctr = oracle.scene(scene).closest_hit(rays, counters=True)
    xs = hip.XpuScene(scene)
    got, got_ctr = xs.closest_hit(rays, counters=True)
    statuses = np.array([r.status for r in expected])
    assert np.count_nonzero(statuses == 1) > n // 10
    assert result_bytes(got) == result_bytes(expected)
    assert np.array_equal(got_ctr, expected_ctr)


def test_occlusion_matches_oracle(bf, abi, oracle, hip):
    scene, _ = bf.scenes.cornell()
    rays = random_rays(abi, 20000, 7, center=(0.0, 0.0, 1.0), radius=2.5, t_max=1.5)
    expected = oracle.scene(scene).occlusion(rays)
    got = hip.XpuScene(scene).occlusion(rays)
    assert result_bytes(got) == result_bytes(expected)
    statuses = np.array([r.status for r in expected])
    assert 0 < np.count_nonzero(statuses == 1) < len(statuses)


def test_extension_enumerates_the_gpu(hip):
    name, devices = hip.xpu_devices()
    assert name.startswith("Blackframe HIP") and len(devices) >= 1
    kind, device_name, vendor, identifier = devices[0]
    assert kind == 2 and vendor == "AMD" and "gfx950" in device_name and identifier != (0, 0)
