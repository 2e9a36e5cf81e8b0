"""The stand-alone traversal in the reference's launcher form (blackframe_hip_launch_scene_closest_hit_blob
/ _occlusion_blob: a "BLKFSOA4" scene blob and the "BLKFBVH1" BVH built from it, both in device
memory) against the library's traversal of the same scene (blackframe_hip_launch_scene_closest_hit /
_occlusion on the scene made from the blob): every result byte for byte, on a Cornell blob (per-lane
spectra, a mixture, a parent chain) and on the reference's rich test scene (three punctual lights, two
emissive instances, a three-level parent chain), with axis-snapped directions among the rays. Broken
layouts answer every ray with the reference's invalid_scene / invalid_bvh (wrong magic, a byte count
other than the header's, a BVH whose source hash names another scene), an invalid ray answers
invalid_ray, and null or short blobs are rejected before any launch."""
import struct

import numpy as np
import pytest

from rays import random_rays, result_bytes
from soa_scenes import cornell_description, soa
from test_scene_soa import rich_blob

pytestmark = pytest.mark.gpu


def scene_blob(bf, which):
    return soa.serialize(soa.frame_scene(cornell_description(bf)[0])) if which == "cornell" else rich_blob()


def root_box(scene):
    tlas = scene.bvh()["tlas_nodes"]
    lo = np.frombuffer(tlas[0:12], dtype=np.float32)
    hi = np.frombuffer(tlas[16:28], dtype=np.float32)
    return (lo + hi) / 2.0, float(np.max(hi - lo))


@pytest.mark.parametrize("which", ["cornell", "rich"])
@pytest.mark.parametrize("grazing", [0.0, 0.25])
def test_blob_traversal_equals_the_scene_traversal(bf, abi, hip, which, grazing):
    blob = scene_blob(bf, which)
    scene = hip.XpuScene.from_soa(blob)
    bvh = scene.bvh_blob()
    center, size = root_box(scene)
    rays = random_rays(abi, 4000, 17, center=tuple(center), radius=float(size), grazing=grazing)
    want = scene.closest_hit(rays)
    got = hip.blob_closest_hit(blob, bvh, rays)
    assert result_bytes(got) == result_bytes(want)
    assert sum(r.status == abi.HIT_HIT for r in got) > 200
    assert result_bytes(hip.blob_closest_hit(blob, bvh, rays, occlusion=True)) == result_bytes(scene.occlusion(rays))


def test_broken_layouts_answer_every_ray(bf, abi, hip):
    blob = scene_blob(bf, "cornell")
    scene = hip.XpuScene.from_soa(blob)
    bvh = scene.bvh_blob()
    center, size = root_box(scene)
    rays = random_rays(abi, 300, 3, center=tuple(center), radius=float(size))
    bad_magic = bytearray(blob)
    bad_magic[0] ^= 0xFF
    statuses = {r.status for r in hip.blob_closest_hit(bytes(bad_magic), bvh, rays)}
    assert statuses == {abi.HIT_INVALID_SCENE}
    statuses = {r.status for r in hip.blob_closest_hit(blob, bvh, rays, scene_size=len(blob) - 16)}
    assert statuses == {abi.HIT_INVALID_SCENE}
    other = bytearray(bvh)
    struct.pack_into("<Q", other, 40, struct.unpack_from("<Q", bvh, 40)[0] ^ 1)  # source_scene_hash
    statuses = {r.status for r in hip.blob_closest_hit(blob, bytes(other), rays, occlusion=True)}
    assert statuses == {abi.HIT_INVALID_BVH}
    rays[5].t_min = -1.0
    got = hip.blob_closest_hit(blob, bvh, rays)
    assert got[5].status == abi.HIT_INVALID_RAY
    with pytest.raises(hip.XpuError):
        hip.blob_closest_hit(blob, bvh[:100], rays)
