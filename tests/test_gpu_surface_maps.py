"""Normal and bump maps on the GPU (EWA-filtered image mip chains over each hit's ray-cone
footprint, CudaWavefrontStages.cu:2876-3673 and the shade stage's use of them at :4170-4193)
against the oracle, through the path-input transport whose rays carry cones:

  - a Cornell box whose white walls carry a normal map (repeat wrap) and whose right wall carries a
    bump map (mirror wrap), 24x24x4 paths with cones of spread 1.5e-3, roulette on: every terminal
    record bit-identical to the oracle's, and the maps change the radiance of most paths;
  - the same with a y-flipped normal map, clamp / border wraps and an anisotropy of 2;
  - a texel budget too small for the footprints fails those paths (code 49) as the oracle does;
  - the maps delivered in the reference's scene blob render like the oracle's Scene-builder twin;
  - the frame entry point, whose camera starts each path with its one-pixel cone, renders the
    mapped Cornell box bit-identically to the oracle's frame render.
"""
import numpy as np
import pytest

from path_inputs import as_u32, frame_inputs
from test_gpu_paths import compare

pytestmark = pytest.mark.gpu


def mapped_cornell(bf, abi, y_convention=0, wraps=(0, 2), anisotropy=8, visits=4096):
    scene, cam = bf.scenes.cornell(24, 24)
    n = 64
    y, x = np.mgrid[0:n, 0:n]
    u, v = (x + 0.5) / n, (y + 0.5) / n
    tilt_x = 0.35 * np.sin(2 * np.pi * 3 * u) * np.cos(2 * np.pi * 2 * v)
    tilt_y = 0.35 * np.cos(2 * np.pi * 4 * v)
    normal = np.stack([tilt_x, tilt_y, np.sqrt(1 - tilt_x ** 2 - tilt_y ** 2)], axis=-1)
    height = 0.5 + 0.5 * np.sin(2 * np.pi * 5 * u) * np.sin(2 * np.pi * 3 * v)
    t_normal = scene.add_image_texture(bf.scenes.mip_chain(0.5 * (normal + 1.0)))
    t_height = scene.add_image_texture(bf.scenes.mip_chain(height[..., None]))
    scene.set_surface_maps(0, normal={"texture": t_normal, "u_wrap": wraps[0], "v_wrap": wraps[0],
                                      "max_anisotropy": anisotropy, "max_texel_visits": visits,
                                      "y_convention": y_convention})
    scene.set_surface_maps(2, bump={"texture": t_height, "scale": 0.05, "u_wrap": wraps[1], "v_wrap": wraps[1],
                                    "max_anisotropy": anisotropy, "max_texel_visits": visits})
    return scene.finalize(), cam


@pytest.mark.parametrize("y_convention,wraps,anisotropy", [(0, (0, 2), 8), (1, (1, 3), 2)])
def test_mapped_cornell_matches_the_oracle(bf, abi, oracle, hip, y_convention, wraps, anisotropy):
    scene, cam = mapped_cornell(bf, abi, y_convention, wraps, anisotropy)
    frame = bf.scenes.frame_desc(cam, 4, seed=5)
    options = bf.scenes.transport_options(max_depth=6, rr_first_depth=2)
    inputs = frame_inputs(abi, oracle, frame, options, cone=(0.0, 1.5e-3))
    got, report = compare(hip, oracle, scene, inputs, options)
    assert report.failed_paths <= len(inputs[1]) // 100
    plain, _ = bf.scenes.cornell(24, 24)
    ws = hip.XpuWavefrontTransportWorkspace(len(inputs[1]))
    flat = hip.trace_xpu_wavefront_paths(ws, hip.XpuScene(plain), inputs, options)
    changed = (as_u32(got["states"])[:, 4:8] != as_u32(flat["states"])[:, 4:8]).any(axis=1)
    assert changed.mean() > 0.5


def test_texel_budget_fails_paths_like_the_oracle(bf, abi, oracle, hip):
    scene, cam = mapped_cornell(bf, abi, visits=2)
    frame = bf.scenes.frame_desc(cam, 2, seed=1)
    options = bf.scenes.transport_options(max_depth=4)
    options.failure_policy = abi.FAILURES_COUNT
    inputs = frame_inputs(abi, oracle, frame, options, cone=(0.0, 4e-3))
    got, report = compare(hip, oracle, scene, inputs, options)
    assert report.failed_paths > 0 and report.first_failure_code == 49


def test_blob_maps_match_the_oracle(bf, abi, oracle, hip):
    """The same maps arriving in the reference's scene blob (image texture, mip and texel columns;
    material normal- and bump-map columns): the blob scene's path-input render on the device equals
    the oracle's render of its Scene-builder twin, record for record."""
    from soa_scenes import cornell_description, scene_from_canonical, soa
    from test_gpu_scene_soa import camera_frame, packet_options

    d, wl = cornell_description(bf)
    n = 32
    y, x = np.mgrid[0:n, 0:n]
    u, v = (x + 0.5) / n, (y + 0.5) / n
    tilt = 0.3 * np.sin(2 * np.pi * 2 * u)
    normal = np.stack([tilt, 0.2 * np.cos(2 * np.pi * 3 * v), np.ones_like(u)], axis=-1)
    normal /= np.linalg.norm(normal, axis=-1, keepdims=True)
    height = np.cos(2 * np.pi * 4 * u) * np.sin(2 * np.pi * 2 * v)

    def image(tid, pixels):
        return {"id": tid, "storage_color_space": 0,
                "levels": [{"width": l.shape[1], "height": l.shape[0], "channels": l.shape[2],
                            "pixels": [float(e) for e in l.reshape(-1)]} for l in bf.scenes.mip_chain(pixels)]}

    d["image_textures"] = [image(700, 0.5 * (normal + 1.0)), image(701, height[..., None])]
    by_id = {m["id"]: m for m in d["materials"]}
    by_id[21]["spectral"]["normal_map"] = {"texture": 700, "red_channel": 0, "green_channel": 1, "blue_channel": 2,
                                           "u_wrap": 0, "v_wrap": 2, "maximum_anisotropy": 8,
                                           "maximum_texel_visits": 4096, "y_convention": 1}
    by_id[23]["spectral"]["bump_map"] = {"texture": 701, "channel": 0, "scale": 0.04, "u_wrap": 1, "v_wrap": 0,
                                         "maximum_anisotropy": 4, "maximum_texel_visits": 4096}
    canon = soa.frame_scene(d)
    blob = soa.serialize(canon)
    twin = scene_from_canonical(bf, canon)
    frame = camera_frame(bf, spp=2, seed=3)
    options = packet_options(bf, wl, max_depth=5)
    options.record_queues = 0
    inputs = frame_inputs(abi, oracle, frame, options, cone=(0.0, 1.5e-3))
    ws = hip.XpuWavefrontTransportWorkspace(len(inputs[1]))
    got = hip.trace_xpu_wavefront_paths(ws, hip.XpuScene.from_soa(blob), inputs, options)
    states, rays, cones, controls, report = oracle.scene(twin).render_paths(inputs, options, threads=8)
    assert got["report"].failed_paths == report.failed_paths
    for name, want in (("states", states), ("rays", rays), ("cones", cones), ("controls", controls)):
        assert np.array_equal(as_u32(got[name]), as_u32(want)), name
    flat = oracle.scene(scene_from_canonical(bf, soa.frame_scene(cornell_description(bf)[0]))).render_paths(
        inputs, options, threads=8)
    assert (as_u32(states)[:, 4:8] != as_u32(flat[0])[:, 4:8]).any(axis=1).mean() > 0.3


@pytest.mark.parametrize("y_convention,wraps,anisotropy", [(0, (0, 2), 8), (1, (1, 3), 2)])
def test_frame_entry_renders_mapped_scenes_like_the_oracle(bf, abi, oracle, hip, y_convention, wraps, anisotropy):
    """Camera frames start each path with the camera's one-pixel cone (PinholeCamera.hpp:224-252),
    so the frame entry point renders mapped scenes: every path record and the float film
    bit-identical to the oracle's frame render, and the maps change most paths."""
    from test_gpu_transport import path_arrays

    scene, cam = mapped_cornell(bf, abi, y_convention, wraps, anisotropy)
    frame = bf.scenes.frame_desc(cam, 4, seed=7)
    options = bf.scenes.transport_options(max_depth=6, rr_first_depth=2)
    options.failure_policy = abi.FAILURES_COUNT
    want = oracle.scene(scene).render(frame, options, threads=8)
    ws = hip.XpuWavefrontTransportWorkspace(24 * 24 * 4)
    got = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene(scene), frame, options, want_paths=True)
    er, gr = want["report"], got["report"]
    assert gr.failed_paths == er.failed_paths <= 24 * 24 * 4 // 100
    assert np.array_equal(path_arrays(got["paths"]), path_arrays(want["paths"]))
    film = got["film"].host()
    assert np.array_equal(film["film"].view(np.uint32), np.asarray(want["film"]).view(np.uint32))
    plain, _ = bf.scenes.cornell(24, 24)
    flat = oracle.scene(plain).render(frame, options, threads=8)
    changed = (path_arrays(want["paths"])[:, :4] != path_arrays(flat["paths"])[:, :4]).any(axis=1)
    assert changed.mean() > 0.5
