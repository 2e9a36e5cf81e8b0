"""The seeded multi-scene workloads, built on the device: the uneven batch of plain frames and the
16-scene texture fit. Shared by bench.py and the GPU tests so both run exactly the same frames."""
import time

from . import scenes


def batch_frames(hip, seed=0, count=32):
    """The uneven batch (scenes.batch_scenes): 32 seeded scenes of 10k-500k triangles, 128^2-1024^2,
    4-256 spp. Returns (host scenes, device scenes, frames)."""
    items = scenes.batch_scenes(count=count)
    hosts = [scene for scene, _, _ in items]
    xscenes = [hip.XpuScene(scene) for scene in hosts]
    frames = [scenes.frame_desc(cam, spp, seed=seed) for _, cam, spp in items]
    return hosts, xscenes, frames


def fit_frames(hip, seed=0, count=16, size=1024):
    """The texture-fit batch (scenes.fit_specs): per scene an icosphere grid of 5e4-2e5 triangles
    with one Lambert + anisotropic gold GGX material bound to its own size^2 RGB albedo and
    size^2 (alpha_x, alpha_y) device textures, a square 256^2-512^2 frame of 32 spp. The textures
    start at the seeded target values. Returns a list of dicts with keys spec, scene (XpuScene),
    frame, albedo, widths (float32 CUDA tensors read in place) and the build time."""
    import torch

    t0 = time.perf_counter()
    out = []
    for k, spec in enumerate(scenes.fit_specs(count=count)):
        albedo_target, widths_target = scenes.fit_textures(spec["seed"], size=size)
        albedo = torch.from_numpy(albedo_target).cuda().contiguous()
        widths = torch.from_numpy(widths_target).cuda().contiguous()
        res = spec["resolution"]
        scene, cam = scenes.icosphere_grid(res, res, grid=spec["grid"], subdivisions=3, seed=spec["seed"],
                                           materials="textured", albedo=(albedo.data_ptr(), size, size, 3),
                                           roughness=(widths.data_ptr(), size, size, 2))
        out.append(dict(spec=spec, host=scene, scene=hip.XpuScene(scene), frame=scenes.frame_desc(cam, 32, seed=seed + k),
                        albedo=albedo, widths=widths, triangles=scene.triangle_count))
    return out, time.perf_counter() - t0


def fit_host_scene(spec, albedo, widths):
    """The host twin of one fit scene (same geometry and material) with host texture copies, for
    the oracle."""
    res = spec["resolution"]
    return scenes.icosphere_grid(res, res, grid=spec["grid"], subdivisions=3, seed=spec["seed"], materials="textured",
                                 albedo=albedo, roughness=widths)
