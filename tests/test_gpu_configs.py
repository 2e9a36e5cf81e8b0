"""The benchmark configurations at their full workloads on the GPU, checked against the oracle.

  - Cornell 64x64 x 16 spp (65 536 paths), depth 8, balance MIS, roulette from depth 3,
    Sobol-Owen seed 0: every path result and the float film bit-identical to the oracle's whole
    frame, the compensated film within 1e-4 relative;
  - the uneven batch (32 seeded scenes, 10k-500k triangles, 128^2-1024^2, 4-256 spp, depth 16,
    power MIS) in one batched call with graph-replayed bounce groups: per frame, seeded path slots
    and the first failed path re-rendered one path at a time by the oracle, bit for bit, and the
    batch's ray count equal to the frames rendered one by one;
  - the 16-scene texture fit (50k-200k triangles, 1024^2 albedo and (alpha_x, alpha_y) textures,
    256^2-512^2 x 32 spp): seeded textured paths bit-identical to the oracle on host copies of the
    textures, then one optimisation step (taped forward, loss, adjoint, Adam) whose directional
    derivative over all albedo texels matches a central difference to 1e-2 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sampled_slots_match(oracle, scene, frame, options, rows, report, count, seed):
    """Re-renders `count` seeded path slots (and the first failed one) with the oracle, one path at
    a time, and compares them with the device rows (uint8 [n, 64] on the device)."""
    import torch

    n = frame.camera.width * frame.camera.height * frame.samples_per_pixel
    rng = np.random.default_rng(seed)
    slots = sorted(set(rng.integers(0, n, size=count).tolist()))
    if report is not None and report.failed_paths:
        slots.append(report.first_failure_slot)
    picked = rows[torch.tensor(slots, device=rows.device)].cpu().numpy()
    oscene = oracle.scene(scene)
    for k, slot in enumerate(slots):
        expected = oscene.render(frame, options, path_begin=slot, path_count=1, threads=1, want_film=False)
        want = np.frombuffer(bytes(memoryview(expected["paths"]).cast("B")), dtype=np.uint8)
        assert np.array_equal(picked[k], want), f"path slot {slot} differs"


def test_cornell_whole_frame(bf, abi, oracle, hip):
    scene, cam = bf.scenes.cornell(64, 64)
    frame = bf.scenes.frame_desc(cam, 16, seed=0)
    options = bf.scenes.transport_options(max_depth=8, mis=abi.MIS_BALANCE, rr_first_depth=3)
    ws = hip.XpuWavefrontTransportWorkspace(64 * 64 * 16)
    got = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene(scene), frame, options, want_paths=True)
    expected = oracle.scene(scene).render(frame, options, threads=8)
    assert got["report"].failed_paths == 0 and expected["report"].failed_paths == 0
    a = np.frombuffer(bytes(memoryview(got["paths"]).cast("B")), dtype=np.uint32)
    b = np.frombuffer(bytes(memoryview(expected["paths"]).cast("B")), dtype=np.uint32)
    assert np.array_equal(a, b)
    film = got["film"].host()
    assert np.array_equal(film["film"], expected["film"])
    rel = np.abs(film["film_compensated"] - expected["film_compensated"]) / np.maximum(np.abs(expected["film_compensated"]),
                                                                                      1e-30)
    assert rel.max() <= 1e-4
    assert got["report"].closest_hit_queries == expected["report"].closest_hit_queries
    assert got["report"].shadow_queries == expected["report"].shadow_queries


def test_uneven_batch_sampled_paths(bf, abi, oracle, hip):
    from blackframe_renderer_amd import workloads

    hosts, xscenes, frames = workloads.batch_frames(hip)
    options = bf.scenes.transport_options(max_depth=16, mis=abi.MIS_POWER)
    options.failure_policy = abi.FAILURES_COUNT
    options.launch = abi.LAUNCH_GRAPH | 4
    ws = hip.XpuWavefrontTransportWorkspace(1 << 27)
    got = hip.trace_xpu_wavefront_batch(ws, list(zip(xscenes, frames)), options, want_paths="device")
    assert got["report"].failed_paths < 200
    for k, (scene, frame, rows) in enumerate(zip(hosts, frames, got["paths"])):
        sampled_slots_match(oracle, scene, frame, options, rows.view(-1, 64), None, 6, 1000 + k)
    if got["report"].failed_paths:
        job, slot = got["report"].first_failure_job, got["report"].first_failure_slot
        sampled = got["paths"][job].view(-1, 64)
        expected = oracle.scene(hosts[job]).render(frames[job], options, path_begin=slot, path_count=1, threads=1,
                                                   want_film=False)
        want = np.frombuffer(bytes(memoryview(expected["paths"]).cast("B")), dtype=np.uint8)
        assert np.array_equal(sampled[slot].cpu().numpy(), want)
        assert expected["report"].failed_paths == 1 and expected["report"].first_failure_code == got["report"].first_failure_code
    # the same frames one by one give the same rays (per-path results do not depend on batching)
    rays = 0
    for x, f in zip(xscenes, frames):
        r = hip.trace_xpu_wavefront_transport(ws, x, f, options, synchronize=True)
        rays += r["report"].closest_hit_queries + r["report"].shadow_queries
    assert rays == got["report"].closest_hit_queries + got["report"].shadow_queries


def test_texture_fit_step(bf, abi, oracle, hip):
    import torch

    from blackframe_renderer_amd import workloads
    from blackframe_renderer_amd.autodiff import TextureFit

    built, _ = workloads.fit_frames(hip)
    options = bf.scenes.transport_options(max_depth=8, mis=abi.MIS_BALANCE)
    options.russian_roulette_enabled = 0
    options.failure_policy = abi.FAILURES_COUNT
    options.launch = abi.LAUNCH_GRAPH | 4
    paths = sum(b["frame"].camera.width * b["frame"].camera.height * b["frame"].samples_per_pixel for b in built)
    ws = hip.XpuWavefrontTransportWorkspace(min(paths, 1 << 27))
    # the targets, and seeded textured paths of four frames against the oracle on host texture copies
    rendered = hip.trace_xpu_wavefront_batch(ws, [(b["scene"], b["frame"]) for b in built], options, want_paths="device")
    assert rendered["report"].failed_paths < 100
    for k in (0, 5, 10, 15):
        b = built[k]
        host, _ = workloads.fit_host_scene(b["spec"], b["albedo"].cpu().numpy(), b["widths"].cpu().numpy())
        sampled_slots_match(oracle, host, b["frame"], options, rendered["paths"][k].view(-1, 64), None, 8, 2000 + k)
    targets = []
    for film in rendered["films"]:
        counts = film.counts.to(torch.float32).clamp_min(1.0).view(film.height, film.width, 1)
        targets.append((film.float_sum.view(film.height, film.width, 3) / counts).contiguous())
    with torch.no_grad():
        for b in built:
            b["albedo"].fill_(0.5)
            b["widths"].fill_(0.25)
    params = []
    for b in built:
        params += [(b["albedo"], 0.02, 0.0, 1.0), (b["widths"], 0.01, 0.01, 1.0)]
    fit = TextureFit([(b["scene"], b["frame"], t, [b["albedo"], b["widths"]]) for b, t in zip(built, targets)], options,
                     params, workspace=ws)
    fit.forward()
    first = fit.loss().clone()
    fit.adjoint()
    albedos = [p for p in fit.parameters if p["value"].shape[-1] == 3]
    gen = torch.Generator(device="cuda").manual_seed(13)
    directions = [torch.randn(p["value"].shape, generator=gen, device="cuda") for p in albedos]
    ad = sum(float((p["gradient"].double() * d.double()).sum()) for p, d in zip(albedos, directions))
    base = [p["value"].clone() for p in albedos]
    eps = 1e-2

    def loss_at(sign):
        with torch.no_grad():
            for p, b0, d in zip(albedos, base, directions):
                p["value"].copy_(b0 + sign * eps * d)
        fit.forward()
        return float(fit.loss().sum())

    fd = (loss_at(1.0) - loss_at(-1.0)) / (2 * eps)
    with torch.no_grad():
        for p, b0 in zip(albedos, base):
            p["value"].copy_(b0)
    assert abs(ad - fd) <= 1e-2 * abs(fd), (fd, ad)
    # a whole optimisation step, then the loss has fallen
    fit.step()
    later = fit.step()
    assert float(later.sum()) < float(first.sum())
