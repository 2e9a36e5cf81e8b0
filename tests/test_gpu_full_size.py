"""Parity at the full plain-forward size (1920x1080 x 64 spp, 1.31 M triangles, mixed GGX /
Lambert materials, depth 16, power MIS): the whole frame is rendered on the GPU with per-path
results kept in device memory; a seeded sample of path slots, plus the first numerically failed
path, is re-rendered one path at a time by the CPU restatement and must match bit for bit (paths
are independent, so a path's result does not depend on which other paths share the launch)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_full_frame_sampled_paths_match_oracle(bf, abi, oracle, hip):
    import torch

    scene, cam = bf.scenes.icosphere_grid()
    frame = bf.scenes.frame_desc(cam, 64, seed=0)
    options = bf.scenes.transport_options(max_depth=16, mis=abi.MIS_POWER)
    options.failure_policy = abi.FAILURES_COUNT
    n = cam.width * cam.height * 64
    ws = hip.XpuWavefrontTransportWorkspace(n)
    got = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene(scene), frame, options, want_paths="device")
    report = got["report"]
    assert report.path_count == n
    assert report.failed_paths < 100  # rare not-representable events of the reference lobe arithmetic
    rows = got["paths"].view(-1, 64)
    rng = np.random.default_rng(20261202)
    slots = sorted(set(rng.integers(0, n, size=192).tolist()))
    if report.failed_paths:
        slots.append(report.first_failure_slot)
    picked = rows[torch.tensor(slots, device=rows.device)].cpu().numpy()
    oscene = oracle.scene(scene)
    for k, slot in enumerate(slots):
        expected = oscene.render(frame, options, path_begin=slot, path_count=1, threads=1, want_film=False)
        want = np.frombuffer(bytes(memoryview(expected["paths"]).cast("B")), dtype=np.uint8)
        assert np.array_equal(picked[k], want), f"path slot {slot} differs"
    if report.failed_paths:
        expected = oscene.render(frame, options, path_begin=report.first_failure_slot, path_count=1, threads=1,
                                 want_film=False)["report"]
        assert expected.failed_paths == 1 and expected.first_failure_code == report.first_failure_code
