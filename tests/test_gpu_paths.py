"""The path-input transport on the GPU (blackframe_hip_trace_wavefront_paths: per path a sample
stream, primary ray, ray cone and initial path state in; terminal path state, ray, cone and lane
control out) against the oracle's path-input transport, record for record and bit for bit:

  - Cornell boxes (Lambert and mixed closures, roulette on) with inputs generated from a frame, at
    24x24x4 and at the full 64x64x16 benchmark frame (65 536 paths);
  - ray cones through Lambert bounces and through a mirror sphere (spreads kept at mirrors);
  - a depth limit of 1 (the blocked depth categories in the lane controls);
  - invalid inputs, which fail only their paths (zero records, termination NONE) with the same
    failure count and first failure as the oracle;
  - the path-input transport gives the frame transport's per-path results on the device.
"""
import numpy as np
import pytest

from path_inputs import as_u32, frame_inputs

pytestmark = pytest.mark.gpu


def compare(hip, oracle, scene, inputs, options, threads=8):
    ws = hip.XpuWavefrontTransportWorkspace(len(inputs[1]))
    got = hip.trace_xpu_wavefront_paths(ws, hip.XpuScene(scene), inputs, options)
    states, rays, cones, controls, report = oracle.scene(scene).render_paths(inputs, options, threads=threads)
    r = got["report"]
    assert (r.failed_paths, r.closest_hit_queries, r.shadow_queries) == \
        (report.failed_paths, report.closest_hit_queries, report.shadow_queries)
    if report.failed_paths:
        assert (r.first_failure_slot, r.first_failure_code) == (report.first_failure_slot, report.first_failure_code)
    for name, want in (("states", states), ("rays", rays), ("cones", cones), ("controls", controls)):
        a, b = as_u32(got[name]), as_u32(want)
        bad = np.nonzero((a != b).any(axis=1))[0]
        assert bad.size == 0, f"{name} differ at paths {bad[:8].tolist()} ({bad.size} in all)"
    return got, report


@pytest.mark.parametrize("materials,rr", [("lambert", 3), ("mixed", 2)])
def test_cornell_path_inputs_match_the_oracle(bf, abi, oracle, hip, materials, rr):
    scene, cam = bf.scenes.cornell(24, 24, materials=materials)
    frame = bf.scenes.frame_desc(cam, 4, seed=9)
    options = bf.scenes.transport_options(max_depth=8, rr_first_depth=rr)
    inputs = frame_inputs(abi, oracle, frame, options, cone=(0.0, 1.5e-3))
    got, report = compare(hip, oracle, scene, inputs, options)
    assert report.failed_paths == 0
    ctl = as_u32(got["controls"])
    assert np.all(ctl[:, 0] == abi.PHASE_TERMINATED)
    assert (ctl[:, 1] == abi.TERM_RUSSIAN_ROULETTE).sum() > 10


def test_full_frame_path_inputs_match_the_oracle_and_the_frame_transport(bf, abi, oracle, hip):
    scene, cam = bf.scenes.cornell(64, 64)
    frame = bf.scenes.frame_desc(cam, 16, seed=0)
    options = bf.scenes.transport_options(max_depth=8, mis=abi.MIS_BALANCE, rr_first_depth=3)
    inputs = frame_inputs(abi, oracle, frame, options)
    got, _ = compare(hip, oracle, scene, inputs, options)
    # the frame entry point's per-path results on the same device: one algorithm behind both
    ws = hip.XpuWavefrontTransportWorkspace(64 * 64 * 16)
    framed = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene(scene), frame, options, want_paths=True)
    want = as_u32(framed["paths"])
    st = as_u32(got["states"])
    assert np.array_equal(st[:, 4:8], want[:, 0:4]) and np.array_equal(st[:, 0:4], want[:, 4:8])
    assert np.array_equal(st[:, 8:12], want[:, 8:12]) and np.array_equal(st[:, 21], want[:, 12])
    assert np.array_equal(as_u32(got["controls"])[:, 1], want[:, 13])
    assert got["report"].closest_hit_queries == framed["report"].closest_hit_queries


def test_ray_cones_through_lambert_and_mirror_bounces(bf, abi, oracle, hip):
    from test_oracle_specular import SPEC_R, spec

    scene, cam = bf.scenes.cornell(16, 16)
    frame = bf.scenes.frame_desc(cam, 2, seed=4)
    options = bf.scenes.transport_options(max_depth=6, rr_first_depth=0)
    compare(hip, oracle, scene, frame_inputs(abi, oracle, frame, options, cone=(0.0, 2.0e-3)), options)

    s = bf.scenes.Scene()
    verts, faces = bf.scenes._icosphere(2)
    g = s.add_geometry(1, verts, verts, np.zeros((len(verts), 2)), faces)
    m = s.add_material(2, [spec(abi, SPEC_R, 0.6)])
    s.add_instance(3, 3, g, m)
    s.set_environment(1.0)
    mirror = s.finalize()
    cam = bf.scenes.look_at_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 16, 16, 0.9)
    frame = bf.scenes.frame_desc(cam, 2)
    options = bf.scenes.transport_options(max_depth=8, rr_first_depth=0)
    got, _ = compare(hip, oracle, mirror, frame_inputs(abi, oracle, frame, options, cone=(1.0e-4, 3.0e-3)), options)
    c = np.frombuffer(bytes(memoryview(got["cones"]).cast("B")), dtype=np.float32).reshape(-1, 2)
    assert np.all(c[:, 1] == np.float32(3.0e-3))


def test_depth_limited_paths_report_their_blocked_categories(bf, abi, oracle, hip):
    scene, cam = bf.scenes.cornell(16, 16)
    frame = bf.scenes.frame_desc(cam, 2, seed=1)
    options = bf.scenes.transport_options(max_depth=1, rr_first_depth=0)
    got, _ = compare(hip, oracle, scene, frame_inputs(abi, oracle, frame, options), options)
    ctl = as_u32(got["controls"])
    limited = ctl[:, 1] == abi.TERM_DEPTH_LIMIT
    assert limited.sum() > 50 and np.all(ctl[limited, 3] == abi.LOBE_DIFFUSE)


def test_invalid_inputs_fail_only_their_paths(bf, abi, oracle, hip):
    scene, cam = bf.scenes.cornell(8, 8)
    frame = bf.scenes.frame_desc(cam, 1)
    options = bf.scenes.transport_options(max_depth=4)
    options.failure_policy = abi.FAILURES_COUNT
    inputs = frame_inputs(abi, oracle, frame, options)
    inputs[1][3].t_min = -1.0
    inputs[3][7].wavelength_pdf_measures[2] = 1
    inputs[3][11].current_medium = 2
    inputs[1][11].current_medium = 2
    inputs[3][20].diffuse_depth = 300
    inputs[3][20].depth = 300
    got, report = compare(hip, oracle, scene, inputs, options)
    assert report.failed_paths == 4 and report.first_failure_slot == 3 and report.first_failure_code == 3
    st = as_u32(got["states"])
    assert np.all(st[[3, 7, 11, 20]] == 0) and np.all(as_u32(got["controls"])[[3, 7, 11, 20], 1] == 0)
    assert np.all(as_u32(got["controls"])[:, 0] == abi.PHASE_TERMINATED)


def test_path_input_arguments_are_checked(bf, abi, oracle, hip):
    import ctypes as C

    scene, cam = bf.scenes.cornell(4, 4)
    frame = bf.scenes.frame_desc(cam, 1)
    options = bf.scenes.transport_options(max_depth=2)
    inputs = frame_inputs(abi, oracle, frame, options)
    ws = hip.XpuWavefrontTransportWorkspace(8)
    with pytest.raises(hip.XpuError):  # 16 paths in a workspace of 8
        hip.trace_xpu_wavefront_paths(ws, hip.XpuScene(scene), inputs, options)
    lib = hip.library()
    report = abi.TransportReport()
    empty = abi.PathInputs()
    outs = abi.PathOutputs()
    assert lib.blackframe_hip_trace_wavefront_paths(ws.handle, hip.XpuScene(scene).handle, C.byref(empty),
                                                    C.byref(options), C.byref(outs), C.byref(report), None) != 0
