"""The oracle's path-input transport (trace_cuda_wavefront_transport's contract: per-path sample
stream, primary ray, ray cone and initial path state in; terminal path state, ray, cone and lane
control out), on the CPU.

Pins:
  - inputs generated from a frame as the camera stage generates them give, path for path, the
    frame transport's radiance, throughput, wavelengths, depth counters, eta scale and
    termination, bit for bit (the two entry points run one algorithm);
  - the terminal records follow the reference's stage kernels (CudaWavefrontStages.cu): a path
    that escapes at depth 0 ends on its own input ray; a path ended by roulette or zero throughput
    ends on its sampled continuation ray (t_min 0, t_max infinity, the input ray's time and mask);
    the wavelength pdfs and measures pass through; every control is phase terminated; a path that
    reaches its depth limit reports the depth categories that blocked it (diffuse for a Lambert
    scene);
  - ray cones, against geometry: a Lambert bounce floors the spread at 1 (the reference's diffuse
    floor), so every path past its first bounce leaves with spread 1; a mirror keeps the spread;
    the width at the first hit is spread x hit distance, with the distance from an independent
    closest-hit query of the primary ray;
  - invalid inputs (a negative t_min, a non-unit wavelength measure, a participating medium) fail
    their paths and only those.
"""
import ctypes as C
import math

import numpy as np
import pytest

from path_inputs import as_u32, frame_inputs


def run(bf, abi, oracle, scene, frame, options, cone=(0.0, 0.0)):
    inputs = frame_inputs(abi, oracle, frame, options, cone)
    out = oracle.scene(scene).render_paths(inputs, options, threads=8)
    return inputs, out


@pytest.mark.parametrize("materials,rr", [("lambert", 3), ("mixed", 2)])
def test_path_inputs_reproduce_the_frame_transport(bf, abi, oracle, materials, rr):
    scene, cam = bf.scenes.cornell(24, 24, materials=materials)
    frame = bf.scenes.frame_desc(cam, 4, seed=9)
    options = bf.scenes.transport_options(max_depth=8, rr_first_depth=rr)
    inputs, (states, rays, cones, controls, report) = run(bf, abi, oracle, scene, frame, options)
    expected = oracle.scene(scene).render(frame, options, threads=8)
    assert report.failed_paths == 0 and expected["report"].failed_paths == 0
    assert report.closest_hit_queries == expected["report"].closest_hit_queries
    assert report.shadow_queries == expected["report"].shadow_queries
    want = as_u32(expected["paths"])
    got = as_u32(states)
    assert np.array_equal(got[:, 4:8], want[:, 0:4])     # accumulated radiance
    assert np.array_equal(got[:, 0:4], want[:, 4:8])     # throughput
    assert np.array_equal(got[:, 8:12], want[:, 8:12])   # wavelengths
    counters = got[:, 16] | (got[:, 17] << 8) | (got[:, 18] << 16) | (got[:, 19] << 24)
    assert np.array_equal(counters, want[:, 14]) and np.array_equal(got[:, 21], want[:, 12])
    assert np.array_equal(got[:, 22], want[:, 15])       # eta scale
    ctl = as_u32(controls)
    assert np.all(ctl[:, 0] == abi.PHASE_TERMINATED) and np.array_equal(ctl[:, 1], want[:, 13])
    assert np.array_equal(got[:, 12:16], as_u32(inputs[3])[:, 12:16])  # wavelength pdfs pass through
    assert np.array_equal(got[:, 25], as_u32(inputs[3])[:, 25])        # measures pass through
    terminations = want[:, 13]
    assert (terminations == abi.TERM_RUSSIAN_ROULETTE).sum() > 10
    r_in, r_out = as_u32(inputs[1]), as_u32(rays)
    continued = (terminations == abi.TERM_RUSSIAN_ROULETTE) | (terminations == abi.TERM_ZERO_THROUGHPUT) | \
        ((terminations == abi.TERM_ESCAPED) & (want[:, 12] > 0))
    rr = np.frombuffer(bytes(memoryview(rays).cast("B")), dtype=np.float32).reshape(-1, 12)
    assert np.all(rr[continued, 3] == 0.0) and np.all(np.isinf(rr[continued, 7]))
    assert np.array_equal(r_out[continued, 8:10], r_in[continued, 8:10])  # time and visibility mask


def test_blocked_depth_limits(bf, abi, oracle):
    scene, cam = bf.scenes.cornell(16, 16)
    frame = bf.scenes.frame_desc(cam, 2, seed=1)
    options = bf.scenes.transport_options(max_depth=1, rr_first_depth=0)
    _, (states, rays, cones, controls, report) = run(bf, abi, oracle, scene, frame, options)
    ctl = as_u32(controls)
    limited = ctl[:, 1] == abi.TERM_DEPTH_LIMIT
    assert limited.sum() > 50
    assert np.all(ctl[limited, 3] == abi.LOBE_DIFFUSE)  # the depth category that ran out
    assert np.all(ctl[~limited, 3] == 0)


def first_hit_distance(abi, oracle, scene, rays):
    hits = oracle.scene(scene).closest_hit(rays)
    out = np.full(len(rays), np.nan)
    for k, h in enumerate(hits):
        if h.status == 1:
            d = np.array([rays[k].direction_x, rays[k].direction_y, rays[k].direction_z], dtype=np.float64)
            out[k] = h.hit.parameter * np.linalg.norm(d)
    return out


def test_ray_cones_follow_lambert_floor_and_hit_distance(bf, abi, oracle):
    scene, cam = bf.scenes.cornell(16, 16)
    frame = bf.scenes.frame_desc(cam, 2, seed=4)
    options = bf.scenes.transport_options(max_depth=6, rr_first_depth=0)
    spread0 = 2.0e-3
    inputs, (states, rays, cones, controls, report) = run(bf, abi, oracle, scene, frame, options, cone=(0.0, spread0))
    assert report.failed_paths == 0
    depth = as_u32(states)[:, 21]
    c = np.frombuffer(bytes(memoryview(cones).cast("B")), dtype=np.float32).reshape(-1, 2)
    assert np.all(c[depth == 0] == np.float32([0.0, spread0]))
    assert (depth >= 1).sum() > 100 and np.all(c[depth >= 1, 1] == 1.0)
    # one bounce: width = spread0 x the first hit's distance
    distance = first_hit_distance(abi, oracle, scene, inputs[1])
    one = depth == 1
    assert np.allclose(c[one, 0], spread0 * distance[one], rtol=2e-6, atol=0.0)
    # widths grow along the path
    assert np.all(c[depth >= 2, 0] >= c[depth >= 2, 0].min()) and np.all(c[depth >= 1, 0] > 0.0)


def test_ray_cones_keep_their_spread_at_mirrors(bf, abi, oracle):
    from test_oracle_specular import SPEC_R, spec

    s = bf.scenes.Scene()
    verts, faces = bf.scenes._icosphere(2)
    g = s.add_geometry(1, verts, verts, np.zeros((len(verts), 2)), faces)
    m = s.add_material(2, [spec(abi, SPEC_R, 0.6)])
    s.add_instance(3, 3, g, m)
    s.set_environment(1.0)
    s.finalize()
    cam = bf.scenes.look_at_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 16, 16, 0.9)
    frame = bf.scenes.frame_desc(cam, 2)
    options = bf.scenes.transport_options(max_depth=8, rr_first_depth=0)
    spread0 = 3.0e-3
    inputs, (states, rays, cones, controls, report) = run(bf, abi, oracle, s.finalize(), frame, options,
                                                          cone=(1.0e-4, spread0))
    assert report.failed_paths == 0
    depth = as_u32(states)[:, 21]
    c = np.frombuffer(bytes(memoryview(cones).cast("B")), dtype=np.float32).reshape(-1, 2)
    assert (depth >= 1).sum() > 50 and np.all(c[:, 1] == np.float32(spread0))
    r_in, r_out = as_u32(inputs[1]), as_u32(rays)
    missed = depth == 0  # pixels beside the sphere escape at once and end on their own input ray
    assert missed.sum() > 20 and np.array_equal(r_out[missed], r_in[missed])
    distance = first_hit_distance(abi, oracle, s, inputs[1])
    one = depth == 1
    assert np.allclose(c[one, 0], 1.0e-4 + spread0 * distance[one], rtol=2e-6, atol=0.0)


def test_invalid_inputs_fail_only_their_paths(bf, abi, oracle):
    scene, cam = bf.scenes.cornell(8, 8)
    frame = bf.scenes.frame_desc(cam, 1)
    options = bf.scenes.transport_options(max_depth=4)
    options.failure_policy = abi.FAILURES_COUNT
    inputs = frame_inputs(abi, oracle, frame, options)
    inputs[1][3].t_min = -1.0
    inputs[3][7].wavelength_pdf_measures[2] = 1
    inputs[3][11].current_medium = 2
    inputs[1][11].current_medium = 2
    states, rays, cones, controls, report = oracle.scene(scene).render_paths(inputs, options, threads=2)
    assert report.failed_paths == 3 and report.first_failure_slot == 3 and report.first_failure_code == 3
    ctl = as_u32(controls)
    assert np.all(ctl[[3, 7, 11], 1] == 0)
