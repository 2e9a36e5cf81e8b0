"""Batched transport: frames of different scenes share one set of queues, and every frame's
paths, queue-independent counters and film are bit-identical to rendering it alone (checked
against the oracle frame by frame). Also: bounce groups replayed as hipGraphs, and a workspace
too small for the batch (several launches, a frame split across launches) give the same bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def path_arrays(paths):
    return np.frombuffer(bytes(memoryview(paths).cast("B")), dtype=np.uint32).reshape(len(paths), 16)


def frames(bf):
    cornell_mixed, cam_a = bf.scenes.cornell(32, 32, materials="mixed")
    grid, cam_b = bf.scenes.icosphere_grid(48, 27, grid=4, subdivisions=2)
    cornell, cam_c = bf.scenes.cornell(16, 16)
    return [(cornell_mixed, bf.scenes.frame_desc(cam_a, 4, seed=0)),
            (grid, bf.scenes.frame_desc(cam_b, 2, seed=7)),
            (cornell, bf.scenes.frame_desc(cam_c, 3, seed=3, first_sample_index=5))]


def check_frames(bf, oracle, hip, capacity, launch):
    items = frames(bf)
    options = bf.scenes.transport_options(max_depth=10, mis=bf.abi.MIS_POWER)
    options.launch = launch
    xscenes = [hip.XpuScene(scene) for scene, _ in items]
    ws = hip.XpuWavefrontTransportWorkspace(capacity)
    got = hip.trace_xpu_wavefront_batch(ws, [(x, f) for x, (_, f) in zip(xscenes, items)], options, want_paths=True)
    assert got["report"].failed_paths == 0
    total_rays = 0
    for k, (scene, frame) in enumerate(items):
        expected = oracle.scene(scene).render(frame, options, threads=8)
        assert np.array_equal(path_arrays(got["paths"][k]), path_arrays(expected["paths"])), f"frame {k}"
        film = got["films"][k].host()
        assert np.array_equal(film["film"], expected["film"]), f"frame {k}"
        assert np.array_equal(film["film_counts"], expected["film_counts"])
        total_rays += expected["report"].closest_hit_queries + expected["report"].shadow_queries
    report = got["report"]
    assert report.closest_hit_queries + report.shadow_queries == total_rays


def test_batch_matches_separate_frames(bf, oracle, hip):
    check_frames(bf, oracle, hip, capacity=1 << 20, launch=0)


def test_batch_with_graph_replay(bf, oracle, hip):
    check_frames(bf, oracle, hip, capacity=1 << 20, launch=bf.abi.LAUNCH_GRAPH | 4)


def test_batch_split_across_launches(bf, oracle, hip):
    # 5 000 slots: the 4-pass 32x32 frame takes 4 launches, the others share launches
    check_frames(bf, oracle, hip, capacity=5000, launch=bf.abi.LAUNCH_GRAPH | 2)


def test_single_frame_graph_replay_bit_exact(bf, oracle, hip):
    scene, cam = bf.scenes.cornell(32, 32, materials="mixed")
    frame = bf.scenes.frame_desc(cam, 4, seed=0)
    options = bf.scenes.transport_options(max_depth=8)
    expected = oracle.scene(scene).render(frame, options, threads=8)
    options.launch = bf.abi.LAUNCH_GRAPH | 6
    ws = hip.XpuWavefrontTransportWorkspace(32 * 32 * 4)
    xs = hip.XpuScene(scene)
    for replay in range(3):  # the captured group is replayed across calls too
        got = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, want_paths=True)
        assert np.array_equal(path_arrays(got["paths"]), path_arrays(expected["paths"])), f"call {replay}"
        assert np.array_equal(got["film"].host()["film"], expected["film"])


@pytest.mark.parametrize("lanes,order", [(64, 0), (256, 0), (128, 0x00100000), (128, 0x00200000), (256, 0x00300000)])
def test_traversal_workgroup_sizes_bit_exact(bf, oracle, hip, lanes, order):
    """The autotuner's traversal workgroup sizes and the shade / traversal orders
    (BF_LAUNCH_SHADE_ORDER, BF_LAUNCH_TRACE_ORDER) change the launch shape, never the result."""
    scene, cam = bf.scenes.icosphere_grid(48, 27, grid=4, subdivisions=2)
    frame = bf.scenes.frame_desc(cam, 2, seed=11)
    options = bf.scenes.transport_options(max_depth=16, mis=bf.abi.MIS_POWER)
    options.record_queues = 1
    expected = oracle.scene(scene).render(frame, options, threads=8)
    options.launch = bf.abi.launch_trace_block(lanes) | order
    ws = hip.XpuWavefrontTransportWorkspace(48 * 27 * 2)
    got = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene(scene), frame, options, want_paths=True, want_counters=True)
    assert np.array_equal(path_arrays(got["paths"]), path_arrays(expected["paths"]))
    assert np.array_equal(got["film"].host()["film"], expected["film"])
    assert np.array_equal(got["traversal_counters"], expected["traversal_counters"])
    for b in range(expected["report"].bounces):
        assert got["report"].ray_queue_hash[b] == expected["report"].ray_queue_hash[b]


def test_autotuner_choice_is_valid_and_bit_exact(bf, hip):
    """The autotuner times 18 launch shapes, then three work orders on the best one, then half and
    a quarter of the workspace in flight where a sample pass still fits, and returns the fastest;
    rendering with it gives the same bits as the default launch."""
    items = frames(bf)
    options = bf.scenes.transport_options(max_depth=10, mis=bf.abi.MIS_POWER)
    xscenes = [hip.XpuScene(scene) for scene, _ in items]
    ws = hip.XpuWavefrontTransportWorkspace(1 << 16)
    tuning = hip.autotune_launch(ws, [(x, f) for x, (_, f) in zip(xscenes, items)], options, trial_passes=2)
    largest = max(f.camera.width * f.camera.height for _, f in items)
    fractions = sum(1 for k in (1, 2) if (1 << 16) >> k >= largest)
    assert tuning.candidates == 21 + fractions
    assert tuning.best_ms <= tuning.default_ms and tuning.best_ms > 0.0
    assert tuning.launch == 0 or tuning.launch in list(tuning.tried)[:tuning.candidates]
    for k in range(21, tuning.candidates):
        assert tuning.tried[k] & bf.abi.LAUNCH_INFLIGHT_MASK == bf.abi.launch_in_flight(k - 20)
    reference = hip.trace_xpu_wavefront_batch(ws, [(x, f) for x, (_, f) in zip(xscenes, items)], options, want_paths=True)
    options.launch = tuning.launch
    tuned = hip.trace_xpu_wavefront_batch(ws, [(x, f) for x, (_, f) in zip(xscenes, items)], options, want_paths=True)
    for k in range(len(items)):
        assert np.array_equal(path_arrays(tuned["paths"][k]), path_arrays(reference["paths"][k]))
        assert np.array_equal(tuned["films"][k].host()["film"], reference["films"][k].host()["film"])


def test_paths_in_flight_fraction_is_bit_exact(bf, hip):
    """A launch word holding a quarter of the workspace in flight splits the batch into more
    launches of whole sample passes; every frame's paths and film are those of the default."""
    items = frames(bf)
    options = bf.scenes.transport_options(max_depth=10, mis=bf.abi.MIS_POWER)
    xscenes = [hip.XpuScene(scene) for scene, _ in items]
    ws = hip.XpuWavefrontTransportWorkspace(8192)  # 7 456 paths: one launch, or several of <= 2 048
    jobs = [(x, f) for x, (_, f) in zip(xscenes, items)]
    reference = hip.trace_xpu_wavefront_batch(ws, jobs, options, want_paths=True)
    options.launch = bf.abi.launch_in_flight(2)
    split = hip.trace_xpu_wavefront_batch(ws, jobs, options, want_paths=True)
    for k in range(len(items)):
        assert np.array_equal(path_arrays(split["paths"][k]), path_arrays(reference["paths"][k]))
        assert np.array_equal(split["films"][k].host()["film"], reference["films"][k].host()["film"])
