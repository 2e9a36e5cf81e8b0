"""Full wavefront transport parity: HIP vs the oracle's 7-queue CPU wavefront on the Cornell box.

Bar: per-path radiance, throughput, depth and termination bit-identical; ray/shadow queue contents
per bounce identical (FNV-1a of the slot order); float film identical; compensated film within
1e-4 relative per pixel (it is computed in double on both sides and expected identical)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_both(bf, oracle, hip, width, height, spp, scene_factory=None, **opts):
    scene, cam = (scene_factory or bf.scenes.cornell)(width, height)
    frame = bf.scenes.frame_desc(cam, spp, seed=opts.pop("seed", 0))
    options = bf.scenes.transport_options(**opts)
    options.record_queues = 1
    expected = oracle.scene(scene).render(frame, options, threads=8)
    ws = hip.XpuWavefrontTransportWorkspace(width * height * spp)
    got = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene(scene), frame, options, want_paths=True, want_counters=True)
    return expected, got


def path_arrays(paths):
    return np.frombuffer(bytes(memoryview(paths).cast("B")), dtype=np.uint32).reshape(len(paths), 16)


@pytest.mark.parametrize("opts", [
    dict(),
    dict(sampler=0, wavelength_mode=0),
    dict(mis=1, max_depth=4, rr_first_depth=2),
])
def test_cornell_transport_bit_exact(bf, oracle, hip, opts):
    check_parity(*run_both(bf, oracle, hip, 32, 32, 4, **dict(opts)))


@pytest.mark.parametrize("opts", [dict(), dict(mis=1, max_depth=12)])
def test_cornell_microfacet_transport_bit_exact(bf, oracle, hip, opts):
    """Gold GGX conductor, GGX dielectric, Oren–Nayar wall and a Lambert + anisotropic GGX mixture."""
    factory = lambda w, h: bf.scenes.cornell(w, h, materials="mixed")
    check_parity(*run_both(bf, oracle, hip, 32, 32, 4, scene_factory=factory, **dict(opts)))


def test_icosphere_grid_transport_bit_exact(bf, oracle, hip):
    """A 4x4 grid of the plain-forward-at-scale generator: mixed Lambert / conductor / dielectric,
    64 emissive triangles, environment, depth 16, power MIS."""
    factory = lambda w, h: bf.scenes.icosphere_grid(w, h, grid=4, subdivisions=2)
    check_parity(*run_both(bf, oracle, hip, 48, 27, 2, scene_factory=factory, mis=1, max_depth=16))


def check_parity(expected, got):
    er, gr = expected["report"], got["report"]
    assert er.failed_paths == 0 and gr.failed_paths == 0
    assert gr.bounces == er.bounces
    for b in range(er.bounces):
        assert gr.ray_queue_size[b] == er.ray_queue_size[b]
        assert gr.ray_queue_hash[b] == er.ray_queue_hash[b], f"ray queue order differs at bounce {b}"
        assert gr.shadow_queue_size[b] == er.shadow_queue_size[b]
        assert gr.shadow_queue_hash[b] == er.shadow_queue_hash[b], f"shadow queue order differs at bounce {b}"
    assert np.array_equal(path_arrays(got["paths"]), path_arrays(expected["paths"]))
    film = got["film"].host()
    assert np.array_equal(film["film"], expected["film"])
    assert np.array_equal(film["film_counts"], expected["film_counts"])
    ref = expected["film_compensated"]
    rel = np.abs(film["film_compensated"] - ref) / np.maximum(np.abs(ref), 1e-12)
    assert rel.max() <= 1e-4
    for field in ("closest_hit_queries", "shadow_queries", "shade_lanes", "closure_samples", "light_samples"):
        assert getattr(gr, field) == getattr(er, field), field
    assert np.array_equal(got["traversal_counters"], expected["traversal_counters"])
