"""Punctual lights in the oracle, pinned by closed forms: a Lambert floor under a point, spot or
directional light, rendered at one centred sample per pixel and depth 1, so each path's radiance
is the light's direct contribution rho/pi * I * falloff * cos(theta) / d^2 (point, spot) or
rho/pi * E * cos(theta) (directional), evaluated in float64 at the hit point of the path's own
binary32 camera ray. The sampling follows sample_punctual_light
(Blackframe/XPU/CUDA/Source/CudaWavefrontStages.cu:2240-2376)."""
import ctypes as C
import math

import numpy as np
import pytest

W = H = 24
RHO = 0.5


def floor_render(bf, abi, oracle, add_light):
    s = bf.scenes.Scene()
    m = s.add_material(1, [bf.scenes.lambert(RHO)])
    g = s.add_geometry(10, *bf.scenes.quad([[-8, 0, -8], [-8, 0, 8], [8, 0, 8], [8, 0, -8]], [0, 1, 0]))
    s.add_instance(20, 20, g, m)
    add_light(s)
    s.finalize()
    cam = bf.scenes.look_at_camera((0.0, 3.0, 0.0), (0.0, 0.0, -1e-3), W, H, 1.4, jitter=abi.JITTER_CENTER)
    frame = bf.scenes.frame_desc(cam, 1)
    options = bf.scenes.transport_options(max_depth=1, rr_first_depth=0)
    result = oracle.scene(s).render(frame, options, threads=4)
    assert result["report"].failed_paths == 0
    radiance = np.array([list(p.radiance) for p in result["paths"]], dtype=np.float64)
    hits = np.zeros((W * H, 3))
    ray, idx = abi.TransportRay(), abi.SampleStreamIndex()
    for slot in range(W * H):
        idx.pixel_x, idx.pixel_y = slot % W, slot // W
        assert oracle.lib.bfo_camera_ray(C.byref(cam), 1, C.byref(idx), C.byref(ray)) == 1
        o = np.array([ray.origin_x, ray.origin_y, ray.origin_z], dtype=np.float64)
        d = np.array([ray.direction_x, ray.direction_y, ray.direction_z], dtype=np.float64)
        hits[slot] = o + (-o[1] / d[1]) * d
    # every lane carries the same (constant-spectrum) value
    assert np.all(radiance == radiance[:, :1])
    return radiance[:, 0], hits


def check(values, expected, mask=None, slack=None):
    """Relative 2e-5, plus `slack` (absolute) where the closed form is ill-conditioned."""
    mask = np.ones_like(values, dtype=bool) if mask is None else mask
    slack = np.zeros_like(values) if slack is None else slack
    assert mask.sum() > 50
    excess = np.abs(values[mask] - expected[mask]) - 2e-5 * expected[mask] - slack[mask]
    assert excess.max() <= 0.0, excess.max()


def test_point_light_inverse_square(bf, abi, oracle):
    I, pos = 7.0, np.array([0.4, 1.3, -0.2])
    values, hits = floor_render(bf, abi, oracle, lambda s: s.add_point_light(tuple(pos), I, (1e-6, 1e-6, 1e-6)))
    to_light = pos - hits
    d = np.linalg.norm(to_light, axis=1)
    expected = RHO / math.pi * I * (to_light[:, 1] / d) / d**2
    check(values, expected)


def test_spot_light_smoothstep_falloff(bf, abi, oracle):
    I, pos, inner, outer = 5.0, np.array([0.0, 1.5, 0.0]), 0.35, 0.7
    axis = np.array([0.2, -1.0, 0.1])
    axis /= np.linalg.norm(axis)
    axis32 = axis.astype(np.float32).astype(np.float64)
    values, hits = floor_render(bf, abi, oracle, lambda s: s.add_spot_light(tuple(pos), tuple(axis32), inner, outer, I))
    to_light = pos - hits
    d = np.linalg.norm(to_light, axis=1)
    cone = (-to_light / d[:, None]) @ axis32
    ci, co = math.cos(np.float32(inner)), math.cos(np.float32(outer))
    t = np.clip((cone - co) / (ci - co), 0.0, 1.0)
    falloff = np.where(cone >= ci, 1.0, t * t * (3.0 - 2.0 * t))
    scale = RHO / math.pi * I * (to_light[:, 1] / d) / d**2
    expected = scale * falloff
    # binary32 cone cosines carry ~2^-23 error, magnified by d(falloff)/d(cone) = 6t(1-t)/(ci-co)
    slack = scale * 6.0 * t * (1.0 - t) / (ci - co) * 4.0 * 2.0**-23
    clear = (np.abs(cone - co) > 1e-5) & (np.abs(cone - ci) > 1e-5)
    inside = clear & (cone > co)
    check(values, expected, inside, slack)
    assert np.all(values[clear & (cone < co)] == 0.0)
    assert (clear & (cone < co)).sum() > 20 and (inside & (cone < ci)).sum() > 20


def test_directional_light_cosine(bf, abi, oracle):
    E = 2.5
    prop = np.array([0.3, -1.0, 0.2])
    prop32 = (prop / np.linalg.norm(prop)).astype(np.float32).astype(np.float64)
    values, hits = floor_render(bf, abi, oracle, lambda s: s.add_directional_light(tuple(prop32), E))
    expected = np.full(values.shape, RHO / math.pi * E * -prop32[1])
    check(values, expected)


def test_point_light_is_shadowed(bf, abi, oracle):
    """An occluder between the floor and the light zeroes the contribution under it."""
    def add(s):
        s.add_point_light((0.0, 1.0, 0.0), 4.0)
        blocker = s.add_material(3, [bf.scenes.lambert(0.0)])
        g = s.add_geometry(12, *bf.scenes.quad([[-0.3, 0.5, -0.3], [0.3, 0.5, -0.3], [0.3, 0.5, 0.3], [-0.3, 0.5, 0.3]],
                                                [0, -1, 0]))
        s.add_instance(22, 22, g, blocker, mask=0xFFFFFFFF)
    values, hits = floor_render(bf, abi, oracle, add)
    # the blocker is seen from above (black); floor points in its shadow cone get nothing
    shadowed = (np.abs(hits[:, 0]) < 0.55) & (np.abs(hits[:, 2]) < 0.55)
    lit = (np.abs(hits[:, 0]) > 0.7) | (np.abs(hits[:, 2]) > 0.7)
    assert np.all(values[shadowed] == 0.0)
    assert np.all(values[lit] > 0.0)
