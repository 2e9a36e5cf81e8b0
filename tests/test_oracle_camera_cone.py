"""Pins of the primary-ray cones that camera frames of mapped scenes start with (the reference's
generate_primary_ray_cone, PinholeCamera.hpp:224-252, reduced by ray_cone_from_differential,
RayCone.hpp:194-220), by facts the geometry fixes:

  - every cone has zero width (the differential rays share the camera origin);
  - at the image centre the spread is the image-plane distance of one pixel, 2 tan(fov/2) / height;
  - everywhere else it equals a binary64 evaluation of the same construction whose largest
    singular value comes from numpy's SVD, not the 2x2 Gram form, at uniform-jitter samples over
    the whole film (the last column and row included: their differentials extrapolate past it);
    a planted mistake in that evaluation (the raster point without its sub-pixel offset) fails.
"""
import ctypes as C
import math

import numpy as np


def direction64(hw, hh, frame, nx, ny):
    u = np.array([nx * hw, ny * hh, -1.0])
    u /= np.linalg.norm(u)
    w = frame @ u
    return w / np.linalg.norm(w)


def spread64(cam, raster_x, raster_y):
    hw, hh = float(cam.half_width), float(cam.half_height)
    frame = np.array([cam.tangent[:], cam.bitangent[:], cam.normal[:]], dtype=np.float64).T
    w, h = cam.width, cam.height
    d = direction64(hw, hh, frame, 2 * raster_x / w - 1, 1 - 2 * raster_y / h)
    slopes = []
    for nx, ny in ((2 * (raster_x + 1) / w - 1, 1 - 2 * raster_y / h), (2 * raster_x / w - 1, 1 - 2 * (raster_y + 1) / h)):
        a = direction64(hw, hh, frame, nx, ny)
        slopes.append(a / (a @ d) - d)
    return np.linalg.svd(np.stack(slopes, axis=1), compute_uv=False)[0]


def cone_of(oracle, abi, cam, sampler, px, py, index):
    idx = abi.SampleStreamIndex()
    idx.pixel_x, idx.pixel_y, idx.sample_index, idx.seed = px, py, index, 11
    ray, cone = abi.TransportRay(), abi.RayCone()
    assert oracle.lib.bfo_camera_ray(C.byref(cam), sampler, C.byref(idx), C.byref(ray)) == 1
    assert oracle.lib.bfo_camera_cone(C.byref(cam), sampler, C.byref(idx), C.byref(ray), C.byref(cone)) == 1
    return cone


def test_image_centre_spread_is_one_pixel(bf, abi, oracle):
    cam = bf.scenes.look_at_camera((0.3, 1.0, 2.0), (0.3, -1.0, 0.0), 65, 49, 0.8, jitter=abi.JITTER_CENTER)
    cone = cone_of(oracle, abi, cam, 1, 32, 24, 0)
    assert cone.width == 0.0
    want = 2 * math.tan(0.5 * float(np.float32(0.8))) / 49
    assert abs(cone.spread - want) <= 2e-6 * want, (cone.spread, want)


def test_spread_equals_a_binary64_evaluation(bf, abi, oracle):
    cam = bf.scenes.look_at_camera((0.0, 0.5, 3.0), (0.2, 0.1, -1.0), 96, 54, 1.1, jitter=abi.JITTER_UNIFORM)
    sampler = abi.SAMPLER_INDEXED_HASH
    rng = np.random.default_rng(4)
    pixels = [(0, 0), (95, 53), (95, 0), (0, 53), (48, 27)] + [tuple(p) for p in rng.integers(0, (96, 54), (40, 2))]
    planted = 0
    for k, (px, py) in enumerate(pixels):
        cone = cone_of(oracle, abi, cam, sampler, int(px), int(py), k)
        idx = abi.SampleStreamIndex()
        idx.pixel_x, idx.pixel_y, idx.sample_index, idx.seed = int(px), int(py), k, 11
        ray = abi.TransportRay()
        oracle.lib.bfo_camera_ray(C.byref(cam), sampler, C.byref(idx), C.byref(ray))
        # recover the raster point from the ray itself (its camera-space direction)
        frame = np.array([cam.tangent[:], cam.bitangent[:], cam.normal[:]], dtype=np.float64)
        local = frame @ np.array([ray.direction_x, ray.direction_y, ray.direction_z], dtype=np.float64)
        rx = (local[0] / -local[2] / float(cam.half_width) + 1) * cam.width / 2
        ry = (1 - local[1] / -local[2] / float(cam.half_height)) * cam.height / 2
        assert px <= rx < px + 1 + 1e-4 and py <= ry < py + 1 + 1e-4, (px, py, rx, ry)
        want = spread64(cam, rx, ry)
        assert cone.width == 0.0
        assert abs(cone.spread - want) <= 5e-5 * want, (px, py, cone.spread, want)
        wrong = spread64(cam, float(px), float(py))
        planted += abs(cone.spread - wrong) > 5e-5 * want
    assert planted >= len(pixels) // 2
