"""Closed-form pins of the oracle's transport pieces outside the closure library.

  - wavelengths: four circularly stratified samples 360 + 470 * frac(u + k/4) (WavelengthSampling.hpp:62-84);
  - MIS: balance a/(a+b) and power a^2/(a^2+b^2) (TransportDevice.cuh:592-689);
  - Russian roulette: survival probability clamp(max(beta) * eta, min, max), survivors divided by it
    (TransportDevice.cuh:711-782);
  - triangle area against exact rational arithmetic: the cross product is resolved from exact
    float expansions and only its length is rounded in binary32, so the area is within 4 ulps
    (CudaWavefrontStages.cu:1724-2042);
  - pinhole camera: the centre ray is the view direction and corner rays open the field of view;
  - white furnace: a convex Lambert sphere of albedo rho in a unit environment returns exactly rho
    on every path that hits it and 1 on every path that escapes;
  - direct lighting: a Lambert floor under a parallel emissive rectangle converges to
    rho * L * F with F the closed-form point-to-rectangle form factor, for balance and power MIS.
"""
import ctypes as C
import math
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest


def fbuf(values):
    return (C.c_float * len(values))(*[float(v) for v in values])


def test_wavelength_stratification(oracle):
    lib = oracle.lib
    for u in list(np.linspace(0.0, 1.0, 97, endpoint=False)) + [np.nextafter(np.float32(1), np.float32(0))]:
        out = fbuf([0] * 4)
        lib.bfo_sample_wavelengths(float(u), out)
        lanes = np.array(list(out), dtype=np.float64)
        expected = 360.0 + 470.0 * np.mod(np.float64(np.float32(u)) + np.arange(4) / 4.0, 1.0)
        assert np.all(np.abs(lanes - expected) < 1e-4), (u, lanes, expected)
        assert np.all((lanes >= 360.0) & (lanes <= 830.0))  # fma rounding may reach 830 at u -> 1
        assert abs(np.sort(lanes)[1:] - np.sort(lanes)[:-1] - 117.5).max() < 1e-3


def mis(lib, heuristic, a, b):
    w = C.c_float()
    assert lib.bfo_mis_weight(heuristic, a, b, C.byref(w)) == 1
    return w.value


def test_mis_closed_forms(oracle):
    lib = oracle.lib
    rng = np.random.default_rng(4)
    for _ in range(2000):
        a, b = np.float32(10 ** rng.uniform(-6, 6)), np.float32(10 ** rng.uniform(-6, 6))
        a64, b64 = float(a), float(b)
        assert abs(mis(lib, 0, a, b) - a64 / (a64 + b64)) <= 2.5e-7
        assert abs(mis(lib, 1, a, b) - a64 * a64 / (a64 * a64 + b64 * b64)) <= 2.5e-7
        assert abs(mis(lib, 0, a, b) + mis(lib, 0, b, a) - 1.0) <= 2.5e-7
    assert mis(lib, 0, 0.3, 0.0) == 1.0 and mis(lib, 1, 0.3, 0.0) == 1.0


def test_russian_roulette_policy(oracle):
    lib = oracle.lib
    rng = np.random.default_rng(9)
    for _ in range(2000):
        beta = rng.uniform(0, 1.5, 4).astype(np.float32)
        eta = np.float32(rng.choice([1.0, 2.25, 1 / 2.25]))
        u = np.float32(rng.random())
        buf, survived = fbuf(beta), C.c_int()
        assert lib.bfo_russian_roulette(buf, float(eta), float(u), 0.1, 0.9, C.byref(survived)) == 1
        p = min(0.9, max(0.1, float(beta.max()) * float(eta)))
        if abs(float(u) - p) < 1e-6:
            continue
        assert bool(survived.value) == (float(u) < p)
        if survived.value:
            assert np.allclose(np.array(list(buf)), beta / p, rtol=3e-7)


def exact_area(v):
    f = [[Fraction(float(x)) for x in row] for row in v]
    e1 = [f[1][i] - f[0][i] for i in range(3)]
    e2 = [f[2][i] - f[0][i] for i in range(3)]
    c = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
    sq = sum(x * x for x in c)
    getcontext().prec = 60
    return float(Decimal(sq.numerator).sqrt() / Decimal(sq.denominator).sqrt() / 2)


def test_triangle_area_is_exact(oracle):
    lib = oracle.lib
    rng = np.random.default_rng(12)
    for k in range(500):
        scale = 10 ** rng.uniform(-3, 3)
        v = (rng.uniform(-1, 1, (3, 3)) * scale).astype(np.float32)
        if k % 5 == 0:  # slivers
            v[2] = v[0] + (v[1] - v[0]) * np.float32(rng.random()) + np.float32(1e-4 * scale) * rng.standard_normal(3).astype(np.float32)
        area, normal = C.c_float(), fbuf([0, 0, 0])
        if not lib.bfo_triangle_area(fbuf(v.reshape(-1)), C.byref(area), normal):
            continue
        exact = exact_area(v)
        ulp = np.spacing(np.float32(exact))
        assert abs(area.value - exact) <= 4 * ulp, (k, area.value, exact)
        n = np.array(list(normal), dtype=np.float64)
        cross = np.cross(v[1].astype(np.float64) - v[0], v[2].astype(np.float64) - v[0])
        assert abs(np.linalg.norm(n) - 1) < 1e-6 and np.dot(n, cross / np.linalg.norm(cross)) > 1 - 1e-5


def test_pinhole_camera_rays(bf, abi, oracle):
    cam = bf.scenes.look_at_camera((0.3, 1.0, 2.0), (0.3, -1.0, 0.0), 65, 49, 0.8, jitter=abi.JITTER_CENTER)
    forward = np.array([0.0, -2.0, -2.0]) / math.sqrt(8.0)
    ray = abi.TransportRay()
    idx = abi.SampleStreamIndex()
    idx.pixel_x, idx.pixel_y = 32, 24
    assert oracle.lib.bfo_camera_ray(C.byref(cam), 1, C.byref(idx), C.byref(ray)) == 1
    d = np.array([ray.direction_x, ray.direction_y, ray.direction_z])
    assert np.allclose(d, forward, atol=2e-7)
    assert (ray.origin_x, ray.origin_y, ray.origin_z) == (np.float32(0.3), 1.0, 2.0)
    # the top-left pixel centre sits (64/65, 48/49) of the half extents away from the axis
    idx.pixel_x, idx.pixel_y = 0, 0
    oracle.lib.bfo_camera_ray(C.byref(cam), 1, C.byref(idx), C.byref(ray))
    d = np.array([ray.direction_x, ray.direction_y, ray.direction_z])
    hh = math.tan(np.float32(0.8) / 2)
    hw = hh * 65 / 49
    local = np.array([-hw * 64 / 65, hh * 48 / 49, -1.0])
    assert abs(np.dot(d, forward) - 1 / np.linalg.norm(local)) < 1e-6


def faceted_sphere(bf, s, subdivisions, radius):
    verts, faces = bf.scenes._icosphere(subdivisions)
    p, n = [], []
    for a, b, c in faces:
        tri = verts[[a, b, c]] * radius
        normal = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        p.extend(tri)
        n.extend([normal / np.linalg.norm(normal)] * 3)
    idx = np.arange(len(p)).reshape(-1, 3)
    return s.add_geometry(40, np.array(p), np.array(n), np.zeros((len(p), 2)), idx)


@pytest.mark.parametrize("rho", [0.5, 0.83])
def test_white_furnace_convex_lambert(bf, abi, oracle, rho):
    s = bf.scenes.Scene()
    m = s.add_material(1, [bf.scenes.lambert(rho)])
    g = faceted_sphere(bf, s, 2, 1.0)
    s.add_instance(2, 2, g, m)
    s.set_environment(1.0)
    s.finalize()
    cam = bf.scenes.look_at_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 24, 24, 0.9)
    frame = bf.scenes.frame_desc(cam, 4)
    options = bf.scenes.transport_options(max_depth=8, rr_first_depth=0)
    result = oracle.scene(s).render(frame, options, threads=4)
    radiance = np.array([list(p.radiance) for p in result["paths"]])
    depth = np.array([p.depth for p in result["paths"]])
    assert result["report"].failed_paths == 0
    hit = depth > 0
    assert hit.sum() > 100 and (~hit).sum() > 100
    assert np.all(radiance[~hit] == 1.0)
    assert np.abs(radiance[hit] - np.float32(rho)).max() <= 4 * np.spacing(np.float32(rho))


def rectangle_form_factor(a, b):
    """Form factor from a differential area to a parallel rectangle with one corner above it,
    sides a and b in units of the separation."""
    A, B = math.sqrt(1 + a * a), math.sqrt(1 + b * b)
    return (a / A * math.atan(b / A) + b / B * math.atan(a / B)) / (2 * math.pi)


@pytest.mark.parametrize("mis", [0, 1])
def test_direct_lighting_form_factor(bf, abi, oracle, mis):
    rho, radiance, h, half = 0.5, 3.0, 1.0, 0.6
    s = bf.scenes.Scene()
    floor = s.add_material(1, [bf.scenes.lambert(rho)])
    light = s.add_material(2, [bf.scenes.lambert(0.0)], emission=radiance)
    g_floor = s.add_geometry(10, *bf.scenes.quad([[-4, 0, -4], [-4, 0, 4], [4, 0, 4], [4, 0, -4]], [0, 1, 0]))
    g_light = s.add_geometry(11, *bf.scenes.quad([[-half, h, -half], [half, h, -half], [half, h, half], [-half, h, half]],
                                                 [0, -1, 0]))
    s.add_instance(20, 20, g_floor, floor)
    s.add_instance(21, 21, g_light, light)
    s.finalize()
    cam = bf.scenes.look_at_camera((0.0, 0.5, 0.0), (0.0, 0.0, -1e-3), 32, 32, 0.02)
    frame = bf.scenes.frame_desc(cam, 32)
    options = bf.scenes.transport_options(max_depth=1, mis=mis, rr_first_depth=0)
    result = oracle.scene(s).render(frame, options, threads=8)
    values = np.array([list(p.radiance) for p in result["paths"]], dtype=np.float64).reshape(-1)
    expected = rho * radiance * 4 * rectangle_form_factor(half / h, half / h)
    error = values.std() / math.sqrt(values.size)
    assert abs(values.mean() - expected) < 4 * error + 1e-3 * expected, (values.mean(), expected, error)


RASTER_X, RASTER_Y, WAVELENGTH = 0xA24BAED4963EE407, 0x9FB21C651E98DF25, 3


def test_wavelength_known_answer(oracle, abi):
    """Tests/Renderer/WavelengthSamplingTest.cpp:156-170: the default SampleStreamIndex's
    wavelength dimension and its four stratified wavelengths, exactly."""
    i = abi.SampleStreamIndex(0, 0, 0, 0)
    u = oracle.lib.bfo_sample_1d(abi.SAMPLER_INDEXED_HASH, C.byref(i), WAVELENGTH)
    assert np.float32(u) == np.float32(0.23662316799163818)
    out = fbuf([0] * 4)
    oracle.lib.bfo_sample_wavelengths(u, out)
    assert [np.float32(v) for v in out] == [np.float32(v) for v in (471.212890625, 588.712890625, 706.212890625,
                                                                    823.712890625)]


def test_pixel_jitter_known_answer(oracle, abi):
    """Tests/Renderer/PixelJitterTest.cpp:142-152: uniform jitter of pixel (17, 29), sample 27,
    seed 42 on the raster dimensions, exactly."""
    i = abi.SampleStreamIndex(17, 29, 27, 42)
    ox = oracle.lib.bfo_sample_1d(abi.SAMPLER_INDEXED_HASH, C.byref(i), RASTER_X)
    oy = oracle.lib.bfo_sample_1d(abi.SAMPLER_INDEXED_HASH, C.byref(i), RASTER_Y)
    assert np.float32(ox) == np.float32(0x4D67A7 * 2.0**-24)
    assert np.float32(oy) == np.float32(0x6B1427 * 2.0**-24)


def _camera(abi, origin, tangent, bitangent, normal, width, height, fov, t_min=0.0, t_max=float("inf"), time=0.0,
            mask=0xFFFFFFFF):
    cam = abi.CameraDesc()
    for k in range(3):
        cam.origin[k], cam.tangent[k], cam.bitangent[k], cam.normal[k] = origin[k], tangent[k], bitangent[k], normal[k]
    cam.half_height = float(np.float32(math.tan(fov / 2.0)))
    cam.half_width = float(np.float32(cam.half_height * width / height))
    cam.t_min, cam.t_max, cam.time = t_min, t_max, time
    cam.width, cam.height, cam.visibility_mask, cam.jitter_mode = width, height, mask, abi.JITTER_CENTER
    return cam


def _ray(oracle, abi, cam, px, py):
    i = abi.SampleStreamIndex(px, py, 0, 0)
    ray = abi.TransportRay()
    assert oracle.lib.bfo_camera_ray(C.byref(cam), abi.SAMPLER_INDEXED_HASH, C.byref(i), C.byref(ray)) == 1
    return ray


def test_pinhole_camera_known_answers(oracle, abi):
    """Tests/Renderer/PinholeCameraTest.cpp:19-76, with the pixel centre (jitter off) standing for
    the raster samples (1.5, 1.5), (0.5, 0.5) and (1.5, 1.5):
      - the centre of a 3x3 image looks exactly along -normal from the origin, with the camera's
        t range, time and visibility mask;
      - on a 2x2 image with a 90 degree field of view the upper-left sample maps to -x, +y, -z and
        the lower-right one to the exact mirror in x and y (an x or y flip fails here);
      - the camera basis maps to world space: normal (1, 0, 0) gives the centre ray (-1, 0, 0)."""
    cam = _camera(abi, (1.0, -2.0, 3.0), (1, 0, 0), (0, 1, 0), (0, 0, 1), 3, 3, 1.0, 0.25, 100.0, 0.75, 0x00FF00FF)
    r = _ray(oracle, abi, cam, 1, 1)
    assert (r.origin_x, r.origin_y, r.origin_z) == (1.0, -2.0, 3.0)
    assert (r.direction_x, r.direction_y, r.direction_z) == (0.0, 0.0, -1.0)
    assert (np.float32(r.t_min), np.float32(r.t_max), np.float32(r.time), r.visibility_mask) == \
        (np.float32(0.25), np.float32(100.0), np.float32(0.75), 0x00FF00FF)
    cam = _camera(abi, (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), 2, 2, math.pi / 2)
    upper_left, lower_right = _ray(oracle, abi, cam, 0, 0), _ray(oracle, abi, cam, 1, 1)
    assert upper_left.direction_x < 0 and upper_left.direction_y > 0 and upper_left.direction_z < 0
    assert lower_right.direction_x == -upper_left.direction_x and lower_right.direction_y == -upper_left.direction_y
    assert lower_right.direction_z == upper_left.direction_z
    cam = _camera(abi, (4.0, 5.0, 6.0), (0, 1, 0), (0, 0, 1), (1, 0, 0), 1, 1, 1.0, 0.0, 1.0)
    r = _ray(oracle, abi, cam, 0, 0)
    assert (r.origin_x, r.origin_y, r.origin_z) == (4.0, 5.0, 6.0)
    assert (r.direction_x, r.direction_y, r.direction_z) == (-1.0, 0.0, 0.0)
