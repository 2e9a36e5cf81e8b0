"""Pins of the oracle's surface maps (normal and bump maps over EWA-filtered image mip chains,
CudaWavefrontStages.cu:2992-3673) by facts the mathematics fixes:

  - exp, expm1 and log of blackframe_math.h equal binary64 libm results rounded to binary32;
  - a constant image filters to exactly its value at every footprint and level;
  - the filter equals an independent binary64 evaluation of the same elliptical Gaussian (principal
    axes from an eigen-decomposition of the texel-space covariance, not the atan2 form), for
    isotropic, anisotropic and rotated footprints, the four wrap modes, and a two-level mip blend;
    planting a transposed Jacobian in that evaluation makes the anisotropic case fail;
  - the ray cone's (u, v) differentials on a square patch seen head-on are width / side length;
  - a flat normal map leaves the shading normal; a constant tilted one tilts it along the tangent
    (or the bitangent, flipped by the y convention) by the encoded angle;
  - a linear height ramp bumps the normal to normalize((-scale * slope, 0, side length));
  - a footprint wider than the whole image reads its 1x1 level (the mean) under repeat wrap;
  - a footprint over the texel budget is refused (status 3).
"""
import math

import numpy as np
import pytest

F32 = np.float32


def test_math_functions_are_rounded_binary64(oracle):
    lib = oracle.lib
    xs = np.concatenate([np.linspace(-19.0, 2.5, 4001), np.geomspace(1e-7, 1e-2, 400)]).astype(np.float32)
    for x in xs:
        assert lib.bfo_expm1f(float(x)) == F32(math.expm1(float(x))), x
        assert lib.bfo_expf(float(x)) == F32(math.exp(float(x))), x
    for x in np.geomspace(1e-30, 1e30, 5000).astype(np.float32):
        assert lib.bfo_logf(float(x)) == F32(math.log(float(x))), x


def chain(bf, image):
    return bf.scenes.mip_chain(np.asarray(image, np.float64))


@pytest.mark.parametrize("g", [(1e-3, 0.0, 0.0, 1e-3), (0.02, 0.004, 0.0, 0.01), (0.3, 0.1, 0.0, 0.2)])
def test_constant_image_filters_to_its_value(bf, oracle, g):
    levels = chain(bf, np.full((32, 24, 3), 0.625))
    for u, v in ((0.5, 0.5), (0.01, 0.99), (1.7, -0.3)):
        for ch in range(3):
            st, value = oracle.map_filter(levels, u, v, g, channel=ch)
            assert st == 0 and value == F32(0.625), (g, u, v, st, value)


def reference_ewa(levels, u, v, g, channel, u_wrap, v_wrap, anisotropy, transpose=False):
    """The same filter in binary64 from an eigen-decomposition of the footprint covariance."""

    def footprint(level):
        h, w = level.shape[:2]
        jac = np.array([[g[0] * w, g[2] * w], [g[1] * h, g[3] * h]], dtype=np.float64)
        if transpose:
            jac = jac.T
        cov = jac @ jac.T
        vals, vecs = np.linalg.eigh(cov)
        big = vals[1]
        small = max(np.linalg.det(jac) ** 2 / big if big > 0 else 0.0, big / anisotropy ** 2)
        return vecs[:, 1], vecs[:, 0], math.sqrt(1 + big), math.sqrt(1 + small), math.sqrt(small)

    def wrap(i, n, mode):
        if mode == 0:
            return i % n, True
        if mode == 1:
            return min(max(i, 0), n - 1), True
        if mode == 2:
            p = i % (2 * n)
            return (p if p < n else 2 * n - 1 - p), True
        return (i, True) if 0 <= i < n else (0, False)

    def sample(level, fp):
        a, b, ra, rb, _ = fp
        h, w = level.shape[:2]
        cx, cy = u * w - 0.5, v * h - 0.5
        reach = int(math.ceil(max(ra, rb))) + 1
        num = den = 0.0
        for y in range(int(math.floor(cy)) - reach, int(math.ceil(cy)) + reach + 1):
            for x in range(int(math.floor(cx)) - reach, int(math.ceil(cx)) + reach + 1):
                o = np.array([x - cx, y - cy])
                r2 = (o @ a / ra) ** 2 + (o @ b / rb) ** 2
                if r2 >= 1.0:
                    continue
                weight = math.exp(-2.0 * r2) - math.exp(-2.0)
                wx, inx = wrap(x, w, u_wrap)
                wy, iny = wrap(y, h, v_wrap)
                num += weight * (level[wy, wx, channel] if inx and iny else 0.0)
                den += weight
        return num / den

    fps = [footprint(l) for l in levels]
    lo = 0
    if fps[0][4] > 1.0:
        for k in range(1, len(levels)):
            if fps[k][4] <= 1.0:
                t = math.log(fps[lo][4]) / (math.log(fps[lo][4]) - math.log(fps[k][4]))
                return (1 - t) * sample(levels[lo], fps[lo]) + t * sample(levels[k], fps[k])
            lo = k
    return sample(levels[lo], fps[lo])


CASES = [
    # (u, v, differentials, u_wrap, v_wrap, anisotropy)
    (0.41, 0.57, (0.01, 0.0, 0.0, 0.01), 0, 0, 8),
    (0.3, 0.6, (0.05, 0.02, 0.0, 0.012), 1, 1, 16),   # anisotropic and rotated
    (0.02, 0.97, (0.03, -0.01, 0.0, 0.02), 0, 2, 8),  # crosses the u and v edges: repeat / mirror
    (0.98, 0.03, (0.03, 0.01, 0.0, 0.03), 3, 1, 8),   # border reads zero past u = 1
    (0.5, 0.5, (0.09, 0.0, 0.0, 0.09), 0, 0, 8),      # minor axis of 5.8 texels: blends levels 2 and 3
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_filter_equals_a_binary64_gaussian(bf, oracle, case):
    u, v, g, uw, vw, aniso = CASES[case]
    rng = np.random.default_rng(case)
    levels = chain(bf, rng.uniform(0.0, 1.0, (64, 64, 2)))
    st, value = oracle.map_filter(levels, u, v, g, channel=1, u_wrap=uw, v_wrap=vw, anisotropy=aniso)
    want = reference_ewa([l.astype(np.float64) for l in levels], u, v, g, 1, uw, vw, aniso)
    assert st == 0 and abs(value - want) <= 2e-5 * max(1.0, abs(want)), (value, want)
    if case == 1:  # a transposed Jacobian is a different footprint
        wrong = reference_ewa([l.astype(np.float64) for l in levels], u, v, g, 1, uw, vw, aniso, transpose=True)
        assert abs(value - wrong) > 1e-4


def test_footprint_wider_than_the_image_reads_the_coarsest_level(bf, oracle):
    """Past the 1x1 level the filter falls back to a one-texel Gaussian there (the reference's
    EwaFootprint defaults), which under repeat wrap reads only the image's mean."""
    levels = chain(bf, np.random.default_rng(5).uniform(0, 1, (16, 8, 1)))
    for u, v in ((0.5, 0.5), (0.13, 0.71)):
        st, value = oracle.map_filter(levels, u, v, (5.0, 0.0, 0.0, 5.0))
        assert st == 0 and value == levels[-1][0, 0, 0], (value, levels[-1])


def test_texel_budget(bf, oracle):
    # centre 31.5 texels, radius hypot(1, 0.64) = 1.19: texels 31..32 in each axis, four reads
    levels = chain(bf, np.random.default_rng(3).uniform(0, 1, (64, 64, 1)))
    assert oracle.map_filter(levels, 0.5, 0.5, (0.01, 0.0, 0.0, 0.01), visits=3)[0] == 3
    assert oracle.map_filter(levels, 0.5, 0.5, (0.01, 0.0, 0.0, 0.01), visits=4)[0] == 0


def test_cone_differentials_head_on(oracle):
    side = 2.0
    g = oracle.cone_differentials((0.01, 0.0), (0.0, 0.0, -1.0), (0.0, 0.0, 1.0), (side, 0, 0), (0, side, 0))
    assert g is not None and np.allclose(g, [0.005, 0.0, 0.0, 0.005], rtol=1e-6, atol=0)


def maps_scene(bf, abi, normal_pixel=None, y_convention=0, ramp=None):
    s = bf.scenes.Scene()
    g = s.add_geometry(1, *bf.scenes.quad([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], [0, 0, 1]))
    m = s.add_material(2, [bf.scenes.lambert(0.5)])
    s.add_instance(3, 3, g, m)
    normal = bump = None
    if normal_pixel is not None:
        t = s.add_image_texture(chain(bf, np.tile(np.asarray(normal_pixel, np.float64), (16, 16, 1))))
        normal = {"texture": t, "y_convention": y_convention}
    if ramp is not None:
        slope, scale = ramp
        u = (np.arange(64) + 0.5) / 64
        t = s.add_image_texture(chain(bf, np.tile((slope * u)[None, :, None], (64, 1, 1))))
        bump = {"texture": t, "scale": scale, "u_wrap": abi.TEXTURE_WRAP_CLAMP, "v_wrap": abi.TEXTURE_WRAP_CLAMP}
    s.set_surface_maps(m, normal=normal, bump=bump)
    return s.finalize()


SURFACE = dict(cone=(0.004, 0.0), direction=(0.0, 0.0, -1.0), uv=(0.5, 0.5), ng=(0, 0, 1), dpdu=(1.0, 0, 0),
               dpdv=(0, 1.0, 0), ns=(0, 0, 1))


def test_flat_normal_map_keeps_the_normal(bf, abi, oracle):
    o = oracle.scene(maps_scene(bf, abi, normal_pixel=(0.5, 0.5, 1.0)))
    st, n = o.apply_maps(0, **SURFACE)
    assert st == 0 and np.array_equal(n, np.float32([0, 0, 1]))


@pytest.mark.parametrize("axis,y_convention", [(0, 0), (1, 0), (1, 1)])
def test_tilted_normal_map(bf, abi, oracle, axis, y_convention):
    theta = 0.3
    enc = [0.5, 0.5, 0.5 * (math.cos(theta) + 1)]
    enc[axis] = 0.5 * (math.sin(theta) + 1)
    o = oracle.scene(maps_scene(bf, abi, normal_pixel=enc, y_convention=y_convention))
    st, n = o.apply_maps(0, **SURFACE)
    want = np.zeros(3)
    want[axis] = math.sin(theta) * (-1 if y_convention else 1)
    want[2] = math.cos(theta)
    assert st == 0 and np.allclose(n, want, atol=2e-6), n


def test_linear_bump_ramp(bf, abi, oracle):
    slope, scale = 0.5, 0.8
    o = oracle.scene(maps_scene(bf, abi, ramp=(slope, scale)))
    st, n = o.apply_maps(0, **SURFACE)
    want = np.array([-scale * slope, 0.0, 1.0])
    assert st == 0 and np.allclose(n, want / np.linalg.norm(want), atol=2e-5), n
