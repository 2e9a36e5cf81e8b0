"""Pins of the oracle's closure library against closed forms and invariants.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a swapped
index of refraction, a transposed width, a missing cosine) fails it:
  - Fresnel: dielectric and complex-IOR conductor reflectance against the textbook amplitude
    formulas evaluated in complex128 (and k = 0 reducing the conductor to the dielectric);
  - GGX: D against its closed form and projected-area normalisation, Smith G1 / Lambda / G2
    (height-correlated) against their closed forms, the visible-normal density against
    G1(wo) (wo.m) D(m) / wo.z and its normalisation, and sampled normals reproducing it;
  - rough conductor / dielectric: the Monte-Carlo estimator E[f cos / pdf] of the sampler agrees
    with deterministic quadrature of f cos; reciprocity of reflection, the generalised
    (eta^2) reciprocity of radiance-mode transmission, and pdf normalisation;
  - energy-preserving Oren–Nayar: roughness 0 is exactly Lambert, the white-furnace albedo is 1
    for rho = 1 at every view angle, and the lobe is reciprocal.
"""
import ctypes as C
import math

import numpy as np
import pytest

OK, INVALID, NOT_REPR, OUTSIDE = 0, 1, 2, 3


@pytest.fixture(scope="module")
def lib(oracle):
    return oracle.lib


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def record(abi, kind, lobes, weight, params):
    r = abi.ClosureRecord()
    r.kind, r.lobes = kind, lobes
    for i, w in enumerate(np.broadcast_to(np.asarray(weight, dtype=np.float32), (4,))):
        r.weight[i] = float(w)
    for i, p in enumerate(params):
        r.parameters[i] = float(p)
    return r


def unit(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)], dtype=np.float32)


def call(fn, *args):
    out = C.c_float()
    st = fn(*args, C.byref(out))
    return st, out.value


# ---- Fresnel ------------------------------------------------------------------------------------
def fresnel_exact(cos_i, n1, n2):
    """Unpolarised Fresnel reflectance for a complex relative index n2/n1 (complex128)."""
    n = complex(n2) / n1
    sin2 = 1.0 - cos_i * cos_i
    cos_t = np.sqrt(n * n - sin2 + 0j)
    rs = (cos_i - cos_t) / (cos_i + cos_t)
    rp = (n * n * cos_i - cos_t) / (n * n * cos_i + cos_t)
    return 0.5 * (abs(rs) ** 2 + abs(rp) ** 2)


@pytest.mark.parametrize("eta_i,eta_t", [(1.0, 1.5), (1.5, 1.0), (1.0, 1.33), (2.4, 1.0)])
def test_dielectric_fresnel_textbook(lib, eta_i, eta_t):
    for cos_i in np.linspace(0.0, 1.0, 41):
        st, r = call(lib.bfo_dielectric_fresnel, float(cos_i), eta_i, eta_t)
        assert st == OK
        sin_t = eta_i / eta_t * math.sqrt(max(0.0, 1 - cos_i * cos_i))
        expected = 1.0 if sin_t >= 1.0 else fresnel_exact(cos_i, eta_i, eta_t)
        assert abs(r - expected) <= 2e-6 + 1e-5 * expected, (cos_i, r, expected)


@pytest.mark.parametrize("eta,k", [(0.27, 2.78), (1.34, 1.94), (0.18, 3.35), (2.9, 3.1), (1.5, 0.0), (0.7, 0.0)])
def test_conductor_fresnel_complex_ior(lib, eta, k):
    for cos_i in np.linspace(0.0, 1.0, 41):
        st, r = call(lib.bfo_conductor_fresnel, float(cos_i), eta, k)
        assert st == OK
        if k == 0.0 and eta < 1.0:
            continue  # total internal reflection region is the dielectric's business
        expected = fresnel_exact(cos_i, 1.0, complex(eta, k))
        assert abs(r - expected) <= 5e-6 + 2e-5 * expected, (cos_i, r, expected)
    # k = 0 above the critical angle: the conductor reduces to the dielectric
    if k == 0.0 and eta > 1.0:
        for cos_i in np.linspace(0.05, 1.0, 12):
            _, rc = call(lib.bfo_conductor_fresnel, float(cos_i), eta, 0.0)
            _, rd = call(lib.bfo_dielectric_fresnel, float(cos_i), 1.0, eta)
            assert abs(rc - rd) <= 2e-6


# ---- GGX ----------------------------------------------------------------------------------------
def ggx_d_exact(ax, ay, m):
    return 1.0 / (math.pi * ax * ay * ((m[0] / ax) ** 2 + (m[1] / ay) ** 2 + m[2] ** 2) ** 2)


def lambda_exact(ax, ay, d):
    if d[0] == 0 and d[1] == 0:
        return 0.0
    a2t2 = ((ax * d[0]) ** 2 + (ay * d[1]) ** 2) / (d[2] * d[2])
    return (math.sqrt(1 + a2t2) - 1) / 2


@pytest.mark.parametrize("ax,ay", [(0.05, 0.05), (0.3, 0.3), (0.6, 0.6), (1.7, 1.7), (0.2, 0.5), (0.7, 0.15)])
def test_ggx_closed_forms(lib, ax, ay):
    rng = np.random.default_rng(7)
    for _ in range(200):
        theta, phi = rng.uniform(0, math.pi / 2 * 0.999), rng.uniform(0, 2 * math.pi)
        m = unit(theta, phi)
        st, d = call(lib.bfo_ggx_d, ax, ay, f3(m))
        assert st == OK
        ref = ggx_d_exact(ax, ay, m.astype(np.float64))
        assert abs(d - ref) <= 2e-5 * ref + 1e-30
        st, lam = call(lib.bfo_smith_lambda, ax, ay, f3(m))
        ref_lam = lambda_exact(ax, ay, m.astype(np.float64))
        assert st == OK and abs(lam - ref_lam) <= 2e-5 * ref_lam + 1e-7
        st, g1 = call(lib.bfo_smith_g1, ax, ay, f3(m))
        assert st == OK and abs(g1 - 1 / (1 + ref_lam)) <= 2e-6
        wi = unit(rng.uniform(0, math.pi / 2 * 0.999), rng.uniform(0, 2 * math.pi))
        st, g2 = call(lib.bfo_smith_g2, ax, ay, f3(m), f3(wi))
        ref_g2 = 1 / (1 + ref_lam + lambda_exact(ax, ay, wi.astype(np.float64)))
        assert st == OK and abs(g2 - ref_g2) <= 4e-6


def hemisphere_grid(n_theta=256, n_phi=256):
    """Midpoint quadrature of the hemisphere in (cos theta, phi)."""
    mu = (np.arange(n_theta) + 0.5) / n_theta
    phi = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
    mu, phi = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1 - mu * mu)
    dirs = np.stack([s * np.cos(phi), s * np.sin(phi), mu], axis=-1).reshape(-1, 3).astype(np.float32)
    weight = 2 * math.pi / (n_theta * n_phi)
    return dirs, weight


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.6, 0.6), (0.25, 0.5)])
def test_ggx_normalisation_and_vndf(lib, ax, ay):
    dirs, w = hemisphere_grid()
    # projected-area normalisation: integral of D(m) m.z over the hemisphere is 1
    total = 0.0
    for m in dirs[::3]:
        _, d = call(lib.bfo_ggx_d, ax, ay, f3(m))
        total += d * m[2]
    assert abs(total * w * 3 - 1.0) < 2e-2
    wo = unit(1.1, 0.4)
    # visible-normal density: closed form G1(wo) (wo.m) D(m) / wo.z, normalised over m
    _, g1 = call(lib.bfo_smith_g1, ax, ay, f3(wo))
    total = 0.0
    for m in dirs[::3]:
        st, p = call(lib.bfo_vndf_pdf, ax, ay, f3(wo), f3(m))
        assert st == OK
        _, d = call(lib.bfo_ggx_d, ax, ay, f3(m))
        scale = g1 * d / float(wo[2])
        ref = scale * max(0.0, float(np.dot(wo.astype(np.float64), m.astype(np.float64))))
        assert abs(p - ref) <= 5e-5 * ref + 1e-6 * scale  # (wo.m) cancels near the silhouette
        total += p
    assert abs(total * w * 3 - 1.0) < 2e-2
    # sampled normals: density returned equals the density function, and the mean of m.z matches quadrature
    rng = np.random.default_rng(3)
    mz = []
    for _ in range(4000):
        m = (C.c_float * 3)()
        pdf = C.c_float()
        st = lib.bfo_sample_vndf(ax, ay, f3(wo), float(rng.random()), float(rng.random()), m, C.byref(pdf))
        assert st == OK
        _, p = call(lib.bfo_vndf_pdf, ax, ay, f3(wo), f3(list(m)))
        assert p == pdf.value
        mz.append(m[2])
    expected = 0.0
    for m in dirs[::3]:
        _, p = call(lib.bfo_vndf_pdf, ax, ay, f3(wo), f3(m))
        expected += p * m[2]
    expected *= w * 3
    assert abs(np.mean(mz) - expected) < 4 * np.std(mz) / math.sqrt(len(mz)) + 2e-3


# ---- lobes --------------------------------------------------------------------------------------
def eval_lobe(lib, rec, wo, wi, allowed=0x18):
    out = (C.c_float * 4)()
    st = lib.bfo_closure_eval(C.byref(rec), allowed, f3(wo), f3(wi), out)
    return st, np.array(list(out))


def pdf_lobe(lib, rec, wo, wi, allowed=0x18):
    p = C.c_float()
    st = lib.bfo_closure_pdf(C.byref(rec), allowed, f3(wo), f3(wi), C.byref(p))
    return st, p.value


def sample_lobe(lib, rec, wo, event, u, v, allowed=0x18):
    out = (C.c_float * 9)()
    lobes = C.c_uint32()
    st = lib.bfo_closure_sample(C.byref(rec), allowed, f3(wo), event, u, v, out, C.byref(lobes))
    return st, np.array(list(out)), lobes.value


def sphere_grid(n_theta=200, n_phi=200):
    dirs, w = hemisphere_grid(n_theta, n_phi)
    lower = dirs.copy()
    lower[:, 2] = -lower[:, 2]
    return np.concatenate([dirs, lower]), w


def conductor(abi, alpha_x, alpha_y, eta=(0.27, 0.2, 0.18, 0.16), k=(2.78, 3.1, 3.35, 3.86)):
    return record(abi, abi.CLOSURE_ROUGH_CONDUCTOR, abi.LOBE_GLOSSY | abi.LOBE_REFLECTION, 1.0,
                  list(eta) + list(k) + [alpha_x, alpha_y])


def dielectric(abi, alpha, exterior=1.0, interior=1.5):
    return record(abi, abi.CLOSURE_ROUGH_DIELECTRIC, abi.LOBE_GLOSSY | abi.LOBE_REFLECTION | abi.LOBE_TRANSMISSION,
                  1.0, [exterior, interior, alpha, alpha])


@pytest.mark.parametrize("alphas", [(0.2, 0.2), (0.5, 0.5), (0.15, 0.45)])
def test_rough_conductor_estimator_matches_quadrature(lib, abi, alphas):
    rec = conductor(abi, *alphas)
    wo = unit(0.9, 0.3)
    dirs, w = hemisphere_grid(160, 160)
    quad = np.zeros(4)
    pdf_total = 0.0
    for wi in dirs:
        st, f = eval_lobe(lib, rec, wo, wi)
        assert st == OK
        quad += f * wi[2]
        _, p = pdf_lobe(lib, rec, wo, wi)
        pdf_total += p
    quad *= w
    assert 0.7 < pdf_total * w <= 1.0 + 1e-2  # reflected normals below the horizon are lost
    rng = np.random.default_rng(11)
    est, n = [], 20000
    for _ in range(n):
        st, out, lobes = sample_lobe(lib, rec, wo, 0.5, float(rng.random()), float(rng.random()))
        if st == OUTSIDE:
            est.append(np.zeros(4))
            continue
        assert st == OK and lobes == abi.LOBE_GLOSSY | abi.LOBE_REFLECTION
        est.append(out[:4] * out[6] / out[7])
    est = np.array(est)
    mean, err = est.mean(axis=0), est.std(axis=0) / math.sqrt(n)
    assert np.all(np.abs(mean - quad) < 5 * err + 5e-3 * quad), (mean, quad)
    # reciprocity of reflection
    for wi in dirs[::997]:
        _, a = eval_lobe(lib, rec, wo, wi)
        _, b = eval_lobe(lib, rec, wi, wo)
        assert np.allclose(a, b, rtol=2e-5, atol=1e-7)


@pytest.mark.parametrize("alpha,outside", [(0.2, True), (0.45, True), (0.3, False)])
def test_rough_dielectric_estimator_and_reciprocity(lib, abi, alpha, outside):
    rec = dielectric(abi, alpha)
    wo = unit(0.7, 1.9)
    if not outside:
        wo = -wo  # arriving from inside the medium
    dirs, w = sphere_grid(160, 160)
    quad = np.zeros(4)
    pdf_total = 0.0
    for wi in dirs:
        st, f = eval_lobe(lib, rec, wo, wi)
        assert st == OK
        quad += f * abs(wi[2])
        _, p = pdf_lobe(lib, rec, wo, wi)
        pdf_total += p
    quad *= w
    assert 0.85 < pdf_total * w <= 1.0 + 1e-2
    rng = np.random.default_rng(5)
    est, n = [], 20000
    for _ in range(n):
        st, out, lobes = sample_lobe(lib, rec, wo, float(rng.random()), float(rng.random()), float(rng.random()))
        if st in (OUTSIDE, NOT_REPR):
            est.append(np.zeros(4))
            continue
        assert st == OK
        est.append(out[:4] * abs(out[6]) / out[7])
    est = np.array(est)
    mean, err = est.mean(axis=0), est.std(axis=0) / math.sqrt(n)
    assert np.all(np.abs(mean - quad) < 5 * err + 1e-2 * quad), (mean, quad)
    # reflection is reciprocal; radiance-mode transmission obeys f(o,i)/eta_o^2 = f(i,o)/eta_i^2
    eta_o = 1.0 if wo[2] > 0 else 1.5
    for wi in dirs[::613]:
        _, a = eval_lobe(lib, rec, wo, wi)
        _, b = eval_lobe(lib, rec, wi, wo)
        if (wi[2] > 0) == (wo[2] > 0):
            assert np.allclose(a, b, rtol=5e-5, atol=1e-7)
        else:
            eta_i = 1.0 if wi[2] > 0 else 1.5
            assert np.allclose(a * eta_i ** 2, b * eta_o ** 2, rtol=5e-4, atol=1e-6), (a, b)


def test_rough_dielectric_direction_masks(lib, abi):
    rec = dielectric(abi, 0.3)
    wo = unit(0.6, 0.2)
    rng = np.random.default_rng(2)
    for allowed, want in ((abi.LOBE_REFLECTION, abi.LOBE_REFLECTION), (abi.LOBE_TRANSMISSION, abi.LOBE_TRANSMISSION)):
        for _ in range(50):
            st, out, lobes = sample_lobe(lib, rec, wo, float(rng.random()), float(rng.random()), float(rng.random()),
                                         allowed)
            if st != OK:
                continue
            assert lobes & 0x18 == want
            # the masked density is the full density conditioned on the branch (divided by its probability)
            _, full = pdf_lobe(lib, rec, wo, out[4:7])
            assert out[7] >= full
        wi_other = unit(2.4, 0.1) if want == abi.LOBE_REFLECTION else unit(0.5, 0.4)
        st, f = eval_lobe(lib, rec, wo, wi_other, allowed)
        assert st == OK and not f.any()


def test_oren_nayar_energy_preserving(lib, abi):
    rho1 = record(abi, abi.CLOSURE_ROUGH_DIFFUSE, abi.LOBE_DIFFUSE | abi.LOBE_REFLECTION, 1.0, [0.0])
    lam = record(abi, abi.CLOSURE_LAMBERTIAN, abi.LOBE_DIFFUSE | abi.LOBE_REFLECTION, 1.0, [])
    wo, wi = unit(0.4, 0.1), unit(1.0, 2.0)
    assert np.array_equal(eval_lobe(lib, rho1, wo, wi)[1], eval_lobe(lib, lam, wo, wi)[1])
    dirs, w = hemisphere_grid(200, 200)
    for sigma in (0.3, 0.7, 1.0):
        rec = record(abi, abi.CLOSURE_ROUGH_DIFFUSE, abi.LOBE_DIFFUSE | abi.LOBE_REFLECTION, 1.0, [sigma])
        for theta in (0.1, 0.8, 1.4):
            wo = unit(theta, 0.3)
            albedo = sum(eval_lobe(lib, rec, wo, d)[1][0] * d[2] for d in dirs[::2]) * w * 2
            assert abs(albedo - 1.0) < 5e-3, (sigma, theta, albedo)
        for d in dirs[::4001]:
            a, b = eval_lobe(lib, rec, wo, d)[1], eval_lobe(lib, rec, d, wo)[1]
            assert np.allclose(a, b, rtol=1e-5)


def test_invalid_closures_rejected(lib, abi):
    wo, wi = unit(0.3, 0.0), unit(0.5, 1.0)
    bad = [
        dielectric(abi, 0.3, 1.5, 1.5),  # equal indices
        record(abi, abi.CLOSURE_ROUGH_DIELECTRIC, abi.LOBE_GLOSSY | abi.LOBE_REFLECTION, 1.0, [1.0, 1.5, 0.3, 0.3]),
        conductor(abi, 0.0, 0.3),
        record(abi, abi.CLOSURE_LAMBERTIAN, abi.LOBE_DIFFUSE | abi.LOBE_REFLECTION, 1.2, []),
        record(abi, abi.CLOSURE_ROUGH_DIFFUSE, abi.LOBE_DIFFUSE | abi.LOBE_REFLECTION, 0.5, [1.5]),
    ]
    for rec in bad:
        assert eval_lobe(lib, rec, wo, wi)[0] == INVALID
        assert pdf_lobe(lib, rec, wo, wi)[0] == INVALID
        assert sample_lobe(lib, rec, wo, 0.2, 0.3, 0.4)[0] == INVALID
